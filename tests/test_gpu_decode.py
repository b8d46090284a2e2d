"""GPU parity of the fused decoder against the CPU oracle (bit-exact hard decisions and APP).

Every case prints its parameters before running so a failing GPU step names itself.
"""
import numpy as np
import pytest

from tests.oracle_binding import OR_BG1, OR_BG2, OR_WIFI

pytestmark = pytest.mark.gpu

KIND = {1: OR_BG1, 2: OR_BG2, 3: OR_WIFI}


def _noisy_llr(nr, code, n_cw, ebn0, rate, seed):
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, (n_cw, code.K), dtype=np.uint8)
    cw = code.encode_host(info)
    sigma = nr.noise_stddev(ebn0, rate)
    x = 1.0 - 2.0 * cw[:, code.punct * code.Z:].astype(np.float32)
    y = (x + sigma * rng.standard_normal(x.shape).astype(np.float32)).astype(np.float32)
    llr = np.zeros((n_cw, code.n_vn), dtype=np.float32)
    llr[:, code.punct * code.Z:] = (y * np.float32(2.0)) / np.float32(sigma * sigma)
    return llr, cw


def _gpu_decode(torch, code, llr, n_iter, flags=0, want_app=True):
    d_llr = torch.from_numpy(llr).cuda()
    n_cw = llr.shape[0]
    d_hd = torch.zeros((n_cw, code.hd_bytes), dtype=torch.uint8, device="cuda")
    d_app = torch.zeros((n_cw, code.n_vn), dtype=torch.float32, device="cuda") if want_app else None
    d_it = torch.zeros(n_cw, dtype=torch.int32, device="cuda")
    code.decode(d_llr.data_ptr(), n_cw, n_iter, d_hd.data_ptr(), d_app.data_ptr() if want_app else 0,
                d_it.data_ptr(), flags=flags, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    hd = np.unpackbits(d_hd.cpu().numpy(), axis=1, bitorder="little")[:, :code.n_vn]
    app = d_app.cpu().numpy() if want_app else None
    return hd, app, d_it.cpu().numpy()


CASES = [
    # kind, Z, m_rows, n_cw, n_iter, ebn0, rate
    (1, 384, 0, 6, 10, 1.25, 1 / 3),
    (1, 384, 0, 3, 1, 0.5, 1 / 3),
    (2, 208, 0, 6, 10, 0.0, 1 / 5),
    (3, 96, 0, 24, 5, 2.0, 1 / 2),
    (1, 2, 0, 40, 4, 1.0, 1 / 3),
    (1, 15, 0, 40, 6, 1.0, 1 / 3),
    (1, 52, 10, 20, 6, 2.0, 22 / 30),
    (2, 7, 0, 40, 6, 1.0, 1 / 5),
    (2, 104, 20, 10, 8, 1.5, 10 / 28),
    (2, 384, 0, 3, 5, 0.5, 1 / 5),
    (1, 320, 0, 3, 5, 1.0, 1 / 3),
    (3, 13, 0, 40, 5, 2.0, 1 / 2),
]


@pytest.mark.parametrize("kind,Z,m,n_cw,n_iter,ebn0,rate", CASES)
def test_decode_matches_oracle(nr, oracle, torch_gpu, kind, Z, m, n_cw, n_iter, ebn0, rate):
    print(f"case kind={kind} Z={Z} m={m} n_cw={n_cw} T={n_iter}", flush=True)
    code = nr.Code(kind, Z, m)
    llr, _ = _noisy_llr(nr, code, n_cw, ebn0, rate, seed=Z * 7 + kind)
    hd, app, iters = _gpu_decode(torch_gpu, code, llr, n_iter)
    ohd, oapp, oit = oracle.decode(oracle.code(KIND[kind], Z, m), llr, n_iter)
    assert np.array_equal(hd, ohd), f"hard decisions differ in {int((hd != ohd).sum())} bits"
    assert np.array_equal(app, oapp), f"APP differs at {int((app != oapp).sum())} positions"
    assert (iters == n_iter).all()


@pytest.mark.parametrize("cpb", [1, 2, 3])
def test_codewords_per_block(nr, oracle, torch_gpu, cpb):
    print(f"case BG2 Z=52 cpb={cpb}", flush=True)
    code = nr.Code(2, 52)
    code.set_launch(cpb)
    llr, _ = _noisy_llr(nr, code, 25, 1.0, 0.2, seed=cpb)
    hd, app, _ = _gpu_decode(torch_gpu, code, llr, 7)
    ohd, oapp, _ = oracle.decode(oracle.code(OR_BG2, 52), llr, 7)
    assert np.array_equal(hd, ohd) and np.array_equal(app, oapp)


# (kind, Z, codewords per workgroup or 0 for the default, codewords, Eb/N0, rate): the Z=52 and
# Z=24 cases put lanes of two codewords into one wavefront, whose syndrome check must not stop
# at the first failing row of either.
EARLY_CASES = [(1, 64, 0, 48, 2.0, 1 / 3), (2, 52, 3, 40, 1.0, 0.2), (2, 24, 0, 60, 1.5, 0.2),
               (1, 384, 1, 8, 1.6, 1 / 3)]


@pytest.mark.parametrize("kind,Z,cpb,n_cw,ebn0,rate", EARLY_CASES)
def test_early_exit_matches_oracle(nr, oracle, torch_gpu, kind, Z, cpb, n_cw, ebn0, rate):
    print(f"case early exit kind={kind} Z={Z} cpb={cpb}", flush=True)
    code = nr.Code(kind, Z)
    if cpb:
        code.set_launch(cpb)
    llr, _ = _noisy_llr(nr, code, n_cw, ebn0, rate, seed=5 + Z)
    hd, app, iters = _gpu_decode(torch_gpu, code, llr, 20, flags=nr.EARLY_EXIT)
    ohd, oapp, oit = oracle.decode(oracle.code(KIND[kind], Z), llr, 20, early_exit=True)
    assert np.array_equal(iters, oit)
    assert np.array_equal(hd, ohd) and np.array_equal(app, oapp)
    assert iters.min() < 20


@pytest.mark.parametrize("Z,cpb", [(24, 2), (13, 4), (52, 3)])
def test_early_exit_mixed_convergence_in_one_wavefront(nr, oracle, torch_gpu, Z, cpb):
    """Codewords that converge at once (noiseless), late (moderate noise) and never (pure
    noise) side by side in one workgroup and wavefront: iteration counts, decisions and APP equal
    the oracle's (the syndrome walk of a wavefront stops only when every codeword in it has a
    failing row)."""
    print(f"case early exit mixed convergence BG2 Z={Z} cpb={cpb}", flush=True)
    code = nr.Code(2, Z)
    code.set_launch(cpb)
    n = 4 * cpb
    good, _ = _noisy_llr(nr, code, n, 40.0, 0.2, seed=Z)
    mid, _ = _noisy_llr(nr, code, n, 1.5, 0.2, seed=Z + 1)
    rng = np.random.default_rng(Z + 2)
    bad = (rng.standard_normal((n, code.n_vn)) * 3).astype(np.float32)
    bad[:, :code.punct * Z] = 0
    pick = rng.integers(0, 3, n)
    pick[:3] = [0, 2, 1]
    llr = np.where((pick == 0)[:, None], good, np.where((pick == 1)[:, None], mid, bad)).astype(np.float32)
    hd, app, iters = _gpu_decode(torch_gpu, code, llr, 20, flags=nr.EARLY_EXIT)
    ohd, oapp, oit = oracle.decode(oracle.code(OR_BG2, Z), llr, 20, early_exit=True)
    assert np.array_equal(iters, oit)
    assert np.array_equal(hd, ohd) and np.array_equal(app, oapp)
    assert iters.min() <= 2 and iters.max() == 20


def test_special_values(nr, oracle, torch_gpu):
    """Signed zeros, exact ties and magnitudes above the 1000.0f start value."""
    print("case special values QC12x24 Z=8", flush=True)
    code = nr.Code(3, 8)
    rng = np.random.default_rng(3)
    n = code.n_vn
    rows = []
    rows.append(np.full(n, 2000.0, np.float32))
    rows.append(np.where(rng.random(n) < 0.5, -1.0, 1.0).astype(np.float32))      # all |Q| tie
    rows.append(np.where(rng.random(n) < 0.5, -0.0, 0.0).astype(np.float32))      # signed zeros
    r = rng.standard_normal(n).astype(np.float32) * 3000
    rows.append(r)
    rows.append(np.round(rng.standard_normal(n) * 3).astype(np.float32))         # many ties
    llr = np.stack(rows)
    hd, app, _ = _gpu_decode(torch_gpu, code, llr, 6)
    ohd, oapp, _ = oracle.decode(oracle.code(OR_WIFI, 8), llr, 6)
    assert np.array_equal(hd, ohd)
    # Values are identical; bit patterns too, except that the decoder canonicalises -0.0
    # channel LLRs to +0.0 on load, so an APP that stays exactly zero may differ in the sign
    # of that zero (both give hard decision 1, and every comparison treats them as equal).
    assert np.array_equal(app, oapp)
    nz = oapp != 0
    assert np.array_equal(app.view(np.uint32)[nz], oapp.view(np.uint32)[nz])


@pytest.mark.parametrize("kind,Z", [(1, 384), (2, 13)])
def test_special_values_slim_layout(nr, oracle, torch_gpu, kind, Z):
    """The special values above in the slim layout (single read copy with wrap-selected reads,
    degree-1 posteriors in device slots), with compiled-in offsets (BG1 Z=384) and the per-edge
    table (BG2 Z=13), through the persistent loop (more codewords than one workgroup holds)."""
    print(f"case special values slim kind={kind} Z={Z}", flush=True)
    code = nr.Code(kind, Z)
    code.set_launch_layout(nr.LAYOUT_SLIM, 0)
    cpb = code.launch_layout(1)[1]
    rng = np.random.default_rng(5 + Z)
    n = code.n_vn
    base = [np.full(n, 2000.0, np.float32),
            np.where(rng.random(n) < 0.5, -1.0, 1.0).astype(np.float32),
            np.where(rng.random(n) < 0.5, -0.0, 0.0).astype(np.float32),
            (rng.standard_normal(n) * 3000).astype(np.float32),
            np.round(rng.standard_normal(n) * 3).astype(np.float32)]
    n_cw = max(len(base), cpb + 1)
    llr = np.stack([base[i % len(base)] for i in range(n_cw)])
    llr[:, :code.punct * Z] = 0.0
    assert code.launch_layout(n_cw)[0] == nr.LAYOUT_SLIM
    hd, app, _ = _gpu_decode(torch_gpu, code, llr, 6)
    ohd, oapp, _ = oracle.decode(oracle.code(KIND[kind], Z), llr, 6)
    assert np.array_equal(hd, ohd)
    assert np.array_equal(app, oapp)
    nz = oapp != 0
    assert np.array_equal(app.view(np.uint32)[nz], oapp.view(np.uint32)[nz])


def test_noiseless_roundtrip_full_batch(nr, torch_gpu):
    """BG1 Z=384, 4096 codewords at the bench size: noiseless codewords decode to themselves."""
    print("case BG1 Z=384 4096 noiseless", flush=True)
    torch = torch_gpu
    code = nr.Code(1, 384)
    n_cw = 4096
    d_llr = torch.empty((n_cw, code.n_vn), dtype=torch.float32, device="cuda")
    d_bits = torch.empty((n_cw, code.n_vn), dtype=torch.uint8, device="cuda")
    code.channel(11, 0, n_cw, d_llr.data_ptr(), sigma=1e-3, cw_bits_ptr=d_bits.data_ptr())
    d_hd = torch.zeros((n_cw, code.hd_bytes), dtype=torch.uint8, device="cuda")
    code.decode(d_llr.data_ptr(), n_cw, 10, d_hd.data_ptr())
    d_w = torch.zeros(n_cw, dtype=torch.int32, device="cuda")
    code.syndrome(d_hd.data_ptr(), n_cw, d_w.data_ptr())
    cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
    code.count_errors(d_hd.data_ptr(), d_bits.data_ptr(), n_cw, cnt.data_ptr())
    torch.cuda.synchronize()
    assert int(d_w.sum()) == 0
    assert cnt.tolist() == [0, 0]
    hd = np.unpackbits(d_hd[:64].cpu().numpy(), axis=1, bitorder="little")[:, :code.n_vn]
    assert np.array_equal(hd, d_bits[:64].cpu().numpy())


def test_bg1_z384_ebn0_sweep_points_match_oracle(nr, oracle, torch_gpu):
    """The benchmark's Eb/N0 sweep (BG1 Z=384, 10 fixed iterations, 0..3 dB in 0.25 dB steps,
    LLRs from the device channel): at every point two codewords decode bit-exact (hard decisions
    and APP) against the oracle, plain and learned with random weights."""
    torch = torch_gpu
    code = nr.Code(1, 384)
    oc = oracle.code(OR_BG1, 384)
    T = 10
    rng = np.random.default_rng(5)
    w = np.stack([rng.uniform(0.5, 1.0, (T, code.nnz_full)), rng.uniform(-0.3, 0.5, (T, code.nnz_full))], -1)
    w = w.astype(np.float32)
    code.set_weights(w)
    d_llr = torch.empty((2, code.n_vn), dtype=torch.float32, device="cuda")
    for i in range(13):
        ebn0 = 0.25 * i
        print(f"case sweep BG1 Z=384 Eb/N0={ebn0}", flush=True)
        code.channel(7000 + 100 * i, 0, 2, d_llr.data_ptr(), sigma=nr.noise_stddev(ebn0, 1 / 3),
                     stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        llr = d_llr.cpu().numpy()
        hd, app, _ = _gpu_decode(torch, code, llr, T)
        ohd, oapp, _ = oracle.decode(oc, llr, T)
        assert np.array_equal(hd, ohd) and np.array_equal(app, oapp), ebn0
        d_hd = torch.zeros((2, code.hd_bytes), dtype=torch.uint8, device="cuda")
        d_app = torch.zeros((2, code.n_vn), dtype=torch.float32, device="cuda")
        code.decode_learned(d_llr.data_ptr(), 2, T, d_hd.data_ptr(), d_app.data_ptr(),
                            stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        lhd = np.unpackbits(d_hd.cpu().numpy(), axis=1, bitorder="little")[:, :code.n_vn]
        lo_hd, lo_app = oracle.decode_noms(oc, llr, T, w[..., 0].reshape(-1), w[..., 1].reshape(-1))
        assert np.array_equal(lhd, lo_hd) and np.array_equal(d_app.cpu().numpy(), lo_app), ebn0
