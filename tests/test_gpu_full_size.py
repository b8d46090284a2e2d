"""Full-size parity: the decoder at the benchmark's batch sizes and beyond 2^31 LLR elements,
checked against the CPU oracle on sampled codewords (bit-exact hard decisions, APP and
iteration counts), plus the empty batch.

Every case prints its parameters before running so a failing GPU step names itself.
"""
import numpy as np
import pytest

from nr_gpu_ldpc_amd import mixed as mx
from tests.oracle_binding import OR_BG1, OR_BG2

pytestmark = pytest.mark.gpu

KIND = {1: OR_BG1, 2: OR_BG2}


def _channel(nr, torch, code, n_cw, ebn0, seed):
    llr = torch.empty((n_cw, code.n_vn), dtype=torch.float32, device="cuda")
    code.channel(seed, 0, n_cw, llr.data_ptr(), sigma=nr.noise_stddev(ebn0, code.K / code.N),
                 stream=torch.cuda.current_stream().cuda_stream)
    return llr


def test_bg1_z384_batch_4096_samples_match_oracle(nr, oracle, torch_gpu):
    """The headline launch (BG1 Z=384, 4096 codewords, 10 iterations): sampled codewords,
    including both ends of the batch, equal the oracle bit for bit."""
    torch = torch_gpu
    print("case BG1 Z=384 batch 4096 T=10", flush=True)
    code = nr.Code(1, 384)
    n = 4096
    llr = _channel(nr, torch, code, n, 1.5, seed=404)
    hd = torch.zeros((n, code.hd_bytes), dtype=torch.uint8, device="cuda")
    app = torch.zeros((n, code.n_vn), dtype=torch.float32, device="cuda")
    its = torch.zeros(n, dtype=torch.int32, device="cuda")
    code.decode(llr.data_ptr(), n, 10, hd.data_ptr(), app.data_ptr(), its.data_ptr(),
                stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    pick = np.array([0, 1, 777, 2047, 2048, 3001, 4094, 4095])
    sel = torch.from_numpy(pick).cuda()
    ohd, oapp, _ = oracle.decode(oracle.code(OR_BG1, 384), llr[sel].cpu().numpy(), 10)
    ghd = np.unpackbits(hd[sel].cpu().numpy(), axis=1, bitorder="little")[:, :code.n_vn]
    assert np.array_equal(ghd, ohd)
    assert np.array_equal(app[sel].cpu().numpy().view(np.uint32), oapp.view(np.uint32))
    assert (its.cpu().numpy() == 10).all()


@pytest.mark.parametrize("ebn0,early", [(1.5, False), (2.25, True)])
def test_bg1_z384_char_batch_4096_samples_match_oracle(nr, oracle, torch_gpu, ebn0, early):
    """The CHAR_PRC launch of the same workload (the reference's default build: LLRs as
    quantize(0.5, x, 8), 2048 workgroups of one codeword pair in two row groups): sampled
    codewords from both halves of pairs and both ends of the batch equal the oracle's CHAR_PRC
    decode bit for bit, with early exit too (iteration counts included)."""
    torch = torch_gpu
    print(f"case CHAR BG1 Z=384 batch 4096 T=10 ebn0={ebn0} early={early}", flush=True)
    code = nr.Code(1, 384)
    n = 4096
    llr = _channel(nr, torch, code, n, ebn0, seed=405)
    q = torch.clamp(torch.floor(llr / 0.5), -128, 127).to(torch.int8)  # quantize(0.5, x, 8)
    del llr
    hd = torch.zeros((n, code.hd_bytes), dtype=torch.uint8, device="cuda")
    app = torch.zeros((n, code.n_vn), dtype=torch.float32, device="cuda")
    its = torch.zeros(n, dtype=torch.int32, device="cuda")
    code.decode_char(q.data_ptr(), n, 10, hd.data_ptr(), app.data_ptr(), its.data_ptr(),
                     flags=nr.EARLY_EXIT if early else 0, stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    pick = np.array([0, 1, 776, 777, 2046, 2047, 3001, 4094, 4095])
    sel = torch.from_numpy(pick).cuda()
    ohd, oapp, oit, _ = oracle.decode_char(oracle.code(OR_BG1, 384), q[sel].cpu().numpy(), 10, early_exit=early)
    ghd = np.unpackbits(hd[sel].cpu().numpy(), axis=1, bitorder="little")[:, :code.n_vn]
    assert np.array_equal(ghd, ohd)
    assert np.array_equal(app[sel].cpu().numpy(), oapp)
    git = its.cpu().numpy()
    assert np.array_equal(git[pick], oit)
    if early:
        assert git.min() < 10 and git.max() == 10  # codewords stop at different iterations
    else:
        assert (git == 10).all()


def test_more_than_2_pow_31_llr_elements(nr, oracle, torch_gpu):
    """82,300 BG1 Z=384 codewords hold 2.149e9 LLRs (8.6 GB): offsets past 2^31 elements are
    addressed correctly; sampled codewords from the far end of the batch equal the oracle."""
    torch = torch_gpu
    n = 82300
    print(f"case BG1 Z=384 batch {n} (LLR elements beyond 2^31) T=5", flush=True)
    code = nr.Code(1, 384)
    assert n * code.n_vn > 2 ** 31
    llr = _channel(nr, torch, code, n, 2.0, seed=31)
    hd = torch.zeros((n, code.hd_bytes), dtype=torch.uint8, device="cuda")
    code.decode(llr.data_ptr(), n, 5, hd.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    # codeword 82241 straddles element 2^31, 82242 is the first to start past it
    pick = np.array([0, 41000, 82241, 82242, 82299])
    assert 82242 * code.n_vn > 2 ** 31 > 82241 * code.n_vn
    sel = torch.from_numpy(pick).cuda()
    ohd, _, _ = oracle.decode(oracle.code(OR_BG1, 384), llr[sel].cpu().numpy(), 5)
    ghd = np.unpackbits(hd[sel].cpu().numpy(), axis=1, bitorder="little")[:, :code.n_vn]
    assert np.array_equal(ghd, ohd)
    del llr, hd
    torch.cuda.empty_cache()


def test_empty_batch_touches_nothing(nr, torch_gpu):
    torch = torch_gpu
    print("case empty batch", flush=True)
    code = nr.Code(2, 208)
    llr = torch.zeros((1, code.n_vn), dtype=torch.float32, device="cuda")
    hd = torch.full((1, code.hd_bytes), 0xAB, dtype=torch.uint8, device="cuda")
    its = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    code.decode(llr.data_ptr(), 0, 10, hd.data_ptr(), 0, its.data_ptr(), flags=nr.EARLY_EXIT)
    torch.cuda.synchronize()
    assert (hd.cpu().numpy() == 0xAB).all() and int(its.item()) == -7


def test_mixed_batch_16384_samples_match_oracle(nr, oracle, torch_gpu):
    """The mixed-batch benchmark's launch (16,384 codewords of random base graph, Z and rate,
    at most 20 iterations with early exit): sampled codewords equal the oracle decoding each
    with its own code and parity-row count, iteration counts included."""
    torch = torch_gpu
    print("case mixed batch 16384 T<=20 early exit", flush=True)
    n, T = 16384, 20
    b = mx.draw_batch(n, seed=2025)
    llr, _ = mx.make_channel(torch, b, seed=77)
    m = nr.Mixed(n)
    m.plan(b["desc"])
    hd = torch.zeros(b["hd_total"], dtype=torch.uint8, device="cuda")
    app = torch.zeros(b["llr_floats"], dtype=torch.float32, device="cuda")
    its = torch.zeros(n, dtype=torch.int32, device="cuda")
    m.decode(llr.data_ptr(), T, hd.data_ptr(), app.data_ptr(), its.data_ptr(), flags=nr.EARLY_EXIT)
    torch.cuda.synchronize()
    rng = np.random.default_rng(5)
    pick = np.concatenate([[0, n - 1], rng.choice(np.arange(1, n - 1), 30, replace=False)])
    d = b["desc"]
    h_llr, h_hd, h_app, h_its = llr.cpu().numpy(), hd.cpu().numpy(), app.cpu().numpy(), its.cpu().numpy()
    for i in pick:
        kind, Z, mr = int(b["kind"][i]), int(b["Z"][i]), int(b["m_rows"][i])
        nv, a0, h0 = int(b["n_vn"][i]), int(d["llr_offset"][i]), int(d["hd_offset"][i])
        ohd, oapp, oit = oracle.decode(oracle.code(KIND[kind], Z, mr), h_llr[a0:a0 + nv][None, :], T, early_exit=True)
        ghd = np.unpackbits(h_hd[h0:h0 + int(b["hd_bytes"][i])], bitorder="little")[:nv]
        assert np.array_equal(ghd, ohd[0]), f"codeword {i} (kind {kind}, Z {Z}, rows {mr})"
        assert np.array_equal(h_app[a0:a0 + nv].view(np.uint32), oapp[0].view(np.uint32)), f"codeword {i}"
        assert h_its[i] == oit[0], f"codeword {i}"
