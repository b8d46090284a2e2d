"""Every TS 38.212 lifting size of both base graphs (102 (BG, Z) pairs) and the 12x24 code, decoded
on the GPU by each schedule and compared with the oracle bit for bit: plain FP32 (with and without
early exit, full and truncated parity rows), HALF_PRC and CHAR_PRC (each with and without early
exit), learned offset/normalized
min-sum with random per-(iteration, edge) weights, in the layout each batch gets and in the slim
layout forced (one workgroup's codewords plus one); and the device encoder against the host one.
A few noisy codewords per pair keep the oracle's share to a few seconds.  Each case prints its
parameters before it runs."""
import numpy as np
import pytest

from tests.oracle_binding import OR_BG1, OR_BG2, OR_WIFI
from tests.test_gpu_char import _char_decode
from tests.test_gpu_decode import _gpu_decode, _noisy_llr
from tests.test_gpu_half import _half_decode
from tests.test_gpu_learned import _learned
from tests.test_oracle import _quantize

pytestmark = pytest.mark.gpu
KIND = {1: OR_BG1, 2: OR_BG2, 3: OR_WIFI}
LIFTING = sorted({a * 2 ** j for a in (2, 3, 5, 7, 9, 11, 13, 15) for j in range(8) if a * 2 ** j <= 384})
ROWS = {1: 46, 2: 42}
KB = {1: 22, 2: 10}
PAIRS = [(k, z) for k in (1, 2) for z in LIFTING] + [(3, z) for z in (4, 17, 96, 127)]


def _case(kind, Z, salt):
    """Parity rows and nominal rate of one case: every third pair keeps a random row count."""
    rng = np.random.default_rng(1000 * kind + Z + salt)
    if kind == 3:
        return 0, 0.5
    m = int(rng.integers(4, ROWS[kind] + 1)) if (Z + salt) % 3 == 0 else 0
    rows = m or ROWS[kind]
    return m, KB[kind] / (KB[kind] + rows - 2.0)


def test_lifting_set_is_complete():
    assert len(LIFTING) == 51 and len(PAIRS) == 106


def test_plain_every_lifting_size(nr, oracle, torch_gpu):
    for kind, Z in PAIRS:
        m, rate = _case(kind, Z, 0)
        print(f"case plain kind={kind} Z={Z} m={m}", flush=True)
        code = nr.Code(kind, Z, m)
        oc = oracle.code(KIND[kind], Z, m)
        llr, _ = _noisy_llr(nr, code, 3, 1.0, rate, seed=7 * Z + kind)
        hd, app, it = _gpu_decode(torch_gpu, code, llr, 4)
        ohd, oapp, _ = oracle.decode(oc, llr, 4)
        assert np.array_equal(hd, ohd), (kind, Z, m)
        assert np.array_equal(app, oapp), (kind, Z, m)
        assert (it == 4).all()
        # early exit at a higher Eb/N0: iteration counts and outputs equal the oracle's
        llr, _ = _noisy_llr(nr, code, 3, 3.0, rate, seed=11 * Z + kind)
        hd, app, it = _gpu_decode(torch_gpu, code, llr, 12, flags=nr.EARLY_EXIT)
        ohd, oapp, oit = oracle.decode(oc, llr, 12, early_exit=True)
        assert np.array_equal(it, oit), (kind, Z, m, it, oit)
        assert np.array_equal(hd, ohd) and np.array_equal(app, oapp), (kind, Z, m)
        code.close()


def test_plain_slim_layout_every_lifting_size(nr, oracle, torch_gpu):
    """The slim layout (single LDS read copy, degree-1 posteriors in device slots, persistent grid)
    forced at every pair: one workgroup's codewords plus one, so that a second block runs through
    the persistent loop, with early exit at an Eb/N0 where codewords stop at different
    iterations."""
    for kind, Z in PAIRS:
        m, rate = _case(kind, Z, 1)
        code = nr.Code(kind, Z, m)
        code.set_launch_layout(nr.LAYOUT_SLIM, 0)
        cpb = code.launch_layout(1)[1]
        n = min(cpb + 1, 400)
        print(f"case plain slim kind={kind} Z={Z} m={m} cpb={cpb} n={n}", flush=True)
        assert code.launch_layout(n)[:2] == (nr.LAYOUT_SLIM, cpb)
        oc = oracle.code(KIND[kind], Z, m)
        llr, _ = _noisy_llr(nr, code, n, 2.0, rate, seed=13 * Z + kind)
        hd, app, it = _gpu_decode(torch_gpu, code, llr, 8, flags=nr.EARLY_EXIT)
        ohd, oapp, oit = oracle.decode(oc, llr, 8, early_exit=True)
        assert np.array_equal(it, oit), (kind, Z, m)
        assert np.array_equal(hd, ohd), (kind, Z, m)
        assert np.array_equal(app.view(np.uint32), oapp.view(np.uint32)), (kind, Z, m)
        code.close()


def test_half_every_lifting_size(nr, oracle, torch_gpu):
    for kind, Z in PAIRS:
        m, rate = _case(kind, Z, 1)
        print(f"case half kind={kind} Z={Z} m={m}", flush=True)
        code = nr.Code(kind, Z, m)
        llr, _ = _noisy_llr(nr, code, 2, 1.0, rate, seed=5 * Z + kind)
        llr_h = llr.astype(np.float16).view(np.uint16)
        hd, app, _ = _half_decode(torch_gpu, code, llr_h, 4)
        ohd, oapp, _ = oracle.decode_half(oracle.code(KIND[kind], Z, m), llr_h, 4)
        assert np.array_equal(hd, ohd) and np.array_equal(app, oapp), (kind, Z, m)
        code.close()


def test_char_every_lifting_size(nr, oracle, torch_gpu):
    for kind, Z in PAIRS:
        m, rate = _case(kind, Z, 2)
        gain = np.float32(4.0 if Z % 2 else 1.0)  # odd Z: scaled LLRs, so the saturating rules fire
        print(f"case char kind={kind} Z={Z} m={m} gain={gain}", flush=True)
        code = nr.Code(kind, Z, m)
        llr, _ = _noisy_llr(nr, code, 2, 1.0, rate, seed=3 * Z + kind)
        q = _quantize(llr * gain)
        hd, app, _ = _char_decode(torch_gpu, code, q, 4)
        ohd, oapp, _, _ = oracle.decode_char(oracle.code(KIND[kind], Z, m), q, 4)
        assert np.array_equal(hd, ohd) and np.array_equal(app, oapp), (kind, Z, m)
        code.close()


def test_half_and_char_early_exit_every_lifting_size(nr, oracle, torch_gpu):
    """HALF_PRC and CHAR_PRC with early exit at every pair: five codewords at an Eb/N0 where they
    stop at different iterations (an odd count leaves the CHAR pair kernel's last codeword without
    a partner); iteration counts, decisions and posteriors equal the oracle's restatements."""
    stopped = {"half": 0, "char": 0}
    for kind, Z in PAIRS:
        m, rate = _case(kind, Z, 4)
        print(f"case half+char early exit kind={kind} Z={Z} m={m}", flush=True)
        code = nr.Code(kind, Z, m)
        oc = oracle.code(KIND[kind], Z, m)
        llr, _ = _noisy_llr(nr, code, 5, 3.0, rate, seed=23 * Z + kind)
        llr_h = llr.astype(np.float16).view(np.uint16)
        hd, app, it = _half_decode(torch_gpu, code, llr_h, 12, flags=nr.EARLY_EXIT)
        ohd, oapp, oit = oracle.decode_half(oc, llr_h, 12, early_exit=True)
        assert np.array_equal(it, oit), ("half", kind, Z, m, it, oit)
        assert np.array_equal(hd, ohd) and np.array_equal(app, oapp), ("half", kind, Z, m)
        stopped["half"] += int(it.min() < 12)
        q = _quantize(llr)
        hd, app, it = _char_decode(torch_gpu, code, q, 12, flags=nr.EARLY_EXIT)
        ohd, oapp, oit, _ = oracle.decode_char(oc, q, 12, early_exit=True)
        assert np.array_equal(it, oit), ("char", kind, Z, m, it, oit)
        assert np.array_equal(hd, ohd) and np.array_equal(app, oapp), ("char", kind, Z, m)
        stopped["char"] += int(it.min() < 12)
        code.close()
    # the Eb/N0 lets codewords stop early at most pairs, so the early exit itself is exercised
    assert min(stopped.values()) > len(PAIRS) // 2, stopped


def test_learned_every_lifting_size(nr, oracle, torch_gpu):
    T = 3
    for kind, Z in PAIRS:
        m, rate = _case(kind, Z, 0)
        print(f"case learned kind={kind} Z={Z} m={m}", flush=True)
        code = nr.Code(kind, Z, m)
        rng = np.random.default_rng(kind * 1000 + Z)
        w = np.stack([rng.uniform(0.5, 1.0, (T, code.nnz_full)), rng.uniform(-0.3, 0.5, (T, code.nnz_full))], -1)
        w = w.astype(np.float32)
        code.set_weights(w)
        llr, _ = _noisy_llr(nr, code, 2, 1.0, rate, seed=13 * Z + kind)
        hd, app, _ = _learned(torch_gpu, code, llr, T)
        # the library keeps weights for every edge of the full base graph; the oracle takes the
        # kept rows' edges only, which are the leading code.nnz ones (edges are numbered row-major)
        wk = w[:, :code.nnz, :]
        ohd, oapp = oracle.decode_noms(oracle.code(KIND[kind], Z, m), llr, T, wk[..., 0].reshape(-1),
                                       wk[..., 1].reshape(-1))
        assert np.array_equal(hd, ohd) and np.array_equal(app, oapp), (kind, Z, m)
        code.close()


def test_half_and_learned_slim_layout_every_lifting_size(nr, oracle, torch_gpu):
    """HALF_PRC and learned min-sum in the slim layout forced at every pair (a workgroup's
    codewords plus one, capped), against the oracle's restatements."""
    T = 3
    for kind, Z in PAIRS:
        m, rate = _case(kind, Z, 2)
        code = nr.Code(kind, Z, m)
        code.set_launch_layout(nr.LAYOUT_SLIM, 0)
        cpb = code.launch_layout(1)[1]
        n = min(cpb + 1, 200)
        print(f"case half+learned slim kind={kind} Z={Z} m={m} cpb={cpb} n={n}", flush=True)
        assert code.launch_layout(n)[:2] == (nr.LAYOUT_SLIM, cpb)
        oc = oracle.code(KIND[kind], Z, m)
        llr, _ = _noisy_llr(nr, code, n, 1.5, rate, seed=19 * Z + kind)
        llr_h = llr.astype(np.float16).view(np.uint16)
        hd, app, _ = _half_decode(torch_gpu, code, llr_h, 4)
        ohd, oapp, _ = oracle.decode_half(oc, llr_h, 4)
        assert np.array_equal(hd, ohd) and np.array_equal(app, oapp), ("half", kind, Z, m)
        rng = np.random.default_rng(kind * 1000 + Z + 7)
        w = np.stack([rng.uniform(0.5, 1.0, (T, code.nnz_full)), rng.uniform(-0.3, 0.5, (T, code.nnz_full))], -1)
        w = w.astype(np.float32)
        code.set_weights(w)
        hd, app, _ = _learned(torch_gpu, code, llr, T)
        wk = w[:, :code.nnz, :]
        ohd, oapp = oracle.decode_noms(oc, llr, T, wk[..., 0].reshape(-1), wk[..., 1].reshape(-1))
        assert np.array_equal(hd, ohd) and np.array_equal(app, oapp), ("learned", kind, Z, m)
        code.close()


def test_device_encoder_every_lifting_size(nr, oracle, torch_gpu):
    """ldpc_encode_async equals the host encoder on every pair; only bit 0 of an info byte counts;
    the codewords satisfy every kept parity check (oracle syndrome)."""
    torch = torch_gpu
    for kind, Z in PAIRS:
        m, _ = _case(kind, Z, 3)
        print(f"case encode kind={kind} Z={Z} m={m}", flush=True)
        code = nr.Code(kind, Z, m)
        rng = np.random.default_rng(17 * Z + kind)
        raw = rng.integers(0, 256, (5, code.K), dtype=np.uint8)  # any byte; bit 0 is the bit
        want = code.encode_host(raw & 1)
        d_info = torch.from_numpy(raw).cuda()
        d_cw = torch.full((5, code.n_vn), 7, dtype=torch.uint8, device="cuda")
        code.encode(d_info.data_ptr(), 5, d_cw.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        got = d_cw.cpu().numpy()
        assert np.array_equal(got, want), (kind, Z, m)
        oc = oracle.code(KIND[kind], Z, m)
        assert all(oracle.syndrome(oc, h) == 0 for h in got), (kind, Z, m)
        code.close()
