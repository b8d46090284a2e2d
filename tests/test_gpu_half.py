"""The reference's HALF_PRC schedule on the GPU against the oracle (bit-exact hard decisions and
posteriors): binary16 channel LLRs; R, dt and the APP between iterations stored in binary16."""
import numpy as np
import pytest

from tests.oracle_binding import OR_BG1, OR_BG2, OR_WIFI
from tests.test_gpu_decode import _noisy_llr

pytestmark = pytest.mark.gpu
KIND = {1: OR_BG1, 2: OR_BG2, 3: OR_WIFI}


def _half_decode(torch, code, llr_h, n_iter, flags=0):
    n = llr_h.shape[0]
    d_llr = torch.from_numpy(llr_h.view(np.int16)).cuda()
    d_hd = torch.zeros((n, code.hd_bytes), dtype=torch.uint8, device="cuda")
    d_app = torch.zeros((n, code.n_vn), dtype=torch.float32, device="cuda")
    d_it = torch.zeros(n, dtype=torch.int32, device="cuda")
    code.decode_half(d_llr.data_ptr(), n, n_iter, d_hd.data_ptr(), d_app.data_ptr(), d_it.data_ptr(), flags=flags)
    torch.cuda.synchronize()
    hd = np.unpackbits(d_hd.cpu().numpy(), axis=1, bitorder="little")[:, :code.n_vn]
    return hd, d_app.cpu().numpy(), d_it.cpu().numpy()


@pytest.mark.parametrize("kind,Z,n_cw,T,ebn0,rate", [(1, 384, 4, 10, 1.25, 1 / 3), (2, 208, 6, 10, 0.5, 0.2),
                                                     (3, 96, 16, 5, 2.0, 0.5), (1, 15, 20, 6, 1.0, 1 / 3),
                                                     (2, 13, 20, 6, 1.0, 0.2)])
def test_half_matches_oracle(nr, oracle, torch_gpu, kind, Z, n_cw, T, ebn0, rate):
    print(f"case half kind={kind} Z={Z} T={T}", flush=True)
    code = nr.Code(kind, Z)
    llr, _ = _noisy_llr(nr, code, n_cw, ebn0, rate, seed=Z + 3 * kind)
    llr_h = llr.astype(np.float16).view(np.uint16)
    hd, app, it = _half_decode(torch_gpu, code, llr_h, T)
    ohd, oapp, oit = oracle.decode_half(oracle.code(KIND[kind], Z), llr_h, T)
    assert np.array_equal(hd, ohd)
    assert np.array_equal(app, oapp)
    assert (it == T).all()


def test_half_early_exit_and_conversion(nr, oracle, torch_gpu):
    torch = torch_gpu
    print("case half early exit BG1 Z=64", flush=True)
    code = nr.Code(1, 64)
    llr, _ = _noisy_llr(nr, code, 40, 2.0, 1 / 3, seed=9)
    d_f = torch.from_numpy(llr).cuda()
    d_h = torch.zeros(llr.shape, dtype=torch.int16, device="cuda")
    nr.llr_to_half(d_f.data_ptr(), llr.size, d_h.data_ptr())
    torch.cuda.synchronize()
    llr_h = d_h.cpu().numpy().view(np.uint16)
    assert np.array_equal(llr_h, llr.astype(np.float16).view(np.uint16))  # round to nearest even
    hd, app, it = _half_decode(torch, code, llr_h, 20, flags=nr.EARLY_EXIT)
    ohd, oapp, oit = oracle.decode_half(oracle.code(OR_BG1, 64), llr_h, 20, early_exit=True)
    assert np.array_equal(it, oit) and it.min() < 20
    assert np.array_equal(hd, ohd) and np.array_equal(app, oapp)


@pytest.mark.parametrize("kind,Z,layout_name", [(1, 384, "slim"), (1, 384, "dup"), (2, 13, "slim"), (3, 8, "dup")])
def test_half_special_values(nr, oracle, torch_gpu, kind, Z, layout_name):
    """binary16 edge cases as channel LLRs: signed zeros, exact ties, subnormals, values near the
    largest finite binary16 (65504) and their mixture, in both LDS layouts; decisions and
    posteriors equal the oracle's HALF_PRC restatement (posteriors up to the sign of an exactly
    zero value, which the decoder canonicalises)."""
    torch = torch_gpu
    print(f"case half special values kind={kind} Z={Z} layout={layout_name}", flush=True)
    code = nr.Code(kind, Z)
    code.set_launch_layout(nr.LAYOUT_SLIM if layout_name == "slim" else nr.LAYOUT_DUP, 0)
    rng = np.random.default_rng(Z + kind)
    n = code.n_vn
    sign = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    rows = [np.full(n, 60000.0), sign * 1.0, sign * 0.0, sign * 6e-8 * rng.integers(1, 1000, n),
            sign * rng.choice([0.0, 6e-6, 1.0, 3.0, 3e4, 65504.0], n), rng.standard_normal(n) * 2000]
    llr = np.stack(rows).astype(np.float16)
    llr[:, :code.punct * Z] = 0
    llr_h = llr.view(np.uint16)
    assert (np.abs(llr[3]) < 6.2e-5).all() and (np.abs(llr[3]) > 0).any()  # subnormal binary16
    hd, app, _ = _half_decode(torch, code, llr_h, 6)
    ohd, oapp, _ = oracle.decode_half(oracle.code(KIND[kind], Z), llr_h, 6)
    assert np.array_equal(hd, ohd)
    assert np.array_equal(app, oapp)
    nz = oapp != 0
    assert np.array_equal(app.view(np.uint32)[nz], oapp.view(np.uint32)[nz])
