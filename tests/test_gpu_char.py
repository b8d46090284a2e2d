"""The reference's CHAR_PRC schedule (its default build) on the GPU against the oracle
(bit-exact hard decisions, posteriors and iteration counts): signed-char channel LLRs from
quantize(0.5, x, 8); R, dt and the APP between iterations stored as signed chars."""
import numpy as np
import pytest

from tests.oracle_binding import OR_BG1, OR_BG2, OR_WIFI
from tests.test_gpu_decode import _noisy_llr
from tests.test_oracle import _quantize

pytestmark = pytest.mark.gpu
KIND = {1: OR_BG1, 2: OR_BG2, 3: OR_WIFI}


def _char_decode(torch, code, llr_q, n_iter, flags=0):
    n = llr_q.shape[0]
    d_llr = torch.from_numpy(llr_q).cuda()
    d_hd = torch.zeros((n, code.hd_bytes), dtype=torch.uint8, device="cuda")
    d_app = torch.zeros((n, code.n_vn), dtype=torch.float32, device="cuda")
    d_it = torch.zeros(n, dtype=torch.int32, device="cuda")
    code.decode_char(d_llr.data_ptr(), n, n_iter, d_hd.data_ptr(), d_app.data_ptr(), d_it.data_ptr(), flags=flags)
    torch.cuda.synchronize()
    hd = np.unpackbits(d_hd.cpu().numpy(), axis=1, bitorder="little")[:, :code.n_vn]
    return hd, d_app.cpu().numpy(), d_it.cpu().numpy()


# (kind, Z, codewords, iterations, Eb/N0, rate, LLR gain): a gain above 1 scales the channel LLRs
# before quantisation so that saturated values and out-of-range conversions occur.
# Odd codeword counts leave the last codeword of the two-per-lane kernel without a partner.
CASES = [(1, 384, 4, 10, 1.25, 1 / 3, 1.0), (1, 384, 4, 10, 1.5, 1 / 3, 8.0), (2, 208, 6, 10, 0.5, 0.2, 1.0),
         (3, 96, 16, 5, 2.0, 0.5, 1.0), (1, 15, 20, 6, 1.0, 1 / 3, 4.0), (2, 13, 20, 6, 1.0, 0.2, 1.0),
         (1, 384, 5, 10, 1.5, 1 / 3, 1.0), (2, 208, 7, 10, 1.0, 0.2, 8.0), (3, 96, 9, 5, 3.0, 0.5, 1.0),
         (2, 52, 3, 8, 1.5, 0.2, 2.0), (1, 7, 1, 4, 1.0, 1 / 3, 1.0)]


@pytest.mark.parametrize("kind,Z,n_cw,T,ebn0,rate,gain", CASES)
def test_char_matches_oracle(nr, oracle, torch_gpu, kind, Z, n_cw, T, ebn0, rate, gain):
    print(f"case char kind={kind} Z={Z} T={T} gain={gain}", flush=True)
    code = nr.Code(kind, Z)
    llr, _ = _noisy_llr(nr, code, n_cw, ebn0, rate, seed=Z + 5 * kind)
    q = _quantize(llr * np.float32(gain))
    hd, app, it = _char_decode(torch_gpu, code, q, T)
    ohd, oapp, oit, oor = oracle.decode_char(oracle.code(KIND[kind], Z), q, T)
    assert np.array_equal(hd, ohd)
    assert np.array_equal(app, oapp)
    assert (it == T).all()
    if gain > 4:
        assert oor > 0 and (np.abs(app) >= 127).any()  # the saturating rules were exercised


@pytest.mark.parametrize("kind,Z,n_cw,ebn0,rate", [(1, 64, 40, 2.0, 1 / 3), (1, 384, 9, 1.75, 1 / 3),
                                                   (2, 208, 11, 1.5, 0.2), (3, 96, 17, 3.5, 0.5)])
def test_char_early_exit(nr, oracle, torch_gpu, kind, Z, n_cw, ebn0, rate):
    """Codewords that converge at different iterations (the two codewords of one lane included)
    stop at the oracle's iteration with the oracle's output."""
    print(f"case char early exit kind={kind} Z={Z} n={n_cw}", flush=True)
    code = nr.Code(kind, Z)
    llr, _ = _noisy_llr(nr, code, n_cw, ebn0, rate, seed=9 + Z)
    q = _quantize(llr)
    hd, app, it = _char_decode(torch_gpu, code, q, 20, flags=nr.EARLY_EXIT)
    ohd, oapp, oit, _ = oracle.decode_char(oracle.code(KIND[kind], Z), q, 20, early_exit=True)
    assert np.array_equal(it, oit) and it.min() < 20
    assert len(set(it.tolist())) > 1
    assert np.array_equal(hd, ohd) and np.array_equal(app, oapp)
