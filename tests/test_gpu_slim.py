"""The slim decoder layout at BG1 Z=384: two codewords per 12-wavefront workgroup (one per CU),
a single LDS read copy and the degree-1 APPs in per-workgroup device slots.  It computes the same
arithmetic in the same order as the default layout, so its outputs are compared bit for bit with
the oracle (small batches) and with the default layout (full batches).  The reference's loop is
src/gpu_decoder_top.cu:16-52 with the kernels of src/ldpc_kernels.cu:72-468.
"""
import numpy as np
import pytest

from tests.oracle_binding import OR_BG1
from tests.test_gpu_decode import _gpu_decode, _noisy_llr
from tests.test_gpu_learned import _learned

pytestmark = pytest.mark.gpu
SLIM = 2     # codewords per workgroup of the slim layout at BG1 Z=384
DEFAULT = 1  # the default layout's


@pytest.mark.parametrize("m,n_cw,n_iter,ebn0,rate", [
    (0, 1, 10, 1.25, 1 / 3),   # one workgroup with a lone codeword
    (0, 7, 10, 1.25, 1 / 3),   # odd count: the last workgroup's second slot is empty
    (0, 4, 1, 0.5, 1 / 3),
    (10, 9, 8, 2.5, 22 / 30),  # parity rows truncated: unused degree-1 slots
])
def test_slim_matches_oracle(nr, oracle, torch_gpu, m, n_cw, n_iter, ebn0, rate):
    print(f"case slim BG1 Z=384 m={m} n_cw={n_cw} T={n_iter}", flush=True)
    code = nr.Code(1, 384, m)
    code.set_launch(SLIM)
    llr, _ = _noisy_llr(nr, code, n_cw, ebn0, rate, seed=31 + n_cw)
    hd, app, iters = _gpu_decode(torch_gpu, code, llr, n_iter)
    ohd, oapp, _ = oracle.decode(oracle.code(OR_BG1, 384, m), llr, n_iter)
    assert np.array_equal(hd, ohd), f"hard decisions differ in {int((hd != ohd).sum())} bits"
    assert np.array_equal(app.view(np.uint32), oapp.view(np.uint32))
    assert (iters == n_iter).all()


def test_slim_early_exit_matches_oracle(nr, oracle, torch_gpu):
    """Codewords that converge at once, late and never, paired in workgroups in every
    combination: iteration counts, decisions and APP equal the oracle's."""
    print("case slim early exit BG1 Z=384", flush=True)
    code = nr.Code(1, 384)
    code.set_launch(SLIM)
    n = 21
    good, _ = _noisy_llr(nr, code, n, 40.0, 1 / 3, seed=1)
    mid, _ = _noisy_llr(nr, code, n, 1.6, 1 / 3, seed=2)
    rng = np.random.default_rng(3)
    bad = (rng.standard_normal((n, code.n_vn)) * 3).astype(np.float32)
    bad[:, :2 * 384] = 0
    pick = np.array([0, 1, 1, 0, 2, 0, 1, 2, 2, 1, 0, 0, 1, 1, 2, 2, 0, 2, 1, 0, 1])
    llr = np.where((pick == 0)[:, None], good, np.where((pick == 1)[:, None], mid, bad)).astype(np.float32)
    hd, app, iters = _gpu_decode(torch_gpu, code, llr, 20, flags=nr.EARLY_EXIT)
    ohd, oapp, oit = oracle.decode(oracle.code(OR_BG1, 384), llr, 20, early_exit=True)
    assert np.array_equal(iters, oit)
    assert np.array_equal(hd, ohd) and np.array_equal(app.view(np.uint32), oapp.view(np.uint32))
    assert iters.min() <= 2 and iters.max() == 20


def test_slim_learned_matches_oracle(nr, oracle, torch_gpu):
    print("case slim learned BG1 Z=384", flush=True)
    code = nr.Code(1, 384)
    code.set_launch(SLIM)
    T = 6
    rng = np.random.default_rng(384)
    w = np.stack([rng.uniform(0.5, 1.0, (T, code.nnz_full)), rng.uniform(-0.3, 0.5, (T, code.nnz_full))], -1)
    w = w.astype(np.float32)
    code.set_weights(w)
    llr, _ = _noisy_llr(nr, code, 5, 1.0, 1 / 3, seed=385)
    hd, app, _ = _learned(torch_gpu, code, llr, T)
    ohd, oapp = oracle.decode_noms(oracle.code(OR_BG1, 384), llr, T, w[..., 0].reshape(-1), w[..., 1].reshape(-1))
    assert np.array_equal(hd, ohd)
    assert np.array_equal(app.view(np.uint32), oapp.view(np.uint32))


@pytest.mark.parametrize("flags_name", ["none", "early"])
def test_slim_full_batch_equals_default_layout(nr, oracle, torch_gpu, flags_name):
    """4097 codewords (the persistent grid's workgroups take 8 or 9 blocks each, one slot holds a
    lone codeword): the batch-size choice (the slim layout here) and the forced slim layout equal
    the default layout bit for bit; sampled codewords equal the oracle."""
    print(f"case slim full batch BG1 Z=384 flags={flags_name}", flush=True)
    flags = nr.EARLY_EXIT if flags_name == "early" else 0
    code = nr.Code(1, 384)
    n = 4097
    llr, _ = _noisy_llr(nr, code, n, 1.5, 1 / 3, seed=77)
    code.set_launch(DEFAULT)
    hd1, app1, it1 = _gpu_decode(torch_gpu, code, llr, 10, flags=flags)
    code.set_launch(0)  # chosen per batch: slim at this size
    hd0, app0, it0 = _gpu_decode(torch_gpu, code, llr, 10, flags=flags)
    code.set_launch(SLIM)
    hd2, app2, it2 = _gpu_decode(torch_gpu, code, llr, 10, flags=flags)
    for hd, app, it in ((hd0, app0, it0), (hd2, app2, it2)):
        assert np.array_equal(hd, hd1) and np.array_equal(it, it1)
        assert np.array_equal(app.view(np.uint32), app1.view(np.uint32))
    pick = np.array([0, 1, 2047, 2048, 4095, 4096])
    ohd, oapp, oit = oracle.decode(oracle.code(OR_BG1, 384), llr[pick], 10, early_exit=flags != 0)
    assert np.array_equal(hd2[pick], ohd) and np.array_equal(app2[pick].view(np.uint32), oapp.view(np.uint32))
    assert np.array_equal(it2[pick], oit)


def test_slim_learned_fallback_equals_default_layout(nr, torch_gpu):
    """Learned decode with plain fallback and early exit (the re-decode runs over a compacted
    list): the slim layout equals the default layout on a full batch."""
    from nr_gpu_ldpc_amd import noms
    print("case slim learned fallback BG1 Z=384", flush=True)
    code = nr.Code(1, 384)
    rec = noms.load_weights(noms.weights_path(1, 384, 10))
    code.set_weights(rec["weights"])
    llr, _ = _noisy_llr(nr, code, 1500, 1.5, 1 / 3, seed=78)
    flags = nr.EARLY_EXIT | nr.FALLBACK_PLAIN
    code.set_launch(DEFAULT)
    ref = _learned(torch_gpu, code, llr, 10, flags=flags)
    code.set_launch(SLIM)
    got = _learned(torch_gpu, code, llr, 10, flags=flags)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[2], ref[2])
    assert np.array_equal(got[1].view(np.uint32), ref[1].view(np.uint32))
    assert 0 < int((ref[2] == 10).sum()) < 1500  # both early exits and full decodes occurred


def test_slim_decodes_on_two_streams(nr, torch_gpu):
    """One context decoding two batches on two streams at once: its shared degree-1 slots are
    taken one launch after the other, and both outputs equal lone decodes."""
    torch = torch_gpu
    print("case slim two streams BG1 Z=384", flush=True)
    code = nr.Code(1, 384)
    code.set_launch(SLIM)
    llr_a, _ = _noisy_llr(nr, code, 600, 1.4, 1 / 3, seed=90)
    llr_b, _ = _noisy_llr(nr, code, 600, 1.6, 1 / 3, seed=91)
    want_a = _gpu_decode(torch, code, llr_a, 10)
    want_b = _gpu_decode(torch, code, llr_b, 10)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for llr, s in ((llr_a, s1), (llr_b, s2)):
        d_llr = torch.from_numpy(llr).cuda()
        d_hd = torch.zeros((600, code.hd_bytes), dtype=torch.uint8, device="cuda")
        d_app = torch.zeros((600, code.n_vn), dtype=torch.float32, device="cuda")
        outs.append((d_llr, d_hd, d_app, s))
    torch.cuda.synchronize()
    for d_llr, d_hd, d_app, s in outs:
        code.decode(d_llr.data_ptr(), 600, 10, d_hd.data_ptr(), d_app.data_ptr(), 0, stream=s.cuda_stream)
    torch.cuda.synchronize()
    for (d_llr, d_hd, d_app, _), want in zip(outs, (want_a, want_b)):
        hd = np.unpackbits(d_hd.cpu().numpy(), axis=1, bitorder="little")[:, :code.n_vn]
        assert np.array_equal(hd, want[0])
        assert np.array_equal(d_app.cpu().numpy().view(np.uint32), want[1].view(np.uint32))


def test_launch_shape_reports_the_layout(nr, torch_gpu):
    """The slim layout is chosen for batches of more than one codeword per CU (one workgroup per
    CU), the default layout for smaller ones, and ldpc_set_launch fixes either."""
    cus = torch_gpu.cuda.get_device_properties(0).multi_processor_count
    code = nr.Code(1, 384)
    assert code.launch_shape(8 * cus) == (SLIM, cus)
    assert code.launch_shape(cus) == (DEFAULT, cus)
    code.set_launch(SLIM)
    assert code.launch_shape(3) == (SLIM, 2)
    code.set_launch(DEFAULT)
    assert code.launch_shape(8 * cus) == (DEFAULT, 8 * cus)
    bg2 = nr.Code(2, 208)
    lay, cpb, wg = bg2.launch_layout(4096)
    if lay == nr.LAYOUT_DUP:
        assert wg == (4096 + cpb - 1) // cpb
    else:
        assert wg == min((4096 + cpb - 1) // cpb, cus)
