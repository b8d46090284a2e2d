"""Launch layouts of the FP32 decoder at every code shape, and the launch-shape autotuner.

The dup layout (two read copies in LDS, degree-1 posteriors in registers) and the slim layout (one
read copy with a wrap select, degree-1 posteriors in a device slot per CU, up to 768 threads) run
the same arithmetic in the same order, so every shape of either layout must equal the oracle bit
for bit: compiled-in offsets (BG1 Z=384, BG2 Z=208, the 12x24 code at Z=96) and the any-Z kernels,
truncated parity rows, odd counts, early exit and learned weights.  The reference's decode loop is
src/gpu_decoder_top.cu:16-52 with the kernels of src/ldpc_kernels.cu:72-468; its launch shape is
fixed at build time (src/setup_exp.cu:131-166), which ldpc_set_launch_layout / ldpc_autotune
replace.
"""
import numpy as np
import pytest

from tests.oracle_binding import OR_BG1, OR_BG2, OR_WIFI
from tests.test_gpu_decode import _gpu_decode, _noisy_llr
from tests.test_gpu_learned import _learned

pytestmark = pytest.mark.gpu
ORC = {1: OR_BG1, 2: OR_BG2, 3: OR_WIFI}
RATE = {1: 1 / 3, 2: 1 / 5, 3: 1 / 2}

SHAPES = [
    # kind, Z, kept parity rows (0: all)
    (1, 384, 0),
    (1, 104, 0),
    (1, 15, 12),
    (2, 208, 0),
    (2, 13, 0),
    (2, 72, 9),
    (3, 96, 0),
    (3, 27, 0),
]


def _cpb_max(nr, code, layout):
    """The largest codewords per workgroup the library accepts for `layout` (probed downwards)."""
    limit = 512 if layout == nr.LAYOUT_DUP else 768
    for c in range(limit // code.Z, 0, -1):
        try:
            code.set_launch_layout(layout, c)
            return c
        except nr.LdpcError:
            continue
    raise AssertionError("no codewords-per-workgroup count fits")


@pytest.mark.parametrize("kind,Z,m", SHAPES)
@pytest.mark.parametrize("layout_name", ["dup", "slim"])
def test_layouts_match_oracle(nr, oracle, torch_gpu, kind, Z, m, layout_name):
    layout = nr.LAYOUT_DUP if layout_name == "dup" else nr.LAYOUT_SLIM
    code = nr.Code(kind, Z, m)
    cmax = _cpb_max(nr, code, layout)
    ocode = oracle.code(ORC[kind], Z, m)
    llr, _ = _noisy_llr(nr, code, 2 * cmax + 3, 1.0, RATE[kind], seed=7 * Z + kind)
    for cpb in sorted({1, cmax}):
        for early in (False, True):
            print(f"case layout={layout_name} kind={kind} Z={Z} m={m} cpb={cpb} early={early}", flush=True)
            code.set_launch_layout(layout, cpb)
            assert code.launch_layout(llr.shape[0])[:2] == (layout, cpb)
            hd, app, it = _gpu_decode(torch_gpu, code, llr, 6, flags=nr.EARLY_EXIT if early else 0)
            ohd, oapp, oit = oracle.decode(ocode, llr, 6, early_exit=early)
            assert np.array_equal(hd, ohd), f"hard decisions differ in {int((hd != ohd).sum())} bits"
            assert np.array_equal(app.view(np.uint32), oapp.view(np.uint32))
            assert np.array_equal(it, oit)


@pytest.mark.parametrize("kind,Z", [(1, 104), (2, 208), (2, 13)])
def test_slim_learned_matches_oracle_any_size(nr, oracle, torch_gpu, kind, Z):
    print(f"case slim learned kind={kind} Z={Z}", flush=True)
    code = nr.Code(kind, Z)
    T = 5
    rng = np.random.default_rng(Z)
    w = np.stack([rng.uniform(0.5, 1.0, (T, code.nnz_full)), rng.uniform(-0.3, 0.5, (T, code.nnz_full))], -1)
    w = w.astype(np.float32)
    code.set_weights(w)
    code.set_launch_layout(nr.LAYOUT_SLIM, 0)
    llr, _ = _noisy_llr(nr, code, 9, 1.0, RATE[kind], seed=Z + 1)
    hd, app, _ = _learned(torch_gpu, code, llr, T)
    ohd, oapp = oracle.decode_noms(oracle.code(ORC[kind], Z), llr, T, w[..., 0].reshape(-1), w[..., 1].reshape(-1))
    assert np.array_equal(hd, ohd)
    assert np.array_equal(app.view(np.uint32), oapp.view(np.uint32))


def test_slim_persistent_rounds_small_code(nr, oracle, torch_gpu):
    """More codewords than one round of the slim grid (one workgroup per CU at the largest
    count): every workgroup takes several blocks; the whole batch equals the dup layout and
    sampled codewords (both ends, block boundaries) equal the oracle."""
    cus = torch_gpu.cuda.get_device_properties(0).multi_processor_count
    code = nr.Code(2, 13)
    cmax = _cpb_max(nr, code, nr.LAYOUT_SLIM)
    n = 2 * cus * cmax + 5
    print(f"case slim rounds BG2 Z=13 cpb={cmax} n={n}", flush=True)
    llr, _ = _noisy_llr(nr, code, n, 1.5, 1 / 5, seed=1313)
    hd_s, app_s, it_s = _gpu_decode(torch_gpu, code, llr, 8, flags=nr.EARLY_EXIT)
    code.set_launch_layout(nr.LAYOUT_DUP, 0)
    hd_d, app_d, it_d = _gpu_decode(torch_gpu, code, llr, 8, flags=nr.EARLY_EXIT)
    assert np.array_equal(hd_s, hd_d) and np.array_equal(it_s, it_d)
    assert np.array_equal(app_s.view(np.uint32), app_d.view(np.uint32))
    pick = np.array([0, 1, cmax - 1, cmax, cus * cmax - 1, cus * cmax, n - 2, n - 1])
    ohd, oapp, oit = oracle.decode(oracle.code(OR_BG2, 13), llr[pick], 8, early_exit=True)
    assert np.array_equal(hd_s[pick], ohd) and np.array_equal(it_s[pick], oit)
    assert np.array_equal(app_s[pick].view(np.uint32), oapp.view(np.uint32))


@pytest.mark.parametrize("kind,Z", [(2, 13), (1, 384)])
def test_autotune_installs_a_shape_that_decodes_exactly(nr, oracle, torch_gpu, kind, Z):
    """The tuner times both layouts at several codewords-per-workgroup counts; afterwards the
    per-batch choice uses its measurements and still decodes bit-exactly."""
    print(f"case autotune kind={kind} Z={Z}", flush=True)
    code = nr.Code(kind, Z)
    cands = code.autotune(n_iter=4)
    layouts = {c["layout"] for c in cands}
    assert layouts == {nr.LAYOUT_DUP, nr.LAYOUT_SLIM}
    assert all(c["round_us"] > 0 and c["cw_per_cu"] >= c["cpb"] >= 1 for c in cands)
    for lay in layouts:
        cpbs = [c["cpb"] for c in cands if c["layout"] == lay]
        assert len(cpbs) == len(set(cpbs))
    cus = torch_gpu.cuda.get_device_properties(0).multi_processor_count
    for n in (1, cus, 16 * cus):
        lay, cpb, wg = code.launch_layout(n)
        assert lay in layouts and cpb >= 1 and wg >= 1
    llr, _ = _noisy_llr(nr, code, 7, 1.0, RATE[kind], seed=99)
    hd, app, _ = _gpu_decode(torch_gpu, code, llr, 5)
    ohd, oapp, _ = oracle.decode(oracle.code(ORC[kind], Z), llr, 5)
    assert np.array_equal(hd, ohd) and np.array_equal(app.view(np.uint32), oapp.view(np.uint32))


def test_layout_model_restores_an_autotune_result(nr, oracle, torch_gpu):
    """An autotune result read with ldpc_get_layout_model and installed into a fresh context with
    ldpc_set_layout_model (an autotune cache across processes) gives the same per-batch launch
    choices as the tuned context, and the restored slim shape decodes bit-exactly."""
    print("case layout model restore BG2 Z=52", flush=True)
    tuned = nr.Code(2, 52)
    tuned.autotune(n_iter=4)
    models = {lay: tuned.layout_model(lay) for lay in (nr.LAYOUT_DUP, nr.LAYOUT_SLIM)}
    assert models[nr.LAYOUT_DUP]["cpb"] >= 1 and models[nr.LAYOUT_SLIM]["cpb"] >= 1
    fresh = nr.Code(2, 52)
    for lay, m in models.items():
        fresh.set_layout_model(lay, m["cpb"], m["cw_per_cu"], m["round_us"])
        assert fresh.layout_model(lay) == m
    cus = torch_gpu.cuda.get_device_properties(0).multi_processor_count
    for n in (1, 7, cus, 3 * cus + 5, 40 * cus):
        assert fresh.launch_layout(n) == tuned.launch_layout(n), n
    fresh.set_layout_model(nr.LAYOUT_SLIM, models[nr.LAYOUT_SLIM]["cpb"], 1, 1e-3)  # slim always cheaper
    n = 3 * models[nr.LAYOUT_SLIM]["cpb"] + 1
    assert fresh.launch_layout(n)[:2] == (nr.LAYOUT_SLIM, models[nr.LAYOUT_SLIM]["cpb"])
    llr, _ = _noisy_llr(nr, fresh, n, 1.5, RATE[2], seed=5)
    hd, app, its = _gpu_decode(torch_gpu, fresh, llr, 6, flags=nr.EARLY_EXIT)
    ohd, oapp, oit = oracle.decode(oracle.code(ORC[2], 52), llr, 6, early_exit=True)
    assert np.array_equal(its, oit)
    assert np.array_equal(hd, ohd) and np.array_equal(app.view(np.uint32), oapp.view(np.uint32))
    fresh.set_layout_model(nr.LAYOUT_SLIM, 0)
    assert fresh.launch_layout(n)[0] == nr.LAYOUT_DUP


def test_set_launch_selects_layout_by_count(nr, torch_gpu):
    """ldpc_set_launch: counts that fit the dup layout take it, larger ones the slim layout."""
    code = nr.Code(1, 384)
    code.set_launch(2)
    assert code.launch_layout(10)[:2] == (nr.LAYOUT_SLIM, 2)
    code.set_launch(1)
    assert code.launch_layout(10)[:2] == (nr.LAYOUT_DUP, 1)
    code.set_launch(0)
    bg2 = nr.Code(2, 13)
    bg2.set_launch(512 // 13)
    assert bg2.launch_layout(10)[:2] == (nr.LAYOUT_DUP, 512 // 13)
    bg2.set_launch(512 // 13 + 1)
    assert bg2.launch_layout(10)[:2] == (nr.LAYOUT_SLIM, 512 // 13 + 1)
    with pytest.raises(nr.LdpcError):
        bg2.set_launch(768 // 13 + 1)


@pytest.mark.parametrize("kind,Z,m", [(1, 2, 43), (2, 3, 11), (1, 5, 5)])
def test_packed_padding_bits_are_zero(nr, torch_gpu, kind, Z, m):
    """Lifting sizes below 8 with n_vn % 8 != 0: the bits of the last packed byte beyond column
    n_vn are 0 for every decoder (both FP32 layouts, HALF_PRC, CHAR_PRC), whatever the output
    buffer held before."""
    torch = torch_gpu
    code = nr.Code(kind, Z, m)
    assert code.n_vn % 8 != 0
    n = 40
    llr, _ = _noisy_llr(nr, code, n, 1.0, 0.5, seed=Z * 10 + m)
    d_llr = torch.from_numpy(llr).cuda()
    d_half = torch.empty((n, code.n_vn), dtype=torch.int16, device="cuda")
    nr.llr_to_half(d_llr.data_ptr(), d_llr.numel(), d_half.data_ptr())
    d_i8 = torch.clamp(torch.floor(d_llr * 2.0), -128, 127).to(torch.int8)
    pad_mask = (0xFF << (code.n_vn % 8)) & 0xFF

    def decode(kind_name):
        hd = torch.full((n, code.hd_bytes), 0xFF, dtype=torch.uint8, device="cuda")
        if kind_name == "dup":
            code.set_launch_layout(nr.LAYOUT_DUP, 0)
            code.decode(d_llr.data_ptr(), n, 5, hd.data_ptr())
        elif kind_name == "slim":
            code.set_launch_layout(nr.LAYOUT_SLIM, 0)
            code.decode(d_llr.data_ptr(), n, 5, hd.data_ptr())
        elif kind_name == "half":
            code.decode_half(d_half.data_ptr(), n, 5, hd.data_ptr())
        else:
            code.decode_char(d_i8.data_ptr(), n, 5, hd.data_ptr())
        torch.cuda.synchronize()
        return hd.cpu().numpy()

    for name in ("dup", "slim", "half", "char"):
        print(f"case padding kind={kind} Z={Z} m={m} {name}", flush=True)
        last = decode(name)[:, -1]
        assert not (last & pad_mask).any(), f"{name}: padding bits set in {int(((last & pad_mask) != 0).sum())} codewords"


@pytest.mark.parametrize("kind,Z,m", [(1, 384, 0), (1, 104, 12), (2, 208, 0), (2, 13, 0), (3, 96, 0)])
def test_half_prc_slim_layout_matches_oracle(nr, oracle, torch_gpu, kind, Z, m):
    """The HALF_PRC schedule in both layouts: binary16 storage of the degree-1 posteriors happens at
    their single update (slim) instead of in registers after the row loop (dup); equal to the
    oracle's HALF_PRC restatement, with and without early exit."""
    torch = torch_gpu
    code = nr.Code(kind, Z, m)
    n = 2 * _cpb_max(nr, code, nr.LAYOUT_SLIM) + 3
    llr, _ = _noisy_llr(nr, code, n, 1.0, RATE[kind], seed=3 * Z + m)
    d_llr = torch.from_numpy(llr).cuda()
    d_half = torch.empty((n, code.n_vn), dtype=torch.int16, device="cuda")
    nr.llr_to_half(d_llr.data_ptr(), d_llr.numel(), d_half.data_ptr())
    torch.cuda.synchronize()
    half_llr = d_half.cpu().numpy().view(np.uint16)
    ocode = oracle.code(ORC[kind], Z, m)
    for layout, name in ((nr.LAYOUT_SLIM, "slim"), (nr.LAYOUT_DUP, "dup")):
        code.set_launch_layout(layout, 0)
        for early in (False, True):
            print(f"case half layout={name} kind={kind} Z={Z} m={m} early={early}", flush=True)
            d_hd = torch.zeros((n, code.hd_bytes), dtype=torch.uint8, device="cuda")
            d_app = torch.zeros((n, code.n_vn), dtype=torch.float32, device="cuda")
            d_it = torch.zeros(n, dtype=torch.int32, device="cuda")
            code.decode_half(d_half.data_ptr(), n, 6, d_hd.data_ptr(), d_app.data_ptr(), d_it.data_ptr(),
                             flags=nr.EARLY_EXIT if early else 0)
            torch.cuda.synchronize()
            hd = np.unpackbits(d_hd.cpu().numpy(), axis=1, bitorder="little")[:, :code.n_vn]
            ohd, oapp, oit = oracle.decode_half(ocode, half_llr, 6, early_exit=early)
            assert np.array_equal(hd, ohd)
            assert np.array_equal(d_app.cpu().numpy().view(np.uint32), oapp.view(np.uint32))
            assert np.array_equal(d_it.cpu().numpy(), oit)
