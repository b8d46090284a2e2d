"""The persistent mixed-batch decoder in the slim layout (768-thread workgroups, a single LDS read
copy, degree-1 posteriors in the object's device slots): every codeword equals the single-code
decoder of its (base graph, Z, parity rows) bit for bit -- hard decisions, posteriors, iteration
counts -- with and without early exit, with learned weights and the plain fallback, through a
captured graph, and after autotuning.  The reference decodes one code per build
(src/gpu_decoder_top.cu:16-52); mixed batches have no reference counterpart beyond that loop."""
import numpy as np
import pytest

from nr_gpu_ldpc_amd import mixed as mx
from tests.test_gpu_mixed import _compare, _learned_reference, _mixed, _reference

pytestmark = pytest.mark.gpu


def _slim(nr, n):
    m = nr.Mixed(n)
    m.set_layout(nr.LAYOUT_SLIM)
    assert m.get_layout() == nr.LAYOUT_SLIM and m.plan_info()["threads"] == 768
    return m


@pytest.mark.parametrize("early", [False, True])
def test_mixed_slim_matches_single_code_decoder(nr, torch_gpu, early):
    torch = torch_gpu
    print(f"case mixed slim vs single early={early}", flush=True)
    sizes = [2, 3, 7, 12, 15, 26, 52, 60, 104, 144, 208, 240, 352, 384]
    b = mx.draw_batch(240, seed=12, sizes=sizes, ebn0=(0.5, 3.0))
    llr, _ = mx.make_channel(torch, b, seed=6)
    flags = nr.EARLY_EXIT if early else 0
    got = _mixed(nr, torch, b, llr, 12, flags, _slim(nr, 240))
    _compare(b, got, _reference(nr, torch, b, llr, 12, flags, {}))
    if early:
        assert got[2].min() < 12


def test_mixed_slim_large_batch_equals_dup_layout(nr, torch_gpu):
    """8192 codewords over every lifting size with early exit (tiles that converge at different
    iterations, short last tiles, lanes beyond a tile's codewords): the slim layout equals the dup
    layout bit for bit."""
    torch = torch_gpu
    print("case mixed slim 8192 all sizes", flush=True)
    b = mx.draw_batch(8192, seed=51, ebn0=(0.0, 3.0))
    llr, _ = mx.make_channel(torch, b, seed=52)
    want = _mixed(nr, torch, b, llr, 16, nr.EARLY_EXIT)
    got = _mixed(nr, torch, b, llr, 16, nr.EARLY_EXIT, _slim(nr, 8192))
    assert all(np.array_equal(g.view(np.uint8), w.view(np.uint8)) for g, w in zip(got, want))
    assert got[2].min() < 16 and got[2].max() == 16


@pytest.mark.parametrize("flags", ["learned+fallback", "learned+fallback+early"])
def test_mixed_slim_learned_matches_single_code(nr, torch_gpu, flags):
    torch = torch_gpu
    print(f"case mixed slim {flags}", flush=True)
    f = nr.LEARNED | nr.FALLBACK_PLAIN | (nr.EARLY_EXIT if "early" in flags else 0)
    T = 8
    rng = np.random.default_rng(5)
    weights = {}
    for key in ((1, 384), (1, 64), (2, 208), (2, 24)):
        nnz = 316 if key[0] == 1 else 197
        weights[key] = np.stack([rng.uniform(0.4, 1.0, (T, nnz)), rng.uniform(-0.2, 0.6, (T, nnz))],
                                -1).astype(np.float32)
    b = mx.draw_batch(160, seed=9, sizes=[24, 64, 208, 384], ebn0=(1.0, 3.0))
    llr, _ = mx.make_channel(torch, b, seed=10)
    m = _slim(nr, 160)
    for (k, z), w in weights.items():
        m.set_weights(k, z, w)
    got = _mixed(nr, torch, b, llr, T, f, m)
    _compare(b, got, _learned_reference(nr, torch, b, llr, T, f, weights))


def test_mixed_slim_graph_replay_and_layout_switch(nr, torch_gpu):
    """A graph captured in the slim layout replays two plans like direct decodes; switching the
    layout discards it (replay refuses until captured again) and the dup layout decodes the same."""
    torch = torch_gpu
    print("case mixed slim graph replay", flush=True)
    b1 = mx.draw_batch(150, seed=23, sizes=[8, 24, 64, 208, 384])
    b2 = mx.draw_batch(90, seed=24, sizes=[5, 40, 96, 384])
    cap_f = max(b1["llr_floats"], b2["llr_floats"])
    cap_h = max(b1["hd_total"], b2["hd_total"])
    m = _slim(nr, 200)
    d_llr = torch.zeros(cap_f, dtype=torch.float32, device="cuda")
    d_hd = torch.zeros(cap_h, dtype=torch.uint8, device="cuda")
    h_llr = torch.zeros(cap_f, dtype=torch.float32).pin_memory()
    h_hd = torch.zeros(cap_h, dtype=torch.uint8).pin_memory()
    s = torch.cuda.Stream()
    m.plan(b1["desc"])
    m.capture(s.cuda_stream, d_llr.data_ptr(), 10, d_hd.data_ptr(), flags=nr.EARLY_EXIT,
              llr_host_ptr=h_llr.data_ptr(), llr_floats=cap_f, hd_host_ptr=h_hd.data_ptr(), hd_bytes=cap_h)
    for rep, b in enumerate((b1, b2)):
        llr, _ = mx.make_channel(torch, b, seed=40 + rep)
        torch.cuda.synchronize()
        h_llr[:b["llr_floats"]].copy_(llr.cpu())
        m.plan(b["desc"])
        m.replay(s.cuda_stream)
        s.synchronize()
        want = _mixed(nr, torch, b, llr, 10, nr.EARLY_EXIT)
        assert np.array_equal(h_hd.numpy()[:b["hd_total"]], want[0][:b["hd_total"]]), f"replay {rep}"
    m.set_layout(nr.LAYOUT_DUP)
    m.plan(b2["desc"])
    with pytest.raises(nr.LdpcError):
        m.replay(s.cuda_stream)


def test_mixed_slim_autotune(nr, torch_gpu):
    torch = torch_gpu
    print("case mixed slim autotune BG2 Z=52", flush=True)
    m = _slim(nr, 4096)
    best, table = m.autotune(2, 52, 10, 4096)
    assert best in table and table[best] == min(table.values())
    assert max(table) * 52 <= 768 and max(table) > 512 // 52
    b = mx.describe([2] * 30, [52] * 30, [42] * 15 + [10] * 15, ebn0=np.full(30, 1.0))
    llr, _ = mx.make_channel(torch, b, seed=4)
    _compare(b, _mixed(nr, torch, b, llr, 8, 0, m), _reference(nr, torch, b, llr, 8, 0, {}))


def test_mixed_layout_autotune_keeps_the_faster_layout(nr, torch_gpu):
    """The layout tuner times both layouts on the caller's batch and flags, keeps the faster, and
    the object then decodes exactly as before."""
    torch = torch_gpu
    print("case mixed layout autotune", flush=True)
    b = mx.draw_batch(1200, seed=61, ebn0=(0.5, 3.0))
    llr, _ = mx.make_channel(torch, b, seed=62)
    want = _mixed(nr, torch, b, llr, 10, nr.EARLY_EXIT)
    m = nr.Mixed(1200)
    d_hd = torch.zeros(b["hd_total"], dtype=torch.uint8, device="cuda")
    best, us = m.autotune_layout(b["desc"], llr.data_ptr(), 10, d_hd.data_ptr(), flags=nr.EARLY_EXIT)
    assert all(t > 0 for t in us.values())
    assert best == min(us, key=us.get) and m.get_layout() == best
    got = _mixed(nr, torch, b, llr, 10, nr.EARLY_EXIT, m)
    assert all(np.array_equal(g.view(np.uint8), w.view(np.uint8)) for g, w in zip(got, want))
