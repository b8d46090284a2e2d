"""Persistent mixed-batch decoder on the GPU: every codeword equals the single-code decoder of
its (base graph, Z, parity rows) on the same LLRs -- hard decisions, APP and iteration counts
bit for bit -- with and without early exit, through a captured graph, and after autotuning."""
import numpy as np
import pytest

from nr_gpu_ldpc_amd import mixed as mx

pytestmark = pytest.mark.gpu


def _reference(nr, torch, b, llr, n_iter, flags, contexts):
    """Decode every codeword with the single-code decoder (grouped by (kind, Z, m))."""
    n = len(b["kind"])
    hd = [None] * n
    app = [None] * n
    its = np.zeros(n, dtype=np.int32)
    keys = sorted(set(zip(b["kind"].tolist(), b["Z"].tolist(), b["m_rows"].tolist())))
    off = b["desc"]["llr_offset"]
    for key in keys:
        idx = np.nonzero((b["kind"] == key[0]) & (b["Z"] == key[1]) & (b["m_rows"] == key[2]))[0]
        code = contexts.get(key) or contexts.setdefault(key, nr.Code(*key))
        L = code.n_vn
        x = torch.stack([llr[int(off[i]):int(off[i]) + L] for i in idx]).contiguous()
        d_hd = torch.zeros((len(idx), code.hd_bytes), dtype=torch.uint8, device="cuda")
        d_app = torch.zeros((len(idx), L), dtype=torch.float32, device="cuda")
        d_it = torch.zeros(len(idx), dtype=torch.int32, device="cuda")
        code.decode(x.data_ptr(), len(idx), n_iter, d_hd.data_ptr(), d_app.data_ptr(), d_it.data_ptr(), flags=flags)
        h, a, t = d_hd.cpu().numpy(), d_app.cpu().numpy(), d_it.cpu().numpy()
        for k, i in enumerate(idx):
            hd[i], app[i], its[i] = h[k], a[k], t[k]
    return hd, app, its


def _mixed(nr, torch, b, llr, n_iter, flags, m=None):
    m = m or nr.Mixed(len(b["kind"]))
    m.plan(b["desc"])
    d_hd = torch.zeros(b["hd_total"], dtype=torch.uint8, device="cuda")
    d_app = torch.zeros(b["llr_floats"], dtype=torch.float32, device="cuda")
    d_it = torch.zeros(len(b["kind"]), dtype=torch.int32, device="cuda")
    m.decode(llr.data_ptr(), n_iter, d_hd.data_ptr(), d_app.data_ptr(), d_it.data_ptr(), flags=flags)
    torch.cuda.synchronize()
    return d_hd.cpu().numpy(), d_app.cpu().numpy(), d_it.cpu().numpy()


def _compare(b, got, want):
    hd, app, its = got
    rhd, rapp, rits = want
    d = b["desc"]
    for i in range(len(b["kind"])):
        h0, nb = int(d["hd_offset"][i]), int(b["hd_bytes"][i])
        a0, nv = int(d["llr_offset"][i]), int(b["n_vn"][i])
        assert np.array_equal(hd[h0:h0 + nb], rhd[i]), f"hard decisions of codeword {i} {d[i]}"
        assert np.array_equal(app[a0:a0 + nv].view(np.uint32), rapp[i].view(np.uint32)), f"APP of codeword {i}"
    assert np.array_equal(its, rits)


@pytest.mark.parametrize("early", [False, True])
def test_mixed_matches_single_code_decoder(nr, torch_gpu, early):
    torch = torch_gpu
    print(f"case mixed vs single early={early}", flush=True)
    sizes = [2, 3, 7, 12, 15, 26, 52, 60, 104, 144, 208, 240, 352, 384]
    b = mx.draw_batch(240, seed=11, sizes=sizes, ebn0=(0.5, 3.0))
    llr, _ = mx.make_channel(torch, b, seed=5)
    flags = nr.EARLY_EXIT if early else 0
    got = _mixed(nr, torch, b, llr, 12, flags)
    want = _reference(nr, torch, b, llr, 12, flags, {})
    _compare(b, got, want)
    if early:
        assert got[2].min() < 12


@pytest.mark.parametrize("early", [True, False])
def test_large_mixed_batch_matches_single_code_decoder(nr, torch_gpu, early):
    """8192 codewords at Z 40..104 (tiles of 2 to 12 codewords that converge at different
    iterations): every codeword equals the single-code decoder bit for bit, iteration counts
    included."""
    torch = torch_gpu
    print(f"case mixed 8192 Z 40..104 early={early}", flush=True)
    b = mx.draw_batch(8192, seed=41, sizes=[z for z in mx.LIFTING_SIZES if 40 <= z <= 104], ebn0=(0.0, 3.0))
    llr, _ = mx.make_channel(torch, b, seed=43)
    m = nr.Mixed(len(b["kind"]))
    flags = nr.EARLY_EXIT if early else 0
    got = _mixed(nr, torch, b, llr, 20, flags, m)
    want = _reference(nr, torch, b, llr, 20, flags, {})
    _compare(b, got, want)
    if early:
        assert got[2].min() < 20 and got[2].max() == 20


def test_mixed_graph_replay_follows_the_plan(nr, torch_gpu):
    """One captured graph (plan upload, LLR copy in, decode, HD copy out) replayed for two
    different plans gives what direct decodes give."""
    torch = torch_gpu
    print("case mixed graph replay", flush=True)
    b1 = mx.draw_batch(150, seed=21, sizes=[8, 24, 64, 208, 384])
    b2 = mx.draw_batch(90, seed=22, sizes=[5, 40, 96, 384])
    cap_f = max(b1["llr_floats"], b2["llr_floats"])
    cap_h = max(b1["hd_total"], b2["hd_total"])
    m = nr.Mixed(200)
    d_llr = torch.zeros(cap_f, dtype=torch.float32, device="cuda")
    d_hd = torch.zeros(cap_h, dtype=torch.uint8, device="cuda")
    d_it = torch.zeros(200, dtype=torch.int32, device="cuda")
    h_llr = torch.zeros(cap_f, dtype=torch.float32).pin_memory()
    h_hd = torch.zeros(cap_h, dtype=torch.uint8).pin_memory()
    s = torch.cuda.Stream()
    m.plan(b1["desc"])
    m.capture(s.cuda_stream, d_llr.data_ptr(), 10, d_hd.data_ptr(), 0, d_it.data_ptr(), flags=nr.EARLY_EXIT,
              llr_host_ptr=h_llr.data_ptr(), llr_floats=cap_f, hd_host_ptr=h_hd.data_ptr(), hd_bytes=cap_h)
    for rep, b in enumerate((b1, b2, b1)):
        llr, _ = mx.make_channel(torch, b, seed=30 + rep)
        torch.cuda.synchronize()
        h_llr[:b["llr_floats"]].copy_(llr.cpu())
        h_hd.fill_(0xAB)
        m.plan(b["desc"])
        m.replay(s.cuda_stream)
        s.synchronize()
        # the download is cut to the bytes this plan addresses
        assert (h_hd.numpy()[b["hd_total"]:] == 0xAB).all(), f"replay {rep}"
        want = _mixed(nr, torch, b, llr, 10, nr.EARLY_EXIT)
        assert np.array_equal(h_hd.numpy()[:b["hd_total"]], want[0][:b["hd_total"]]), f"replay {rep}"
        assert np.array_equal(d_it.cpu().numpy()[:len(b["kind"])], want[2]), f"replay {rep}"


def test_concurrent_streams_match_lone_decodes(nr, torch_gpu):
    """Three mixed objects, each with its graph captured on its own stream, replayed round-robin
    with several batches in flight at once: every stream's decisions and iteration counts equal
    a decode of its batch with nothing else running."""
    torch = torch_gpu
    print("case mixed concurrent streams", flush=True)
    slots = []
    for k in range(3):
        b = mx.draw_batch(60 + 40 * k, seed=70 + k, sizes=[16, 52, 208, 384], ebn0=(0.5, 3.0))
        llr, _ = mx.make_channel(torch, b, seed=80 + k)
        want = _mixed(nr, torch, b, llr, 10, nr.EARLY_EXIT)
        h_llr = torch.zeros(b["llr_floats"], dtype=torch.float32).pin_memory()
        h_llr.copy_(llr.cpu())
        h_hd = torch.zeros(b["hd_total"], dtype=torch.uint8).pin_memory()
        d_llr = torch.zeros(b["llr_floats"], dtype=torch.float32, device="cuda")
        d_hd = torch.zeros(b["hd_total"], dtype=torch.uint8, device="cuda")
        d_it = torch.zeros(len(b["kind"]), dtype=torch.int32, device="cuda")
        s = torch.cuda.Stream()
        m = nr.Mixed(len(b["kind"]))
        m.plan(b["desc"])
        m.capture(s.cuda_stream, d_llr.data_ptr(), 10, d_hd.data_ptr(), 0, d_it.data_ptr(), flags=nr.EARLY_EXIT,
                  llr_host_ptr=h_llr.data_ptr(), llr_floats=b["llr_floats"], hd_host_ptr=h_hd.data_ptr(),
                  hd_bytes=b["hd_total"])
        # every buffer the captured graph reads or writes stays referenced until the end: a
        # tensor dropped here would go back to torch's caching allocators and be handed to the next
        # slot while this graph still copies into or out of it
        slots.append((b, m, s, h_hd, d_it, want, (d_llr, d_hd, h_llr)))
    torch.cuda.synchronize()  # the zero fills above ran on the default stream; replays use others
    for rep in range(12):
        b, m, s = slots[rep % 3][:3]
        m.plan(b["desc"])
        m.replay(s.cuda_stream)
    torch.cuda.synchronize()
    for k, (b, m, s, h_hd, d_it, want, _) in enumerate(slots):
        assert np.array_equal(h_hd.numpy(), want[0][:b["hd_total"]]), f"stream {k}"
        assert np.array_equal(d_it.cpu().numpy(), want[2]), f"stream {k}"


def test_autotune_picks_a_timed_candidate(nr, torch_gpu):
    torch = torch_gpu
    print("case mixed autotune BG2 Z=52", flush=True)
    m = nr.Mixed(2048)
    best, table = m.autotune(2, 52, 10, 2048)
    assert best in table and all(t > 0 for t in table.values())
    assert table[best] == min(table.values())
    assert m.get_cpb(2, 52) == best
    # decoding after tuning is still exact
    b = mx.describe([2] * 30, [52] * 30, [42] * 15 + [10] * 15, ebn0=np.full(30, 1.0))
    llr, _ = mx.make_channel(torch, b, seed=3)
    _compare(b, _mixed(nr, torch, b, llr, 8, 0, m), _reference(nr, torch, b, llr, 8, 0, {}))


def _learned_reference(nr, torch, b, llr, n_iter, flags, weights):
    n = len(b["kind"])
    hd, app, its = [None] * n, [None] * n, np.zeros(n, dtype=np.int32)
    off = b["desc"]["llr_offset"]
    for key in sorted(set(zip(b["kind"].tolist(), b["Z"].tolist(), b["m_rows"].tolist()))):
        idx = np.nonzero((b["kind"] == key[0]) & (b["Z"] == key[1]) & (b["m_rows"] == key[2]))[0]
        code = nr.Code(*key)
        L = code.n_vn
        x = torch.stack([llr[int(off[i]):int(off[i]) + L] for i in idx]).contiguous()
        d_hd = torch.zeros((len(idx), code.hd_bytes), dtype=torch.uint8, device="cuda")
        d_app = torch.zeros((len(idx), L), dtype=torch.float32, device="cuda")
        d_it = torch.zeros(len(idx), dtype=torch.int32, device="cuda")
        w = weights.get(key[:2])
        if w is None:
            code.decode(x.data_ptr(), len(idx), n_iter, d_hd.data_ptr(), d_app.data_ptr(), d_it.data_ptr(),
                        flags=flags & nr.EARLY_EXIT)
        else:
            code.set_weights(w)
            code.decode_learned(x.data_ptr(), len(idx), n_iter, d_hd.data_ptr(), d_app.data_ptr(), d_it.data_ptr(),
                                flags=flags & (nr.EARLY_EXIT | nr.FALLBACK_PLAIN))
        h, a, t = d_hd.cpu().numpy(), d_app.cpu().numpy(), d_it.cpu().numpy()
        for k, i in enumerate(idx):
            hd[i], app[i], its[i] = h[k], a[k], t[k]
    return hd, app, its


@pytest.mark.parametrize("flags", ["learned", "learned+fallback", "learned+fallback+early"])
def test_mixed_learned_matches_single_code(nr, torch_gpu, flags):
    """Codes with weights set decode with them (others plain min-sum); with fallback, the plain
    re-decode inside the persistent launch replaces exactly the codewords the single-code
    learned path replaces."""
    torch = torch_gpu
    print(f"case mixed learned {flags}", flush=True)
    f = nr.LEARNED
    if "fallback" in flags:
        f |= nr.FALLBACK_PLAIN
    if "early" in flags:
        f |= nr.EARLY_EXIT
    T = 8
    rng = np.random.default_rng(4)
    weights = {}
    for key in ((1, 64), (2, 208), (2, 24)):
        nnz = 316 if key[0] == 1 else 197
        w = np.stack([rng.uniform(0.4, 1.0, (T, nnz)), rng.uniform(-0.2, 0.6, (T, nnz))], -1).astype(np.float32)
        weights[key] = w
    b = mx.draw_batch(160, seed=8, sizes=[24, 64, 208, 384], ebn0=(1.0, 3.0))
    llr, _ = mx.make_channel(torch, b, seed=9)
    m = nr.Mixed(len(b["kind"]))
    for (k, z), w in weights.items():
        m.set_weights(k, z, w)
    got = _mixed(nr, torch, b, llr, T, f, m)
    want = _learned_reference(nr, torch, b, llr, T, f, weights)
    _compare(b, got, want)


@pytest.mark.parametrize("fmt", ["i8", "f16"])
def test_mixed_narrow_inputs_equal_float_decode_of_converted_llrs(nr, oracle, torch_gpu, fmt):
    """int8 (quantize(0.5, x, 8), LLR = 0.5 q) and binary16 channel LLRs, converted on load: the
    decode equals the float decode of the converted values bit for bit (direct launch and a
    replayed graph whose upload moves 1 or 2 bytes per LLR), and sampled codewords equal the
    oracle."""
    torch = torch_gpu
    print(f"case mixed input {fmt}", flush=True)
    b = mx.draw_batch(200, seed=31, sizes=[13, 52, 208, 384], ebn0=(0.5, 3.0))
    llr, _ = mx.make_channel(torch, b, seed=32)
    x = llr.cpu().numpy()
    if fmt == "i8":
        narrow = mx.quantize(x)
        conv = narrow.astype(np.float32) * np.float32(0.5)
        ntype, code, scale = torch.int8, nr.LLR_I8, 0.5
    else:
        narrow = x.astype(np.float16)
        conv = narrow.astype(np.float32)
        ntype, code, scale = torch.float16, nr.LLR_F16, 1.0
    want = _mixed(nr, torch, b, torch.from_numpy(conv).cuda(), 12, nr.EARLY_EXIT)
    m = nr.Mixed(len(b["kind"]))
    m.set_input(code, scale)
    d_narrow = torch.from_numpy(narrow.view(np.int8) if fmt == "i8" else narrow.view(np.int16)).cuda()
    got = _mixed(nr, torch, b, d_narrow, 12, nr.EARLY_EXIT, m)
    assert all(np.array_equal(g.view(np.uint8), w.view(np.uint8)) for g, w in zip(got, want))
    # replayed graph with the narrow host buffer
    h_in = torch.from_numpy(d_narrow.cpu().numpy()).pin_memory()
    h_hd = torch.zeros(b["hd_total"], dtype=torch.uint8).pin_memory()
    d_in = torch.zeros_like(d_narrow)
    d_hd = torch.zeros(b["hd_total"], dtype=torch.uint8, device="cuda")
    s = torch.cuda.Stream()
    m.plan(b["desc"])
    m.capture(s.cuda_stream, d_in.data_ptr(), 12, d_hd.data_ptr(), 0, 0, flags=nr.EARLY_EXIT,
              llr_host_ptr=h_in.data_ptr(), llr_floats=b["llr_floats"], hd_host_ptr=h_hd.data_ptr(),
              hd_bytes=b["hd_total"])
    m.replay(s.cuda_stream)
    s.synchronize()
    assert np.array_equal(h_hd.numpy(), want[0][:b["hd_total"]])
    # the oracle on the converted values, a few codewords of each (kind, Z)
    d = b["desc"]
    for i in range(0, len(b["kind"]), 17):
        a0, nv = int(d["llr_offset"][i]), int(b["n_vn"][i])
        h0, nb = int(d["hd_offset"][i]), int(b["hd_bytes"][i])
        oc = oracle.code(int(b["kind"][i]), int(b["Z"][i]), int(b["m_rows"][i]))
        ohd, oapp, oit = oracle.decode(oc, conv[a0:a0 + nv][None, :], 12, early_exit=True)
        assert np.array_equal(got[0][h0:h0 + nb], np.packbits(ohd[0], bitorder="little")), i
        assert np.array_equal(got[1][a0:a0 + nv].view(np.uint32), oapp[0].view(np.uint32)), i
        assert got[2][i] == oit[0], i


def test_mixed_reduced_grid_decodes_the_same(nr, torch_gpu):
    """A persistent grid of a quarter (and of one workgroup) of the device decodes the same."""
    torch = torch_gpu
    print("case mixed reduced grid", flush=True)
    b = mx.draw_batch(300, seed=41, sizes=[16, 104, 208, 384], ebn0=(0.5, 3.0))
    llr, _ = mx.make_channel(torch, b, seed=42)
    want = _mixed(nr, torch, b, llr, 10, nr.EARLY_EXIT)
    full = nr.Mixed(300).plan_info()["grid"]
    for g in (full // 4, 1):
        m = nr.Mixed(300)
        m.set_grid(g)
        m.plan(b["desc"])
        assert m.plan_info()["grid"] == g
        got = _mixed(nr, torch, b, llr, 10, nr.EARLY_EXIT, m)
        assert all(np.array_equal(x.view(np.uint8), y.view(np.uint8)) for x, y in zip(got, want)), g
