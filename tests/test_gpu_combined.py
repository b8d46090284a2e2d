"""The combined stream's decode against the oracle: BG1 Z=384 and BG2 Z=208 codewords at mixed
rates with the committed learned weights, plain min-sum fallback on syndrome failure and
per-codeword early exit, through one captured graph replayed for batches of 1, 257 and 8192
codewords.

The oracle's rule for each codeword (the reference's decode loop, src/gpu_decoder_top.cu:16-52,
with the learned weights in place of the fixed 0.75 of src/ldpc_kernels.cu:164,295): learned
decode with early exit; when its final hard decision fails the syndrome check, plain min-sum
with early exit decodes the codeword again, and its output (hard decisions, APP, iterations)
replaces the learned one only where its hard decision passes the check.

Also the mixed object's guards: a replay refuses a plan that addresses more than the buffers
given at capture or holds weights for another iteration count, and launches of one object on
two streams are ordered."""

import numpy as np
import pytest

from nr_gpu_ldpc_amd import mixed as mx
from tests.oracle_binding import OR_BG1, OR_BG2

pytestmark = pytest.mark.gpu
T = 10


def _weights():
    from nr_gpu_ldpc_amd import noms
    out = {}
    for kind, Z in ((1, 384), (2, 208)):
        rec = noms.load_weights(noms.weights_path(kind, Z, T))
        out[(kind, Z)] = rec["weights"]
    return out


def _stream_batch(n, seed, ebn0=(1.0, 3.5)):
    """n codewords of the combined stream: BG1 Z=384 or BG2 Z=208 (50/50), rate uniform in
    [1/5, 8/9] by parity-row truncation, Eb/N0 uniform in the waterfall range."""
    rng = np.random.default_rng(seed)
    kind = rng.integers(1, 3, n)
    Z = np.where(kind == 1, 384, 208)
    rate = rng.uniform(0.2, 8 / 9, n)
    m = np.where(kind == 1, mx.rows_for_rate(1, rate), mx.rows_for_rate(2, rate))
    return mx.describe(kind, Z, m, ebn0=rng.uniform(ebn0[0], ebn0[1], n))


def _oracle_stream_decode(oracle, weights, kind, Z, m, llr):
    """The oracle's learned + early exit + plain fallback decode of one codeword."""
    oc = oracle.code(OR_BG1 if kind == 1 else OR_BG2, Z, m)
    nnz = int((oracle.shifts(oc) >= 0).sum())
    w = weights[(kind, Z)][:, :nnz, :]  # edges of the kept rows are the leading ones (row-major)
    x = llr[None, :]
    hd, app, it = oracle.decode_noms(oc, x, T, w[..., 0].reshape(-1), w[..., 1].reshape(-1), early_exit=True,
                                     with_iters=True, threads=1)
    state = "learned"
    if oracle.syndrome(oc, hd[0]) != 0:
        phd, papp, pit = oracle.decode(oc, x, T, early_exit=True, threads=1)
        if oracle.syndrome(oc, phd[0]) == 0:
            hd, app, it, state = phd, papp, pit, "replaced"
        else:
            state = "kept"
    return np.packbits(hd[0], bitorder="little"), app[0], int(it[0]), state


def test_combined_stream_replays_match_oracle(nr, oracle, torch_gpu):
    torch = torch_gpu
    weights = _weights()
    batches = [(1, 501), (257, 502), (8192, 503)]
    drawn = [_stream_batch(n, seed) for n, seed in batches]
    cap_f = max(b["llr_floats"] for b in drawn)
    cap_h = max(b["hd_total"] for b in drawn)
    m = nr.Mixed(8192)
    for (k, z), w in weights.items():
        m.set_weights(k, z, w)
    flags = nr.LEARNED | nr.FALLBACK_PLAIN | nr.EARLY_EXIT
    d_llr = torch.zeros(cap_f, dtype=torch.float32, device="cuda")
    d_app = torch.zeros(cap_f, dtype=torch.float32, device="cuda")
    d_hd = torch.zeros(cap_h, dtype=torch.uint8, device="cuda")
    d_it = torch.zeros(8192, dtype=torch.int32, device="cuda")
    h_llr = torch.zeros(cap_f, dtype=torch.float32).pin_memory()
    h_hd = torch.zeros(cap_h, dtype=torch.uint8).pin_memory()
    s = torch.cuda.Stream()
    m.plan(drawn[0]["desc"])
    m.capture(s.cuda_stream, d_llr.data_ptr(), T, d_hd.data_ptr(), d_app.data_ptr(), d_it.data_ptr(), flags=flags,
              llr_host_ptr=h_llr.data_ptr(), llr_floats=cap_f, hd_host_ptr=h_hd.data_ptr(), hd_bytes=cap_h)
    states = {"learned": 0, "replaced": 0, "kept": 0}
    early = 0
    for (n, seed), b in zip(batches, drawn):
        print(f"case combined stream batch {n} seed {seed}", flush=True)
        llr, _ = mx.make_channel(torch, b, seed=seed + 1000)
        torch.cuda.synchronize()
        h_llr[:b["llr_floats"]].copy_(llr.cpu())
        m.plan(b["desc"])
        m.replay(s.cuda_stream)
        s.synchronize()
        hd = h_hd.numpy()
        app = d_app.cpu().numpy()
        its = d_it.cpu().numpy()[:n]
        host_llr = h_llr.numpy()
        rng = np.random.default_rng(seed)
        # a sample of the batch: both ends, random codewords, codewords that ran every iteration
        # (the fallback candidates), and among those high-rate BG2 ones, where plain min-sum
        # rescues learned failures most often (the weights were trained at rate 1/5)
        full = np.nonzero(its == T)[0]
        hi_bg2 = full[(b["kind"][full] == 2) & (b["m_rows"][full] <= 16)]
        pick = set([0, n - 1])
        pick.update(rng.choice(n, min(n, 24), replace=False).tolist())
        for cand in (full, hi_bg2):
            pick.update(rng.choice(cand, min(len(cand), 24), replace=False).tolist() if len(cand) else [])
        d = b["desc"]
        for i in sorted(pick):
            a0, nv = int(d["llr_offset"][i]), int(b["n_vn"][i])
            h0, nb = int(d["hd_offset"][i]), int(b["hd_bytes"][i])
            whd, wapp, wit, state = _oracle_stream_decode(oracle, weights, int(b["kind"][i]), int(b["Z"][i]),
                                                          int(b["m_rows"][i]), host_llr[a0:a0 + nv])
            assert np.array_equal(hd[h0:h0 + nb], whd), f"batch {n}: hard decisions of codeword {i} ({state})"
            assert np.array_equal(app[a0:a0 + nv].view(np.uint32), wapp.view(np.uint32)), \
                f"batch {n}: APP of codeword {i} ({state})"
            assert its[i] == wit, f"batch {n}: iterations of codeword {i} ({state})"
            states[state] += 1
            early += wit < T
    # every branch of the rule was exercised
    assert states["learned"] > 0 and states["replaced"] > 0 and states["kept"] > 0, states
    assert early > 0


def test_replay_refuses_a_plan_beyond_the_captured_buffers(nr, torch_gpu):
    """A plan that addresses more LLR floats or hard-decision bytes than the capture was given,
    or that holds learned weights for another iteration count, is refused before anything is
    enqueued; a plan that fits replays as before."""
    torch = torch_gpu
    print("case mixed replay guards", flush=True)
    small = mx.describe([2] * 6, [52] * 6, [20] * 6, ebn0=np.full(6, 2.0))
    big = mx.describe([2] * 9, [52] * 9, [20] * 9, ebn0=np.full(9, 2.0))
    m = nr.Mixed(16)
    d_llr = torch.zeros(small["llr_floats"], dtype=torch.float32, device="cuda")
    d_hd = torch.zeros(small["hd_total"], dtype=torch.uint8, device="cuda")
    h_llr = torch.zeros(small["llr_floats"], dtype=torch.float32).pin_memory()
    h_hd = torch.zeros(small["hd_total"], dtype=torch.uint8).pin_memory()
    s = torch.cuda.Stream()
    w8 = np.stack([np.full((8, 197), 0.75), np.zeros((8, 197))], -1).astype(np.float32)
    m.set_weights(2, 52, w8)
    m.plan(small["desc"])
    # a host buffer without its size is refused
    with pytest.raises(nr.LdpcError):
        m.capture(s.cuda_stream, d_llr.data_ptr(), 8, d_hd.data_ptr(), flags=nr.LEARNED,
                  llr_host_ptr=h_llr.data_ptr(), llr_floats=0)
    m.capture(s.cuda_stream, d_llr.data_ptr(), 8, d_hd.data_ptr(), flags=nr.LEARNED,
              llr_host_ptr=h_llr.data_ptr(), llr_floats=small["llr_floats"], hd_host_ptr=h_hd.data_ptr(),
              hd_bytes=small["hd_total"])
    llr, _ = mx.make_channel(torch, small, seed=3)
    torch.cuda.synchronize()
    h_llr.copy_(llr.cpu())
    m.replay(s.cuda_stream)
    s.synchronize()
    first = h_hd.numpy().copy()
    m.plan(big["desc"])
    with pytest.raises(nr.LdpcError, match="given at capture"):
        m.replay(s.cuda_stream)
    # weights for another iteration count (re-planned: tiles point at the new table)
    w5 = np.stack([np.full((5, 197), 0.75), np.zeros((5, 197))], -1).astype(np.float32)
    m.set_weights(2, 52, w5)
    m.plan(small["desc"])
    with pytest.raises(nr.LdpcError, match="iteration count"):
        m.replay(s.cuda_stream)
    m.set_weights(2, 52, w8)
    m.plan(small["desc"])
    h_hd.fill_(0)
    m.replay(s.cuda_stream)
    s.synchronize()
    assert np.array_equal(h_hd.numpy(), first)


def test_one_object_on_two_streams_is_ordered(nr, torch_gpu):
    """Two decodes of one mixed object's plan, enqueued back to back on two different streams
    into different buffers: both equal a lone decode (the second waits for the first, which
    shares the object's work counter)."""
    torch = torch_gpu
    print("case mixed one object two streams", flush=True)
    b = mx.draw_batch(300, seed=61, sizes=[16, 52, 208, 384], ebn0=(0.5, 3.0))
    llr, _ = mx.make_channel(torch, b, seed=62)
    m = nr.Mixed(len(b["kind"]))
    m.plan(b["desc"])
    lone_hd = torch.zeros(b["hd_total"], dtype=torch.uint8, device="cuda")
    lone_it = torch.zeros(len(b["kind"]), dtype=torch.int32, device="cuda")
    m.decode(llr.data_ptr(), 10, lone_hd.data_ptr(), 0, lone_it.data_ptr(), flags=nr.EARLY_EXIT)
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [(torch.zeros_like(lone_hd), torch.zeros_like(lone_it)) for _ in streams]
    torch.cuda.synchronize()
    for rep in range(3):
        for s, (h, it) in zip(streams, outs):
            m.decode(llr.data_ptr(), 10, h.data_ptr(), 0, it.data_ptr(), flags=nr.EARLY_EXIT, stream=s.cuda_stream)
    torch.cuda.synchronize()
    for k, (h, it) in enumerate(outs):
        assert torch.equal(h, lone_hd), f"stream {k}"
        assert torch.equal(it, lone_it), f"stream {k}"
