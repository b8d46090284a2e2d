#!/usr/bin/env python3
"""Benchmark of the decode hot path on one GPU.

Headline (``value``): flooding scaled min-sum, 5G NR BG1 Z=384, T=10 iterations, batches of
4096 codewords, information Gbps.  A step = one decode launch over one batch whose LLRs are
already resident in HBM (generated on the device by the Philox channel kernel before timing).

Also measured, each in its own object of the same JSON line:
  * ``mixed_batch``: 16384 codewords of random base graph / lifting size / rate / Eb/N0 in
    one persistent launch, at most 20 iterations with early exit (autotuned tiles);
  * ``combined_stream``: BG1 Z=384 and BG2 Z=208 codewords at mixed rates in batches of
    1..8192, each batch an LLR upload, a replayed HIP graph (plan upload, decode) and a
    hard-decision download, plain min-sum against learned weights with plain fallback; batch
    by batch (latency) and pipelined over two streams (throughput);
  * ``transport_block``: the TS 38.212 receive chain (de-rate-matching, decode, CRC checks) for
    one 999,576-bit transport block (119 code blocks);
  * ``stream_overlap``: the same traffic as batches of 16 / 256 / 2048 codewords over 1, 2 and 4
    streams round-robin (host planning and PCIe copies overlapped with decoding);
  * ``cpu_baseline``: the C oracle (same schedule) on this host's cores.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--skip-mixed] [--skip-stream]

Prints one JSON line.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BATCH = 4096
Z = 384
N_ITER = 10
EBN0_DB = 1.5          # point of the Eb/N0 0..3 dB sweep used for the FER column
RATE = 1.0 / 3.0       # nominal rate of the BG1 Z=384 code (K=8448, N=25344)
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E
# Non-packed 32-bit VALU lane-ops/s: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz.
PEAK_VALU_GOPS = 256 * 64 * 2.4
VALU_OPS_PER_EDGE = 8  # convention: sub, min, med3, sign, select, sign-apply, dt-sub, APP-add
# counter summaries of the kernel that decodes the headline batch (the slim layout)
# PMC summaries of the headline kernel, per LDS layout of the launch (slim: the per-batch choice at
# 4096 codewords; dup: the layout of batches of at most one codeword per CU)
PROFILE_STEM = {"slim": "k_decode_slim_bg1_z384_t10_b4096", "dup": "k_decode_bg1_z384_t10_b4096"}


def profile_file(layout, kind):
    return os.path.join(ROOT, "profiles", f"{PROFILE_STEM[layout]}_{kind}.json")
# A CU issues at most one 64-lane VALU instruction per clock (4 SIMDs x 16 lanes).
PEAK_VALU_ISSUE = 256 * 2.4e9
STREAM_CLASSES = [1 << k for k in range(14)]  # combined-stream batch-size classes, 1 .. 8192
TUNING_FILE = os.path.join(ROOT, "nr_gpu_ldpc_amd", "tuning", "mixed_gfx950.txt")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--cpu-sample", type=int, default=4096, help="codewords in the CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--skip-mixed", action="store_true")
    p.add_argument("--skip-stream", action="store_true")
    p.add_argument("--layout", choices=("auto", "dup", "slim"), default="auto",
                   help="LDS layout of the headline launch (auto: the per-batch choice of the tuned model)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a 256 MiB device-to-device copy (scale factor for PMC byte counters)")
    return p.parse_args()


def pct(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, max(0, int(round(q * len(v))) - 1))]


def copy_calibration():
    import torch
    n = 64 * 1024 * 1024  # floats = 256 MiB
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n,
                      "launches": 3}))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(llr_host, n_cw):
    """The oracle (the reference's min-sum restated in C) on this host's cores: the whole
    sample with one thread per core of the box's share, and a slice with one thread."""
    from tests import oracle_binding
    orc = oracle_binding.load()
    code = orc.code(oracle_binding.OR_BG1, Z)
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    orc.decode(code, llr_host[:n_cw], N_ITER, threads=threads)
    dt = time.perf_counter() - t0
    n1 = max(1, n_cw // 16)
    t0 = time.perf_counter()
    orc.decode(code, llr_host[:n1], N_ITER, threads=1)
    dt1 = time.perf_counter() - t0
    return {"value": round(n_cw * 22 * Z / dt / 1e9, 6), "unit": "Gbps", "cores": threads, "kind": "port",
            "sample": f"{n_cw} codewords of the same BG1 Z=384 T=10 workload, decoded by the C oracle "
                      f"with {threads} threads ({dt:.2f} s wall, about {dt * threads:.0f} CPU-s)",
            "single_thread_gbps": round(n1 * 22 * Z / dt1 / 1e9, 6),
            "host": {"cpu": cpu_model(), "nproc": os.cpu_count()}}


def traffic_per_launch(layout):
    """HBM bytes per decode launch measured by rocprofv3 PMC counters (FETCH_SIZE + WRITE_SIZE,
    each scaled by the counter's reading for a copy of known size), if profiled."""
    try:
        with open(profile_file(layout, "traffic")) as f:
            t = json.load(f)
        return t["bytes_per_launch"], t
    except (OSError, ValueError, KeyError):
        return None, None


def bench_headline(torch, nr, args):
    code = nr.Code(nr.BG1, Z)
    if args.layout != "auto":
        code.set_launch_layout(nr.LAYOUT_DUP if args.layout == "dup" else nr.LAYOUT_SLIM, 0)
    layout = "slim" if code.launch_layout(BATCH)[0] == nr.LAYOUT_SLIM else "dup"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    llr = torch.empty((BATCH, code.n_vn), dtype=torch.float32, device=dev)
    bits = torch.empty((BATCH, code.n_vn), dtype=torch.uint8, device=dev)
    hd = torch.zeros((BATCH, code.hd_bytes), dtype=torch.uint8, device=dev)
    sigma = nr.noise_stddev(EBN0_DB, RATE)
    code.channel(2024, 0, BATCH, llr.data_ptr(), sigma=sigma, cw_bits_ptr=bits.data_ptr(), stream=s)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        code.decode(llr.data_ptr(), BATCH, N_ITER, hd.data_ptr(), stream=s)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    for k in range(args.steps):
        code.decode(llr.data_ptr(), BATCH, N_ITER, hd.data_ptr(), stream=s)
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [ev[k].elapsed_time(ev[k + 1]) for k in range(args.steps)]
    ms = ev[0].elapsed_time(ev[-1]) / args.steps
    gbps = BATCH * code.K * args.steps / (wall * 1e9)
    counts = torch.zeros(2, dtype=torch.int64, device=dev)
    code.count_errors(hd.data_ptr(), bits.data_ptr(), BATCH, counts.data_ptr(), stream=s)
    torch.cuda.synchronize()
    bit_err, frame_err = counts.tolist()

    edges = code.nnz * Z
    valu_ops = BATCH * edges * N_ITER * VALU_OPS_PER_EDGE
    alg_bytes = BATCH * (4 * code.n_vn + code.hd_bytes)
    valu_gops = valu_ops / (ms * 1e6)
    hbm_gbs = alg_bytes / (ms * 1e6)
    traffic, tsrc = traffic_per_launch(layout)
    try:
        with open(profile_file(layout, "sq")) as f:
            issue = json.load(f)
        n_valu = issue["valu_insts_per_launch"]
        issued = {"valu_insts_per_launch": n_valu, "lane_ops_per_edge_iteration": round(n_valu * 64 / (BATCH * edges * N_ITER), 2),
                  "rate_per_s": round(n_valu / (ms * 1e-3), 1), "peak_per_s": PEAK_VALU_ISSUE,
                  "frac": round(n_valu / (ms * 1e-3) / PEAK_VALU_ISSUE, 4), "source": issue["source"]}
    except (OSError, ValueError, KeyError):
        issued = None
    # the reference's HALF_PRC schedule on the same batch (binary16 LLRs, 2 bytes per LLR)
    llr_h = torch.empty((BATCH, code.n_vn), dtype=torch.int16, device=dev)
    nr.llr_to_half(llr.data_ptr(), llr.numel(), llr_h.data_ptr(), stream=s)
    code.decode_half(llr_h.data_ptr(), BATCH, N_ITER, hd.data_ptr(), stream=s)
    torch.cuda.synchronize()
    hev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    hev[0].record(stream)
    for _ in range(args.steps):
        code.decode_half(llr_h.data_ptr(), BATCH, N_ITER, hd.data_ptr(), stream=s)
    hev[1].record(stream)
    torch.cuda.synchronize()
    half_ms = hev[0].elapsed_time(hev[1]) / args.steps
    # the reference's CHAR_PRC schedule (its default build) on the same batch: LLRs quantised as
    # quantize(0.5, x, 8) does (1 byte per LLR)
    llr_q = torch.clamp(torch.floor(llr * 2.0), -128, 127).to(torch.int8)
    code.decode_char(llr_q.data_ptr(), BATCH, N_ITER, hd.data_ptr(), stream=s)
    torch.cuda.synchronize()
    cev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    cev[0].record(stream)
    for _ in range(args.steps):
        code.decode_char(llr_q.data_ptr(), BATCH, N_ITER, hd.data_ptr(), stream=s)
    cev[1].record(stream)
    torch.cuda.synchronize()
    char_ms = cev[0].elapsed_time(cev[1]) / args.steps
    code.count_errors(hd.data_ptr(), bits.data_ptr(), BATCH, counts.data_ptr(), stream=s)
    torch.cuda.synchronize()
    char_fe = int(counts.tolist()[1]) - frame_err
    res = {
        "value": round(gbps, 4), "ms_per_step": round(ms, 4),
        "launch_shape": dict(zip(("cw_per_block", "workgroups"), code.launch_shape(BATCH)), layout=layout),
        "latency_ms": {"p50": round(statistics.median(per), 4), "p99": round(pct(per, 0.99), 4),
                       "kind": "device time per batch of 4096"},
        "fer": round(frame_err / BATCH, 6), "ber": bit_err / (BATCH * code.K),
        "half_prc": {"info_gbps": round(BATCH * code.K / (half_ms * 1e6), 4), "ms_per_step": round(half_ms, 4),
                     "note": "the reference's HALF_PRC schedule: binary16 LLRs, R, dt and inter-iteration APP"},
        "char_prc": {"info_gbps": round(BATCH * code.K / (char_ms * 1e6), 4), "ms_per_step": round(char_ms, 4),
                     "fer": round(char_fe / BATCH, 6),
                     "note": "the reference's CHAR_PRC schedule (its default build): signed-char LLRs "
                             "(quantize(0.5, x, 8)), R, dt and inter-iteration APP"},
        "roofline": {
            "kernel": (tsrc or {}).get("kernel"),
            "bound": "hbm",
            "achieved": round(hbm_gbs, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(hbm_gbs / PEAK_HBM_GBS, 5),
            "traffic": traffic,
            "per_launch": {"algorithmic_bytes": alg_bytes, "bytes_per_codeword": 4 * code.n_vn + code.hd_bytes,
                           "codewords": BATCH, "edge_iterations": BATCH * edges * N_ITER},
            "traffic_source": (tsrc or {}).get("source"),
            "traffic_note": (tsrc or {}).get("note"),
            "valu": {"note": "the binding limit: min-sum is compare/select/add work on the VALU",
                     "achieved": round(valu_gops, 1), "peak": round(PEAK_VALU_GOPS, 1), "unit": "Gop/s",
                     "frac": round(valu_gops / PEAK_VALU_GOPS, 4), "ops_per_edge": VALU_OPS_PER_EDGE,
                     "issued": issued},
        },
    }
    return res, llr


def bench_mixed(torch, nr, args):
    """16384 codewords: BG1/BG2 50/50, Z uniform over the 51 sizes, rate U[1/5, 8/9] by
    parity-row truncation (clamped per base graph), Eb/N0 U[0, 4] dB, T <= 20, early exit."""
    from nr_gpu_ldpc_amd import mixed as mx
    n, T = 16384, 20
    b = mx.draw_batch(n, seed=2025)
    llr, bits = mx.make_channel(torch, b, seed=77)
    m = nr.Mixed(n)
    tuned = os.path.exists(TUNING_FILE)
    if tuned:
        m.load_tuning(TUNING_FILE)
    hd = torch.zeros(b["hd_total"], dtype=torch.uint8, device="cuda")
    its = torch.zeros(n, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    # LDS layout of the tiles: timed on this batch with its flags, the faster kept
    lay_best, lay_us = m.autotune_layout(b["desc"], llr.data_ptr(), T, hd.data_ptr(), flags=nr.EARLY_EXIT)
    lay_name = {nr.LAYOUT_DUP: "dup", nr.LAYOUT_SLIM: "slim"}
    m.plan(b["desc"])
    for _ in range(2):
        m.decode(llr.data_ptr(), T, hd.data_ptr(), 0, its.data_ptr(), flags=nr.EARLY_EXIT, stream=s)
    torch.cuda.synchronize()
    steps = max(5, args.steps)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * steps)]
    for k in range(steps):
        ev[2 * k].record(stream)
        m.decode(llr.data_ptr(), T, hd.data_ptr(), 0, its.data_ptr(), flags=nr.EARLY_EXIT, stream=s)
        ev[2 * k + 1].record(stream)
    torch.cuda.synchronize()
    per = [ev[2 * k].elapsed_time(ev[2 * k + 1]) for k in range(steps)]
    # the same batch through a captured graph (plan upload + decode), replayed
    gs = torch.cuda.Stream()
    m.capture(gs.cuda_stream, llr.data_ptr(), T, hd.data_ptr(), 0, its.data_ptr(), flags=nr.EARLY_EXIT)
    gev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * steps)]
    for k in range(steps):
        gev[2 * k].record(gs)
        m.replay(gs.cuda_stream)
        gev[2 * k + 1].record(gs)
    torch.cuda.synchronize()
    gper = [gev[2 * k].elapsed_time(gev[2 * k + 1]) for k in range(steps)]
    be, fe, _ = mx.info_errors(torch, b, hd, bits)
    info = int(b["info_bits"].sum())
    tiles, _ = m.tiles()
    cpb = {}
    for kind, z, _, c, _, cnt in tiles.tolist():
        key = f"{kind}/{z}"
        cpb[key] = max(cpb.get(key, 0), c)
    info_ = m.plan_info()
    p50 = statistics.median(per)
    return {
        "workload": "16384 codewords in one persistent launch: BG1/BG2 50/50, Z uniform over the 51 "
                    "TS 38.212 sizes, rate U[1/5, 8/9] by parity-row truncation (clamped per base graph), "
                    "Eb/N0 U[0, 4] dB, at most 20 iterations with per-codeword early exit",
        "info_gbps": round(info / (p50 * 1e6), 3),
        "latency_ms": {"p50": round(p50, 3), "p99": round(pct(per, 0.99), 3), "kind": "device time per batch"},
        "graph_latency_ms": {"p50": round(statistics.median(gper), 3), "p99": round(pct(gper, 0.99), 3),
                             "kind": "replayed graph (plan upload + decode), device time"},
        "mean_iterations": round(float(its.float().mean()), 3),
        "fer": round(fe / n, 6), "ber": be / info,
        "tiling": {"source": os.path.relpath(TUNING_FILE, ROOT) if tuned else "default (most codewords that fit)",
                   "layout": lay_name[lay_best],
                   "layout_autotune_ms": {lay_name[k]: round(v / 1000, 3) for k, v in lay_us.items()},
                   "threads": info_["threads"], "grid": info_["grid"], "lds_bytes": info_["lds_bytes"],
                   "tiles": info_["n_tiles"], "cpb_examples": {k: cpb[k] for k in sorted(cpb)[:6]},
                   "cpb_bg1_384": cpb.get("1/384"), "cpb_bg2_208": cpb.get("2/208")},
        "seed": {"batch": 2025, "channel": 77},
    }


def stream_weights(nr, T=10):
    """The committed learned weights of the stream's two codes (BG1 Z=384, BG2 Z=208)."""
    from nr_gpu_ldpc_amd import noms
    weights = {}
    for kind, z in ((nr.BG1, 384), (nr.BG2, 208)):
        path = noms.weights_path(kind, z, T)
        if os.path.exists(path):
            weights[(kind, z)] = noms.load_weights(path)["weights"]
    return weights


def stream_layouts(torch, nr, classes):
    """The mixed objects' LDS layout for the stream's traffic, per decode mode and batch-size
    class: the layout autotuner timed on a representative batch of the class size (BG1 Z=384 /
    BG2 Z=208 at mixed rates, Eb/N0 1.5 dB, T=10) with the mode's own flags.  Small batches
    favour the dup layout (a tile's latency), large ones the slim layout (throughput)."""
    from nr_gpu_ldpc_amd import mixed as mx
    T = 10
    rng = np.random.default_rng(123)
    names = {nr.LAYOUT_DUP: "dup", nr.LAYOUT_SLIM: "slim"}
    weights = stream_weights(nr, T)
    modes = (("plain", nr.EARLY_EXIT), ("learned", nr.EARLY_EXIT | nr.LEARNED | nr.FALLBACK_PLAIN))
    out = {mode: {} for mode, _ in modes}
    contexts = {}
    for c in classes:
        kind = np.where(rng.random(c) < 0.5, nr.BG1, nr.BG2)
        r_nom = rng.uniform(0.2, 8 / 9, c)
        mr = np.where(kind == nr.BG1, mx.rows_for_rate(nr.BG1, r_nom), mx.rows_for_rate(nr.BG2, r_nom))
        b = mx.describe(kind, np.where(kind == nr.BG1, 384, 208), mr, ebn0=np.full(c, 1.5))
        llr, _ = mx.make_channel(torch, b, seed=123 + c, contexts=contexts)
        hd = torch.zeros(b["hd_total"], dtype=torch.uint8, device="cuda")
        for mode, flags in modes:
            m = nr.Mixed(c)
            if mode == "learned":
                for (k, z), w in weights.items():
                    m.set_weights(k, z, w)
            best, us = m.autotune_layout(b["desc"], llr.data_ptr(), T, hd.data_ptr(), flags=flags)
            out[mode][c] = {"layout": best, "name": names[best],
                            "ms": {names[k]: round(v / 1000, 3) for k, v in us.items()}}
            m.close()
    return out


def bench_stream(torch, nr, args, layouts):
    """BG1 Z=384 / BG2 Z=208 at mixed rates, batches of 1..8192 codewords, each batch one
    replayed graph: LLR upload, plan upload, decode, hard-decision download.  The same sequence
    with the LLRs sent as float32, as int8 (quantize(0.5, x, 8)) and as binary16."""
    from nr_gpu_ldpc_amd import mixed as mx
    T = 10
    rng = np.random.default_rng(99)
    weights = stream_weights(nr, T)
    classes = STREAM_CLASSES
    max_f = 8192 * 68 * 384
    max_h = 8192 * ((68 * 384 + 7) // 8)
    # element type -> (LDPC_LLR_* code, scale, torch dtype, bytes per LLR)
    fmts = {"f32": (nr.LLR_F32, 1.0, torch.float32, 4), "i8": (nr.LLR_I8, 0.5, torch.int8, 1),
            "f16": (nr.LLR_F16, 1.0, torch.float16, 2)}
    d_llr = torch.zeros(max_f, dtype=torch.float32, device="cuda")
    d_hd = torch.zeros(max_h, dtype=torch.uint8, device="cuda")
    h_llr = {f: torch.zeros(max_f, dtype=v[2]).pin_memory() for f, v in fmts.items()}
    h_hd = torch.zeros(max_h, dtype=torch.uint8).pin_memory()
    stream = torch.cuda.Stream()
    s = stream.cuda_stream
    modes = {"plain": nr.EARLY_EXIT, "learned": nr.EARLY_EXIT | nr.LEARNED | nr.FALLBACK_PLAIN}
    objs = {}
    for f, (code, scale, _, nbytes) in fmts.items():
        for mode, flags in modes.items():
            for c in classes:
                m = nr.Mixed(c)
                m.set_layout(layouts[mode][c]["layout"])
                m.set_input(code, scale)
                if mode == "learned":
                    for (kind, z), w in weights.items():
                        m.set_weights(kind, z, w)
                m.plan(mx.describe([nr.BG1] * c, [384] * c, [46] * c)["desc"])
                m.capture(s, d_llr.data_ptr(), T, d_hd.data_ptr(), 0, 0, flags=flags,
                          llr_host_ptr=h_llr[f].data_ptr(), llr_floats=c * 68 * 384, hd_host_ptr=h_hd.data_ptr(),
                          hd_bytes=c * ((68 * 384 + 7) // 8))
                objs[(f, mode, c)] = m
    ebn0_points = (1.0, 1.5, 2.0)
    batches = [(c, int(rng.integers(c // 2 + 1, c + 1)), ebn0_points[i % 3]) for i, c in enumerate(classes * 3)]
    batches += [(8192, 8192, e) for e in ebn0_points for _ in range(2)]  # FER statistics
    stats = {f: {mode: {"lat": [], "info": 0, "bytes": 0, "time_ms": 0.0, "fer": {e: [0, 0] for e in ebn0_points}}
                 for mode in modes} for f in fmts}
    # the bus these batches cross: best of three pinned host-to-device copies of 256 MiB
    n_cal = 64 * 1024 * 1024
    cal = []
    with torch.cuda.stream(stream):
        for _ in range(3):
            c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            c0.record(stream)
            d_llr[:n_cal].copy_(h_llr["f32"][:n_cal], non_blocking=True)
            c1.record(stream)
            stream.synchronize()
            cal.append(c0.elapsed_time(c1))
    h2d_gbs = 4 * n_cal / (min(cal) * 1e6)
    contexts = {}
    drawn = []
    for bi, (c, n, e) in enumerate(batches):
        kind = np.where(rng.random(n) < 0.5, nr.BG1, nr.BG2)
        Zs = np.where(kind == nr.BG1, 384, 208)
        r_nom = rng.uniform(0.2, 8 / 9, n)
        mr = np.where(kind == nr.BG1, mx.rows_for_rate(nr.BG1, r_nom), mx.rows_for_rate(nr.BG2, r_nom))
        drawn.append(mx.describe(kind, Zs, mr, ebn0=np.full(n, e)))
    # every batch's LLRs in one pinned buffer per element type, and room for every batch's
    # decisions, for the pipelined pass below
    l_off = np.concatenate([[0], np.cumsum([b["llr_floats"] for b in drawn])])
    h_off = np.concatenate([[0], np.cumsum([b["hd_total"] for b in drawn])])
    all_llr = {f: torch.empty(int(l_off[-1]), dtype=v[2]).pin_memory() for f, v in fmts.items()}
    all_hd = {(f, mode): torch.zeros(int(h_off[-1]), dtype=torch.uint8).pin_memory() for f in fmts for mode in modes}
    want_hd = {(f, mode): torch.zeros(int(h_off[-1]), dtype=torch.uint8) for f in fmts for mode in modes}
    for bi, (c, n, e) in enumerate(batches):
        b = drawn[bi]
        llr, bits = mx.make_channel(torch, b, seed=1000 + bi, contexts=contexts)
        x = llr.cpu().numpy()
        host = {"f32": torch.from_numpy(x), "i8": torch.from_numpy(mx.quantize(x)),
                "f16": torch.from_numpy(x.astype(np.float16))}
        for f in fmts:
            h_llr[f][:b["llr_floats"]].copy_(host[f])
            all_llr[f][int(l_off[bi]):int(l_off[bi + 1])].copy_(host[f])
            for mode in modes:
                m = objs[(f, mode, c)]
                m.plan(b["desc"])
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                m.replay(s)
                e1.record(stream)
                stream.synchronize()
                ms = e0.elapsed_time(e1)
                want_hd[(f, mode)][int(h_off[bi]):int(h_off[bi + 1])].copy_(h_hd[:b["hd_total"]])
                hd = h_hd[:b["hd_total"]].cuda()
                _, fe, _ = mx.info_errors(torch, b, hd, bits)
                st = stats[f][mode]
                st["lat"].append(ms)
                st["info"] += int(b["info_bits"].sum())
                st["bytes"] += fmts[f][3] * b["llr_floats"] + b["hd_total"]
                st["time_ms"] += ms
                st["fer"][e][0] += fe
                st["fer"][e][1] += n
    out = {"workload": "BG1 Z=384 and BG2 Z=208 (50/50) at rates U[1/5, 8/9] by parity-row truncation, "
                       "batches of 1..8192 codewords (sizes in (2^k/2, 2^k], 3 per power of two, plus 2 x 8192 "
                       "per Eb/N0 point), Eb/N0 1.0 / 1.5 / 2.0 dB, T=10 with early exit; each batch is one "
                       "replayed HIP graph: LLR upload, plan upload, persistent decode, hard-decision download; "
                       "top-level entries send float32 LLRs, llr_int8 / llr_half the same LLRs as "
                       "quantize(0.5, x, 8) / binary16 (converted to float32 on load, decoded in float32)",
           "learned_weights": sorted(f"{k}/{z}" for k, z in weights),
           "layout": {mode: {str(c): v["name"] for c, v in per.items()} for mode, per in layouts.items()},
           "layout_autotune_ms": {mode: {str(c): v["ms"] for c, v in per.items()} for mode, per in layouts.items()},
           "pcie": {"h2d_gbs": round(h2d_gbs, 2),
                    "note": "best pinned host-to-device copy of 256 MiB; every batch moves its LLRs in (4, 1 or 2 "
                            "B each) and its packed decisions out, so info Gbps <= info bits / (bytes / h2d_gbs)"}}
    # The same sequence pipelined over two streams: batch i + 1 is planned, uploaded and
    # decoded while batch i is decoded or downloaded.  Each stream has its own mixed object
    # (graph: plan upload + decode) and device buffers; uploads and downloads are stream copies
    # from and to the pinned buffers holding every batch.
    n_pipe = 2
    pstreams = hip_streams(torch, n_pipe)
    for f, (code, scale, dtype, _) in fmts.items():
        for mode, flags in modes.items():
            # one object per stream serves every batch size: its layout is the faster of the two
            # on this whole pipelined sequence (each run's decisions are checked)
            runs = {}
            for lay, lay_name in ((nr.LAYOUT_DUP, "dup"), (nr.LAYOUT_SLIM, "slim")):
                slots = []
                for k in range(n_pipe):
                    m = nr.Mixed(8192)
                    m.set_layout(lay)
                    m.set_input(code, scale)
                    if mode == "learned":
                        for (kind, z), w in weights.items():
                            m.set_weights(kind, z, w)
                    dl = torch.zeros(max_f, dtype=dtype, device="cuda")
                    dh = torch.zeros(max_h, dtype=torch.uint8, device="cuda")
                    m.plan(drawn[0]["desc"])
                    m.capture(pstreams[k].cuda_stream, dl.data_ptr(), T, dh.data_ptr(), 0, 0, flags=flags,
                              llr_floats=max_f, hd_bytes=max_h)
                    slots.append((m, dl, dh))
                torch.cuda.synchronize()
                all_hd[(f, mode)].zero_()
                t0 = time.perf_counter()
                for bi, b in enumerate(drawn):
                    m, dl, dh = slots[bi % n_pipe]
                    ps = pstreams[bi % n_pipe]
                    m.plan(b["desc"])  # waits until this stream's previous batch has left its buffers
                    with torch.cuda.stream(ps):
                        dl[:b["llr_floats"]].copy_(all_llr[f][int(l_off[bi]):int(l_off[bi + 1])], non_blocking=True)
                    m.replay(ps.cuda_stream)
                    with torch.cuda.stream(ps):
                        all_hd[(f, mode)][int(h_off[bi]):int(h_off[bi + 1])].copy_(dh[:b["hd_total"]],
                                                                                   non_blocking=True)
                torch.cuda.synchronize()
                runs[lay_name] = time.perf_counter() - t0
                if not torch.equal(all_hd[(f, mode)], want_hd[(f, mode)]):
                    raise RuntimeError(f"pipelined {f} {mode} {lay_name} stream decisions differ from the serial pass")
                for m, _, _ in slots:
                    m.close()
            best = min(runs, key=runs.get)
            wall = runs[best]
            stats[f][mode]["pipelined"] = {"streams": n_pipe, "layout": best,
                                           "info_gbps": round(stats[f][mode]["info"] / (wall * 1e9), 3),
                                           "wall_ms": round(wall * 1e3, 2),
                                           "wall_ms_by_layout": {k: round(v * 1e3, 2) for k, v in runs.items()}}
    for f in fmts:
        dst = out if f == "f32" else out.setdefault({"i8": "llr_int8", "f16": "llr_half"}[f], {})
        for mode, st in stats[f].items():
            dst[mode] = {"info_gbps": round(st["info"] / (st["time_ms"] * 1e6), 3),
                         "pipelined": st["pipelined"],
                         "bytes_moved": st["bytes"],
                         "copy_bound_info_gbps": round(st["info"] * h2d_gbs / st["bytes"], 3),
                         "latency_ms": {"p50": round(statistics.median(st["lat"]), 3),
                                        "p99": round(pct(st["lat"], 0.99), 3),
                                        "kind": "device time per batch incl. host copies"},
                         "fer": {f"{e}dB": round(v[0] / max(1, v[1]), 6) for e, v in st["fer"].items()},
                         "frames": {f"{e}dB": v[1] for e, v in st["fer"].items()}}
    return out


def hip_streams(torch, n):
    """n fresh HIP streams created back to back (the runtime hands out its hardware queues
    round-robin, so consecutive streams do not share one), wrapped for torch.  They live until
    the process ends: torch's pinned-memory allocator records events on every stream a pinned
    buffer was copied on, also when the buffer is freed later."""
    import ctypes
    import importlib.util
    spec = importlib.util.find_spec("torch")
    hip = ctypes.CDLL(os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so"))
    out = []
    for _ in range(n):
        h = ctypes.c_void_p()
        if hip.hipStreamCreate(ctypes.byref(h)) != 0:
            raise RuntimeError("hipStreamCreate failed")
        out.append(torch.cuda.ExternalStream(h.value))
    return out


def bench_pipeline(torch, nr, args, layouts):
    """Copy/compute overlap over S = 1, 2, 4 streams: batches of c codewords (BG1 Z=384 / BG2
    Z=208 at mixed rates, Eb/N0 1.5 dB) submitted round-robin, each one replayed graph (LLR
    upload from pinned host memory, plan upload, persistent decode with early exit, hard-decision
    download) on its stream's own buffers and mixed object; the host plans batch i + S while
    the GPU works on batch i.  Every slot's decisions are checked against a decode of the same
    batch with nothing else in flight."""
    from nr_gpu_ldpc_amd import mixed as mx
    T = 10
    S_MAX = 4
    rng = np.random.default_rng(7)
    streams = hip_streams(torch, S_MAX)
    contexts = {}
    out = {"workload": "BG1 Z=384 / BG2 Z=208 (50/50) at rates U[1/5, 8/9], Eb/N0 1.5 dB, T=10 with "
                       "early exit; per batch: plan on the host, then one replayed HIP graph (pinned "
                       "LLR upload, plan upload, decode, hard-decision download); S streams round-robin, each "
                       "stream's persistent grid the whole device or (shared_grid) 1/S of its resident workgroups; "
                       "each configuration timed in both LDS layouts, the faster reported",
           "latency_kind": "device time per batch on its stream (start to end event), under S-way overlap"}
    for c in (16, 256, 2048):
        slots = []
        for k in range(S_MAX):
            kind = np.where(rng.random(c) < 0.5, nr.BG1, nr.BG2)
            Zs = np.where(kind == nr.BG1, 384, 208)
            r_nom = rng.uniform(0.2, 8 / 9, c)
            mr = np.where(kind == nr.BG1, mx.rows_for_rate(nr.BG1, r_nom), mx.rows_for_rate(nr.BG2, r_nom))
            b = mx.describe(kind, Zs, mr, ebn0=np.full(c, 1.5))
            llr, _ = mx.make_channel(torch, b, seed=50000 + 10 * c + k, contexts=contexts)
            h_llr = torch.empty(b["llr_floats"], dtype=torch.float32).pin_memory()
            h_llr.copy_(llr.cpu())
            h_hd = torch.zeros(b["hd_total"], dtype=torch.uint8).pin_memory()
            d_llr = torch.empty_like(llr)
            d_hd = torch.zeros(b["hd_total"], dtype=torch.uint8, device="cuda")
            m = nr.Mixed(c)
            m.set_layout(layouts["plain"][c]["layout"])
            m.plan(b["desc"])
            st = streams[k].cuda_stream
            m.capture(st, d_llr.data_ptr(), T, d_hd.data_ptr(), 0, 0, flags=nr.EARLY_EXIT,
                      llr_host_ptr=h_llr.data_ptr(), llr_floats=b["llr_floats"], hd_host_ptr=h_hd.data_ptr(),
                      hd_bytes=b["hd_total"])
            m.replay(st)
            streams[k].synchronize()
            slots.append({"m": m, "b": b, "h_llr": h_llr, "h_hd": h_hd, "d_llr": d_llr, "d_hd": d_hd,
                          "want": h_hd.clone(), "info": int(b["info_bits"].sum())})
        row = {}
        full_grid = slots[0]["m"].plan_info()["grid"]
        # every stream's persistent grid the whole device, or (shared) 1/S of it; each
        # configuration is timed in both LDS layouts and the faster one is reported (under S-way
        # contention the best layout is not the lone batch's)
        for S, shared in ((1, False), (2, False), (4, False), (2, True), (4, True)):
            by_layout = {}
            for lay_name, lay in (("dup", nr.LAYOUT_DUP), ("slim", nr.LAYOUT_SLIM)):
                for k in range(S):
                    sl = slots[k]
                    sl["m"].set_layout(lay)
                    sl["m"].set_grid(full_grid // S if shared else 0)
                    sl["m"].plan(sl["b"]["desc"])
                    sl["m"].capture(streams[k].cuda_stream, sl["d_llr"].data_ptr(), T, sl["d_hd"].data_ptr(), 0, 0,
                                    flags=nr.EARLY_EXIT, llr_host_ptr=sl["h_llr"].data_ptr(),
                                    llr_floats=sl["b"]["llr_floats"], hd_host_ptr=sl["h_hd"].data_ptr(),
                                    hd_bytes=sl["b"]["hd_total"])
                    sl["h_hd"].zero_()
                reps = max(8 * S, min(256, (1 << 17) // c))
                for k in range(S):  # warm-up, one per stream
                    slots[k]["m"].plan(slots[k]["b"]["desc"])
                    slots[k]["m"].replay(streams[k].cuda_stream)
                torch.cuda.synchronize()
                ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
                info = 0
                t0 = time.perf_counter()
                for i in range(reps):
                    sl, st = slots[i % S], streams[i % S]
                    sl["m"].plan(sl["b"]["desc"])
                    ev[i][0].record(st)
                    sl["m"].replay(st.cuda_stream)
                    ev[i][1].record(st)
                    info += sl["info"]
                torch.cuda.synchronize()
                wall = time.perf_counter() - t0
                for k in range(S):
                    if not torch.equal(slots[k]["h_hd"], slots[k]["want"]):
                        raise RuntimeError(f"pipelined decode differs from the lone decode (c={c}, S={S}, "
                                           f"{lay_name}, slot {k})")
                lat = [a.elapsed_time(b_) for a, b_ in ev]
                by_layout[lay_name] = {
                    "batches": reps, "grid_per_stream": full_grid // S if shared else full_grid,
                    "info_gbps": round(info / wall / 1e9, 3),
                    "batches_per_s": round(reps / wall, 1),
                    "latency_ms": {"p50": round(statistics.median(lat), 3), "p99": round(pct(lat, 0.99), 3)}}
            best = max(by_layout, key=lambda n: by_layout[n]["info_gbps"])
            row[f"streams_{S}" + ("_shared_grid" if shared else "")] = dict(
                by_layout[best], layout=best,
                info_gbps_by_layout={n: v["info_gbps"] for n, v in by_layout.items()})
        out[f"batch_{c}"] = row
        for sl in slots:
            sl["m"].close()
    return out


def bench_transport_block(torch, nr, args):
    """The TS 38.212 receive chain around the decoder for one large transport block: de-rate-
    matching of G channel LLRs into C code blocks, the fused decode (T=10, early exit), the
    code-block CRC checks and the transport-block assembly, all on the device."""
    A, rate, G, Qm = 999_576, 1.0 / 3.0, 3_000_000, 4  # a TS 38.214 size: C = 119, K' = 8424
    t = nr.TransportBlock(A, rate, G, Qm=Qm)
    rng = np.random.default_rng(5)
    a = rng.integers(0, 2, A, dtype=np.uint8)
    f, _ = t.encode_host(a)
    sigma = np.float32(0.7)
    y = (1.0 - 2.0 * f.astype(np.float32)) + sigma * rng.standard_normal(G).astype(np.float32)
    llr = torch.from_numpy(((y * np.float32(2)) / (sigma * sigma)).astype(np.float32)).cuda()
    code = nr.Code(t.bg, t.Z)
    d_cb = torch.empty((t.C, t.n_vn), dtype=torch.float32, device="cuda")
    d_hd = torch.empty((t.C, code.hd_bytes), dtype=torch.uint8, device="cuda")
    ok = torch.zeros(t.C, dtype=torch.int32, device="cuda")
    bits = torch.zeros(t.B, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream

    def chain():
        t.derate(llr.data_ptr(), d_cb.data_ptr(), stream=s)
        code.decode(d_cb.data_ptr(), t.C, N_ITER, d_hd.data_ptr(), flags=nr.EARLY_EXIT, stream=s)
        t.check(d_hd.data_ptr(), ok.data_ptr(), bits.data_ptr(), stream=s)

    chain()
    torch.cuda.synchronize()
    reps = max(5, args.steps)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ev[0].record(stream)
    for k in range(reps):
        chain()
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    per = [ev[k].elapsed_time(ev[k + 1]) for k in range(reps)]
    passed = int(ok.sum().item())
    ok_tb = passed == t.C and bool(torch.equal(bits[:A].cpu(), torch.from_numpy(a)))
    return {"workload": f"one transport block: A={A} bits, rate 1/3, G={G} bits at Qm={Qm}, BPSK-equivalent "
                        f"LLRs at sigma {float(sigma)}; C={t.C} BG{t.bg} Z={t.Z} code blocks, T<=10 with early exit",
            "info_gbps": round(A / (statistics.median(per) * 1e6), 3),
            "latency_ms": {"p50": round(statistics.median(per), 3), "p99": round(pct(per, 0.99), 3),
                           "kind": "device time: de-rate-matching + decode + CRC checks + assembly"},
            "code_blocks": t.C, "code_blocks_passed": passed, "payload_recovered": ok_tb}


def bench_small_batches(torch, nr, args):
    """Device time of single-code decodes of small batches (BG1 Z=384 and BG2 Z=208, T=10 fixed,
    Eb/N0 1.5 dB): a batch below one codeword per CU still takes one round of the fused kernel,
    which bounds the latency of tiny batches."""
    out = {"workload": "single-code decodes of 1, 16 and 256 codewords, T=10 fixed iterations, LLRs "
                       "resident in HBM; device time per launch (event pair on the decode stream)",
           "kind": "device time per batch"}
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    for kind, z, rate in ((nr.BG1, 384, 1 / 3), (nr.BG2, 208, 1 / 5)):
        code = nr.Code(kind, z)
        row = {}
        for n in (1, 16, 256):
            llr = torch.empty((n, code.n_vn), dtype=torch.float32, device="cuda")
            hd = torch.zeros((n, code.hd_bytes), dtype=torch.uint8, device="cuda")
            code.channel(77, 0, n, llr.data_ptr(), sigma=nr.noise_stddev(1.5, rate), stream=s)
            for _ in range(3):
                code.decode(llr.data_ptr(), n, N_ITER, hd.data_ptr(), stream=s)
            reps = 50
            ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
            for a, b_ in ev:
                a.record(stream)
                code.decode(llr.data_ptr(), n, N_ITER, hd.data_ptr(), stream=s)
                b_.record(stream)
            torch.cuda.synchronize()
            lat = [a.elapsed_time(b_) for a, b_ in ev]
            lay, cpb, wg = code.launch_layout(n)
            row[str(n)] = {"latency_ms": {"p50": round(statistics.median(lat), 4), "p99": round(pct(lat, 0.99), 4)},
                           "info_gbps": round(n * code.K / (statistics.median(lat) * 1e6), 4),
                           "layout": "slim" if lay == nr.LAYOUT_SLIM else "dup", "cw_per_block": cpb,
                           "workgroups": wg}
        out[f"bg{kind}_z{z}"] = row
        code.close()
    return out


def bench_operating_point(torch, nr, args):
    """Info throughput at the FER 1e-3 operating point of each decoder (BG1 Z=384, T <= 10 with
    early exit, batches of 4096): plain min-sum at its FER 1e-3 Eb/N0, the learned decoder (with
    plain fallback) at its own, as measured for the committed weights."""
    from nr_gpu_ldpc_amd import noms
    path = noms.weights_path(nr.BG1, Z, N_ITER)
    if not os.path.exists(path):
        return None
    meta = json.load(open(path.replace(".bin", ".eval.json")))["fer1e-3_db"]
    code = nr.Code(nr.BG1, Z)
    code.set_weights(noms.load_weights(path)["weights"])
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    llr = torch.empty((BATCH, code.n_vn), dtype=torch.float32, device=dev)
    bits = torch.empty((BATCH, code.n_vn), dtype=torch.uint8, device=dev)
    hd = torch.zeros((BATCH, code.hd_bytes), dtype=torch.uint8, device=dev)
    its = torch.zeros(BATCH, dtype=torch.int32, device=dev)
    cnt = torch.zeros(2, dtype=torch.int64, device=dev)
    out = {"workload": "BG1 Z=384, batch 4096, at most 10 iterations with early exit, at each decoder's "
                       "FER 1e-3 Eb/N0 (from the committed evaluation)"}
    steps = max(5, args.steps)
    wgt = torch.zeros(BATCH, dtype=torch.int32, device=dev)
    for mode, ebn0 in (("plain", meta["plain"]), ("learned", meta["learned"]),
                       ("learned_no_fallback", meta["learned"])):
        flags = nr.EARLY_EXIT | (nr.FALLBACK_PLAIN if mode == "learned" else 0)

        def run(seed):
            code.channel(seed, 0, BATCH, llr.data_ptr(), sigma=nr.noise_stddev(ebn0, RATE), cw_bits_ptr=bits.data_ptr(),
                         stream=s)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            if mode.startswith("learned"):
                code.decode_learned(llr.data_ptr(), BATCH, N_ITER, hd.data_ptr(), 0, its.data_ptr(), flags=flags, stream=s)
            else:
                code.decode(llr.data_ptr(), BATCH, N_ITER, hd.data_ptr(), 0, its.data_ptr(), flags=flags, stream=s)
            e1.record(stream)
            code.count_errors(hd.data_ptr(), bits.data_ptr(), BATCH, cnt.data_ptr(), stream=s)
            return e0, e1

        run(1)
        cnt.zero_()
        torch.cuda.synchronize()
        evs, iters, sfail = [], 0, 0
        for k in range(steps):
            evs.append(run(100 + k))
            code.syndrome(hd.data_ptr(), BATCH, wgt.data_ptr(), stream=s)
            iters += int(its.sum())
            sfail += int(torch.count_nonzero(wgt))
        torch.cuda.synchronize()
        per = [a.elapsed_time(b) for a, b in evs]
        be, fe = cnt.tolist()
        out[mode] = {"ebn0_db": ebn0, "info_gbps": round(BATCH * code.K / (statistics.mean(per) * 1e6), 3),
                     "latency_ms": {"p50": round(statistics.median(per), 3), "p99": round(pct(per, 0.99), 3)},
                     "mean_iterations": round(iters / (steps * BATCH), 3),
                     "fer_measured": fe / (steps * BATCH), "syndrome_fail": sfail / (steps * BATCH),
                     "frames": steps * BATCH}
    return out


def bench_ebn0_sweep(torch, nr, args):
    """BER/FER of the headline workload (BG1 Z=384, 10 fixed iterations, batches of 4096) over
    Eb/N0 0..3 dB in 0.25 dB steps: plain min-sum and the learned decoder with plain fallback on
    the same LLRs, errors counted over the first K decoded bits as the reference counts them."""
    from nr_gpu_ldpc_amd import noms
    path = noms.weights_path(nr.BG1, Z, N_ITER)
    code = nr.Code(nr.BG1, Z)
    learned = os.path.exists(path)
    if learned:
        code.set_weights(noms.load_weights(path)["weights"])
    dev = torch.device("cuda:0")
    s = torch.cuda.current_stream().cuda_stream
    llr = torch.empty((BATCH, code.n_vn), dtype=torch.float32, device=dev)
    bits = torch.empty((BATCH, code.n_vn), dtype=torch.uint8, device=dev)
    hd = torch.zeros((BATCH, code.hd_bytes), dtype=torch.uint8, device=dev)
    cnt = {m: torch.zeros(2, dtype=torch.int64, device=dev) for m in ("plain", "learned")}
    batches = 2
    points = []
    for i in range(13):
        ebn0 = 0.25 * i
        for m in cnt.values():
            m.zero_()
        for k in range(batches):
            code.channel(7000 + 100 * i + k, 0, BATCH, llr.data_ptr(), sigma=nr.noise_stddev(ebn0, RATE),
                         cw_bits_ptr=bits.data_ptr(), stream=s)
            code.decode(llr.data_ptr(), BATCH, N_ITER, hd.data_ptr(), stream=s)
            code.count_errors(hd.data_ptr(), bits.data_ptr(), BATCH, cnt["plain"].data_ptr(), stream=s)
            if learned:
                code.decode_learned(llr.data_ptr(), BATCH, N_ITER, hd.data_ptr(), flags=nr.FALLBACK_PLAIN, stream=s)
                code.count_errors(hd.data_ptr(), bits.data_ptr(), BATCH, cnt["learned"].data_ptr(), stream=s)
        frames = batches * BATCH
        row = {"ebn0_db": ebn0}
        for m in (("plain", "learned") if learned else ("plain",)):
            be, fe = cnt[m].tolist()
            row[m] = {"ber": be / (frames * code.K), "fer": fe / frames}
        points.append(row)
    torch.cuda.synchronize()
    return {"workload": "BG1 Z=384, 10 fixed iterations, Eb/N0 0..3 dB step 0.25, "
                        f"{batches} x {BATCH} frames per point; learned = committed weights with plain fallback",
            "frames_per_point": batches * BATCH, "points": points}


def bench_reference_example():
    """The reference's shipped example through the BER/FER driver (12 x 24 code, Z=96, T=5,
    80 codewords x 8 streams per SNR point, SNR 0..10 dB; host-generated data, PCIe included)."""
    import subprocess
    exe = os.path.join(ROOT, "nr_gpu_ldpc_amd", "bin", "ldpc_ber")
    if not os.path.exists(exe):
        return None
    args = ["--code", "qc", "--Z", "96", "--iters", "5", "--snr-min", "0", "--snr-max", "10", "--snr-step", "1",
            "--codewords", "80", "--cw-per-mcw", "8", "--mcw", "10", "--streams", "8", "--seed", "69012"]
    out = {"workload": "12 x 24 QC code, Z=96 (K=1152, N=2304), flooding min-sum T=5, 640 frames per SNR point "
                       "(80 unique x 8 streams), host rand/drand48 data as the reference generates it",
           "command": "nr_gpu_ldpc_amd/bin/ldpc_ber " + " ".join(args)}
    # FP32 (the configuration's LLRs) and the reference's default CHAR_PRC build (--prec char)
    for key, extra in (("points", []), ("char_build_points", ["--prec", "char"])):
        r = subprocess.run([exe] + args + extra, capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            return {"error": r.stderr[-500:]}
        out[key] = [{k: x[k] for k in ("snr_db", "ber", "fer", "best_latency_us", "user_throughput_mbps")}
                    for x in (json.loads(line) for line in r.stdout.splitlines())]
    return out


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the decode path would shard "
                                   "codewords across GPUs by codeword range; run with --gpus 1"}))
        sys.exit(2)
    import torch
    import nr_gpu_ldpc_amd as nr

    head, llr = bench_headline(torch, nr, args)
    out = {
        "metric": "info Gbps/GPU at FER 1e-3, learned vs plain min-sum; p99 batch latency",
        "value": head["value"],
        "unit": "Gbps",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Philox-generated info bits, systematic encoding, BPSK over AWGN at "
                f"Eb/N0 {EBN0_DB} dB, generated on the device before timing",
        "config": {"workload": "5G NR BG1 Z=384 (K=8448, N=25344), flooding scaled min-sum, "
                               "10 fixed iterations, batch 4096",
                   "bg": 1, "Z": Z, "n_iter": N_ITER, "batch": BATCH, "ebn0_db": EBN0_DB,
                   "launch_shape": head["launch_shape"]},
        "latency_ms": head["latency_ms"],
        "fer": head["fer"],
        "ber": head["ber"],
        "roofline": head["roofline"],
        "half_prc": head["half_prc"],
        "char_prc": head["char_prc"],
    }
    op = bench_operating_point(torch, nr, args)
    if op:
        out["fer1e-3_operating_point"] = op
    out["ebn0_sweep"] = bench_ebn0_sweep(torch, nr, args)
    out["small_batches"] = bench_small_batches(torch, nr, args)
    if not args.skip_mixed:
        out["mixed_batch"] = bench_mixed(torch, nr, args)
        out["transport_block"] = bench_transport_block(torch, nr, args)
    if not args.skip_stream:
        layouts = stream_layouts(torch, nr, STREAM_CLASSES)
        out["combined_stream"] = bench_stream(torch, nr, args, layouts)
        out["stream_overlap"] = bench_pipeline(torch, nr, args, layouts)
        out["stream_overlap"]["layout"] = {str(c): layouts["plain"][c]["name"] for c in (16, 256, 2048)}
    if not args.skip_stream:
        ex = bench_reference_example()
        if ex:
            out["reference_example"] = ex
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(llr.cpu().numpy(), min(args.cpu_sample, BATCH))
    print(json.dumps(out, default=lambda o: o.item() if hasattr(o, "item") else str(o)))


if __name__ == "__main__":
    main()
