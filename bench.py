#!/usr/bin/env python3
"""Benchmark of the decode hot path: flooding scaled min-sum, 5G NR BG1 Z=384, T=10 iterations,
batches of 4096 codewords (info Gbps per GPU), with the kernel's roofline and a CPU baseline.

A step = one decode launch over one batch of 4096 codewords whose LLRs are already resident
in HBM (generated on the device by the Philox channel kernel before timing).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]

Prints one JSON line.
"""
import argparse
import json
import os
import statistics
import sys
import time

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


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--cpu-sample", type=int, default=1024, help="codewords in the CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a 256 MiB device-to-device copy (scale factor for PMC byte counters)")
    return p.parse_args()


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


def cpu_baseline(llr_host, n_cw):
    """The oracle (the reference's min-sum restated in C) on this host's cores."""
    from tests import oracle_binding
    orc = oracle_binding.load()
    code = orc.code(oracle_binding.OR_BG1, Z)
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    orc.decode(code, llr_host[:n_cw], N_ITER, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(n_cw * 22 * Z / dt / 1e9, 6), "unit": "Gbps", "cores": threads, "kind": "port",
            "sample": f"{n_cw} codewords of the same BG1 Z=384 T=10 workload, decoded by the C oracle "
                      f"with {threads} threads ({dt:.2f} s wall)"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the decode path would shard "
                                   "codewords across GPUs (see DESIGN.md), run with --gpus 1"}))
        sys.exit(2)
    import torch
    import nr_gpu_ldpc_amd as nr

    code = nr.Code(nr.BG1, Z)
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
    total_ms = ev[0].elapsed_time(ev[-1])
    ms = total_ms / args.steps
    info_bits = BATCH * code.K
    gbps = info_bits * args.steps / (wall * 1e9)

    counts = torch.zeros(2, dtype=torch.int64, device=dev)
    code.count_errors(hd.data_ptr(), bits.data_ptr(), BATCH, counts.data_ptr(), stream=s)
    torch.cuda.synchronize()
    bit_err, frame_err = counts.tolist()

    # roofline of the decode kernel (one launch = one batch)
    edges = code.nnz * Z
    valu_ops = BATCH * edges * N_ITER * VALU_OPS_PER_EDGE
    alg_bytes = BATCH * (4 * code.n_vn + code.hd_bytes)
    valu_gops = valu_ops / (ms * 1e6)
    hbm_gbs = alg_bytes / (ms * 1e6)
    p50 = statistics.median(per)
    p99 = sorted(per)[max(0, int(round(0.99 * len(per))) - 1)]

    out = {
        "metric": "info Gbps/GPU at FER 1e-3, learned vs plain min-sum; p99 batch latency",
        "value": round(gbps, 4),
        "unit": "Gbps",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Philox-generated info bits, systematic encoding, BPSK over AWGN at "
                f"Eb/N0 {EBN0_DB} dB, generated on the device before timing",
        "config": {"workload": "5G NR BG1 Z=384 (K=8448, N=25344), flooding scaled min-sum, "
                               "10 fixed iterations, batch 4096",
                   "bg": 1, "Z": Z, "n_iter": N_ITER, "batch": BATCH, "ebn0_db": EBN0_DB,
                   "cw_per_block": code.cw_per_block},
        "latency_ms": {"p50": round(p50, 4), "p99": round(p99, 4), "kind": "device time per batch"},
        "fer": round(frame_err / BATCH, 6),
        "ber": bit_err / (BATCH * code.K),
        "roofline": {
            "bound": "valu",
            "achieved": round(valu_gops, 1),
            "peak": round(PEAK_VALU_GOPS, 1),
            "unit": "Gop/s",
            "frac": round(valu_gops / PEAK_VALU_GOPS, 4),
            "traffic": None,
            "per_launch": {"edge_iterations": BATCH * edges * N_ITER, "ops_per_edge": VALU_OPS_PER_EDGE,
                           "algorithmic_bytes": alg_bytes},
            "hbm": {"achieved": round(hbm_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(hbm_gbs / PEAK_HBM_GBS, 5)},
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(llr.cpu().numpy(), min(args.cpu_sample, BATCH))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
