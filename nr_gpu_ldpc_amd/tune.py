"""Autotuning of the mixed-batch decoder's codewords per workgroup, per (base graph, Z).

    python -m nr_gpu_ldpc_amd.tune [--iters 20] [--out nr_gpu_ldpc_amd/tuning/mixed_gfx950.txt]

For every lifting size of both base graphs the library times each candidate tile size on a
batch of full-rate codewords (fixed iterations, no early exit: the work is data independent)
and keeps the fastest; the table is written as "kind Z cpb" lines that
``Mixed.load_tuning`` reads.  A JSON summary (per candidate microseconds per codeword) goes
to stdout.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import nr_gpu_ldpc_amd as nr  # noqa: E402
from nr_gpu_ldpc_amd.mixed import LIFTING_SIZES  # noqa: E402

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "mixed_gfx950.txt")


def batch_for(Z, cpb_max, n_cu=256, cap=65536):
    """Enough codewords for four full waves of tiles at the largest candidate."""
    return int(min(cap, max(1024, 4 * n_cu * cpb_max)))


def tune_all(iters=20, kinds=(nr.BG1, nr.BG2), sizes=None, device=0, log=print):
    sizes = LIFTING_SIZES if sizes is None else sizes
    m = nr.Mixed(65536, device)
    table = {}
    t0 = time.time()
    for kind in kinds:
        for Z in sizes:
            m.set_cpb(kind, Z, 0)
            cmax = m.get_cpb(kind, Z)
            n = batch_for(Z, cmax)
            best, times = m.autotune(kind, Z, iters, n)
            table[f"{kind}/{Z}"] = {"best": best, "n_cw": n, "us_per_cw": {str(k): round(v, 4) for k, v in times.items()}}
            log(f"kind {kind} Z {Z}: best cpb {best} of {sorted(times)} ({time.time() - t0:.1f} s)")
    return m, table


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--iters", type=int, default=20)
    p.add_argument("--out", default=DEFAULT_PATH)
    a = p.parse_args()
    m, table = tune_all(a.iters, log=lambda s: print(s, file=sys.stderr, flush=True))
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    m.save_tuning(a.out)
    print(json.dumps({"saved": os.path.relpath(a.out, ROOT), "iters": a.iters, "table": table}))


if __name__ == "__main__":
    main()
