"""Autotuning of the decoder launch shapes.

    python -m nr_gpu_ldpc_amd.tune [--iters 20] [--out nr_gpu_ldpc_amd/tuning/mixed_gfx950.txt]
    python -m nr_gpu_ldpc_amd.tune --decode [--iters 10]

Mixed batches: for every lifting size of both base graphs the library times each candidate tile
size on a batch of full-rate codewords (fixed iterations, no early exit: the work is data
independent) and keeps the fastest; the table is written as "kind Z cpb" lines that
``Mixed.load_tuning`` reads.

Single-code decodes (--decode): for every (base graph, Z) and the 12x24 code's sizes,
``ldpc_autotune`` times both LDS layouts (dup, slim) at each codewords-per-workgroup count on whole
rounds of the device; the fastest count of each layout and its round time are written to
csrc/decode_tuning.inc, the table compiled into the library (rebuild to use it), and with
--cache PATH also to a JSON cache that ``apply_cache`` installs into a context at run time
(``ldpc_set_layout_model``), with no rebuild and no timing.

A JSON summary (per candidate timings) goes to stdout.
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
DECODE_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "decode_tuning.inc")
QC_SIZES = (27, 54, 81, 96)


def batch_for(Z, cpb_max, n_cu=256, cap=65536):
    """Enough codewords for four full waves of tiles at the largest candidate."""
    return int(min(cap, max(1024, 4 * n_cu * cpb_max)))


def tune_all(iters=20, kinds=(nr.BG1, nr.BG2), sizes=None, device=0, log=print,
             layouts=(nr.LAYOUT_DUP, nr.LAYOUT_SLIM)):
    """Both layouts' codewords per workgroup for every (base graph, Z); the object ends in the
    first layout, holding the tuned counts of all."""
    sizes = LIFTING_SIZES if sizes is None else sizes
    m = nr.Mixed(65536, device)
    table = {}
    t0 = time.time()
    for layout in reversed(layouts):
        m.set_layout(layout)
        name = "slim" if layout == nr.LAYOUT_SLIM else "dup"
        for kind in kinds:
            for Z in sizes:
                m.set_cpb(kind, Z, 0)
                cmax = m.get_cpb(kind, Z)
                n = batch_for(Z, cmax)
                best, times = m.autotune(kind, Z, iters, n)
                table[f"{name}/{kind}/{Z}"] = {"best": best, "n_cw": n,
                                               "us_per_cw": {str(k): round(v, 4) for k, v in times.items()}}
                log(f"{name} kind {kind} Z {Z}: best cpb {best} of {sorted(times)} ({time.time() - t0:.1f} s)")
    return m, table


def best_of(cands, layout):
    """The candidate of `layout` with the least time per codeword of a round, or None."""
    cs = [c for c in cands if c["layout"] == layout]
    return min(cs, key=lambda c: c["round_us"] / c["cw_per_cu"]) if cs else None


def tune_decode(iters=10, device=0, log=print):
    shapes = [(k, z) for k in (nr.BG1, nr.BG2) for z in LIFTING_SIZES] + [(nr.QC12X24, z) for z in QC_SIZES]
    table, lines, models = {}, [], {}
    t0 = time.time()
    for kind, Z in shapes:
        code = nr.Code(kind, Z, 0, device)
        cands = code.autotune(n_iter=iters)
        models[f"{kind}/{Z}"] = layout_models(code)
        code.close()
        d, s = best_of(cands, nr.LAYOUT_DUP), best_of(cands, nr.LAYOUT_SLIM)
        table[f"{kind}/{Z}"] = cands
        fmt = lambda c: "{0, 0, 0.0f}" if c is None else f"{{{c['cpb']}, {c['cw_per_cu']}, {c['round_us']:.1f}f}}"
        lines.append(f"{{{kind}, {Z}, {fmt(d)}, {fmt(s)}}},")
        log(f"kind {kind} Z {Z}: dup {d and (d['cpb'], round(d['round_us'] / d['cw_per_cu'], 3))} "
            f"slim {s and (s['cpb'], round(s['round_us'] / s['cw_per_cu'], 3))} us/cw ({time.time() - t0:.1f} s)")
    return table, lines, models


def layout_models(code):
    """A context's launch cost model: {"dup": {cpb, cw_per_cu, round_us}, "slim": {...}}."""
    return {"dup": code.layout_model(nr.LAYOUT_DUP), "slim": code.layout_model(nr.LAYOUT_SLIM)}


def save_cache(path, models, device_name=""):
    """Write {"kind/Z": layout_models(...)} as a JSON autotune cache."""
    with open(path, "w") as f:
        json.dump({"format": "nr_gpu_ldpc_amd layout models 1", "device": device_name, "codes": models}, f,
                  indent=1, sort_keys=True)


def apply_cache(code, path):
    """Install the cached model of code's (kind, Z) into the context; False when the cache has none.
    The slim layout is dropped from the model where the cache records none."""
    with open(path) as f:
        doc = json.load(f)
    if doc.get("format") != "nr_gpu_ldpc_amd layout models 1":
        raise ValueError(f"{path}: not a layout-model cache")
    m = doc["codes"].get(f"{code.kind}/{code.Z}")
    if m is None:
        return False
    d, s = m["dup"], m["slim"]
    code.set_layout_model(nr.LAYOUT_DUP, d["cpb"], d["cw_per_cu"], d["round_us"])
    if s["cpb"] > 0:
        code.set_layout_model(nr.LAYOUT_SLIM, s["cpb"], s["cw_per_cu"], s["round_us"])
    else:
        code.set_layout_model(nr.LAYOUT_SLIM, 0)
    return True


def write_decode_table(lines, iters, path=DECODE_TABLE):
    import torch
    props = torch.cuda.get_device_properties(0)
    arch = getattr(props, "gcnArchName", "gfx950").split(":")[0]
    name = "MI355X" if arch == "gfx950" else props.name
    head = ["// Generated by nr_gpu_ldpc_amd/tune.py --decode -- do not edit.",
            f"// Launch shapes of the single-code decoder measured on {name} ({arch}, "
            f"{props.multi_processor_count} CUs), T = {iters}, no early exit:",
            "// per (kind, Z) the fastest codewords per workgroup of each LDS layout,",
            "// {kind, Z, {dup cpb, dup codewords per CU per round, dup round us},",
            "//           {slim cpb, slim codewords per CU per round, slim round us}}"]
    with open(path, "w") as f:
        f.write("\n".join(head + lines) + "\n")


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--iters", type=int, default=None)
    p.add_argument("--out", default=DEFAULT_PATH)
    p.add_argument("--decode", action="store_true", help="tune the single-code decoder's launch layouts")
    p.add_argument("--table", default=DECODE_TABLE, help="--decode: where to write the table")
    p.add_argument("--cache", default="", help="--decode: also write a JSON cache of the models (apply_cache)")
    a = p.parse_args()
    if a.decode:
        iters = a.iters or 10
        table, lines, models = tune_decode(iters, log=lambda s: print(s, file=sys.stderr, flush=True))
        write_decode_table(lines, iters, a.table)
        if a.cache:
            import torch
            save_cache(a.cache, models, torch.cuda.get_device_properties(0).name)
        print(json.dumps({"saved": os.path.relpath(a.table, ROOT), "iters": iters, "table": table}))
        return
    a.iters = a.iters or 20
    m, table = tune_all(a.iters, log=lambda s: print(s, file=sys.stderr, flush=True))
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    m.save_tuning(a.out)
    print(json.dumps({"saved": os.path.relpath(a.out, ROOT), "iters": a.iters, "table": table}))


if __name__ == "__main__":
    main()
