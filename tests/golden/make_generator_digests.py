#!/usr/bin/env python3
"""Write tests/golden/generator_digests.json from the reference's generator tables.

Reads src/Gen_shift_value.h of the reference checkout as text (the OAI encoder tables used by
ldpc_encoder_orig, src/nr_ldpc.cpp:864-891) and records, for every (base graph, Z), a digest
of the canonical generator: for each parity block row i and information block column j the
sorted set of shifts that occur an odd number of times (pairs cancel under XOR).  Only the
digests and counts are stored, not the tables.

Usage: python tests/golden/make_generator_digests.py [/root/reference]
"""
import hashlib
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_arrays(text):
    arrays = {}
    for m in re.finditer(r"short\s+(\w+)\s*\[\s*\d*\s*\]\s*=\s*\{([^}]*)\}", text):
        arrays[m.group(1)] = [int(v) for v in re.findall(r"-?\d+", m.group(2))]
    return arrays


def canonical(no, ptr, gen, rows, ncols, kb):
    parts = []
    for i in range(rows):
        for j in range(kb):
            t = i * ncols + j
            vals = gen[ptr[t]: ptr[t] + no[t]]
            odd = sorted(v for v in set(vals) if vals.count(v) % 2 == 1)
            parts.append(f"{i},{j}:" + ",".join(map(str, odd)))
    return ";".join(parts)


def digest(s):
    return hashlib.sha256(s.encode()).hexdigest()[:16]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    text = open(os.path.join(ref, "src", "Gen_shift_value.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    arr = parse_arrays(text)
    out = {}
    for bg, rows, ncols in ((1, 46, 22), (2, 42, 10)):
        no = arr[f"no_shift_values_BG{bg}"]
        ptr = arr[f"pointer_shift_values_BG{bg}"]
        for name, gen in arr.items():
            m = re.fullmatch(rf"Gen_shift_values_BG{bg}_Z_(\d+)", name)
            if not m:
                continue
            z = int(m.group(1))
            can = canonical(no, ptr, gen, rows, ncols, ncols)
            out[f"BG{bg}_Z{z}"] = {"digest": digest(can), "raw_count": len(gen), "raw_sum": sum(gen)}
    with open(os.path.join(HERE, "generator_digests.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(len(out), "entries")


if __name__ == "__main__":
    main()
