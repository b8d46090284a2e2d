#!/usr/bin/env python3
"""Generate the base-graph data file ``nr_gpu_ldpc_amd/csrc/bg_tables.inc``.

One-time build step. It reads the TS 38.212 base matrices as *text* from the
reference checkout (``src/h_matrices/h_bg_{1,2}_i_{0..7}.h``) and the 12x24
example matrix from ``Overview.md:22-33``, repairs the known defects of the
shipped BG1 tables, verifies every table against fixed sha256 digests (as
shipped and corrected), and writes a compact table in the layout TS 38.212
itself uses (Table 5.3.2-2/3): one record per nonzero block position
``(row, col)`` carrying the eight shift values V for i_LS = 0..7.

Defects repaired (both verified by H.c = 0 for every lifting size):
  * BG1 i_LS in {2,3,5,6,7}: 16 spurious ``1`` entries in column 0 (rows listed
    in COL0_TYPO_ROWS) that must be ``-1`` (a lost minus sign).
  * BG1 i_LS = 2 (64 entries) and i_LS = 4 (18 entries) hold V - 256 instead
    of V (values stored modulo 256); the true values are restored.

Usage: python gen_bg_tables.py [/root/reference]
"""
import hashlib
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "..", "csrc", "bg_tables.inc")

COL0_TYPO_ROWS = [10, 16, 18, 21, 23, 25, 27, 29, 31, 33, 35, 37, 39, 41, 43, 45]

# (row, col, shipped, true) for BG1 i_LS=2 and i_LS=4 entries stored mod 256.
MOD256_I2 = [
    (0, 15, 42, 298), (1, 0, 47, 303), (1, 2, 38, 294), (1, 4, 5, 261), (1, 12, 44, 300),
    (1, 15, 10, 266), (1, 19, 4, 260), (1, 21, 45, 301), (2, 4, 24, 280), (2, 13, 39, 295),
    (2, 14, 27, 283), (2, 15, 45, 301), (2, 20, 20, 276), (3, 10, 25, 281), (3, 11, 47, 303),
    (3, 20, 63, 319), (5, 1, 36, 292), (5, 12, 62, 318), (5, 21, 11, 267), (5, 22, 23, 279),
    (6, 0, 33, 289), (6, 10, 37, 293), (6, 17, 46, 302), (8, 0, 39, 295), (8, 16, 40, 296),
    (8, 21, 13, 269), (9, 11, 30, 286), (9, 17, 11, 267), (10, 14, 55, 311), (11, 12, 34, 290),
    (11, 16, 33, 289), (11, 23, 24, 280), (13, 3, 46, 302), (13, 7, 47, 303), (14, 12, 43, 299),
    (14, 21, 2, 258), (15, 1, 34, 290), (16, 22, 44, 300), (17, 0, 1, 257), (18, 1, 4, 260),
    (18, 12, 38, 294), (18, 13, 35, 291), (18, 19, 39, 295), (19, 8, 14, 270), (20, 0, 45, 301),
    (21, 20, 27, 283), (22, 17, 60, 316), (24, 0, 33, 289), (24, 3, 24, 280), (24, 22, 25, 281),
    (25, 6, 39, 295), (26, 0, 14, 270), (26, 4, 62, 318), (27, 8, 48, 304), (28, 4, 37, 293),
    (29, 14, 52, 308), (31, 7, 52, 308), (32, 14, 15, 271), (34, 7, 40, 296), (35, 22, 29, 285),
    (36, 0, 3, 259), (37, 1, 42, 298), (44, 9, 3, 259), (44, 22, 4, 260),
]
MOD256_I4 = [
    (0, 21, 5, 261), (1, 4, 0, 256), (2, 0, 2, 258), (2, 18, 13, 269), (5, 0, 5, 261),
    (5, 12, 27, 283), (10, 7, 11, 267), (11, 12, 18, 274), (11, 21, 17, 273), (16, 1, 14, 270),
    (20, 0, 9, 265), (21, 16, 10, 266), (22, 17, 1, 257), (24, 11, 10, 266), (26, 15, 21, 277),
    (28, 0, 19, 275), (35, 6, 22, 278), (43, 25, 14, 270),
]

# sha256[:16] over the int32 little-endian row-major array: (as shipped, corrected)
DIGESTS = {
    (1, 0): ("89db242aff60e42e", "89db242aff60e42e"),
    (1, 1): ("814ffd07ea0110da", "814ffd07ea0110da"),
    (1, 2): ("f64f59d1c2c7361d", "608227bc8842e1e1"),
    (1, 3): ("972929eecded6131", "f9afcfb70543cfe6"),
    (1, 4): ("2cc225fcabef871a", "4b33630c787db1a4"),
    (1, 5): ("d303461510be2866", "6256e9e3369195ce"),
    (1, 6): ("ec6b2dc0b3ef1485", "00f0c248bac6a2a2"),
    (1, 7): ("d36c37ccef37f156", "784c8d906475b205"),
    (2, 0): ("9e8a11386b9faea9",) * 2,
    (2, 1): ("e499f93ab7e48955",) * 2,
    (2, 2): ("eb25f59479c3425d",) * 2,
    (2, 3): ("5060200a3efab28d",) * 2,
    (2, 4): ("3fc54220663c4836",) * 2,
    (2, 5): ("8f925ed2aac20659",) * 2,
    (2, 6): ("d3eaa280a0b1ff42",) * 2,
    (2, 7): ("d7d561321e03fb39",) * 2,
}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()[:16]


def parse_header(path, rows, cols):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    body = text[text.index("= {") + 3: text.rindex("}")]
    vals = [int(v) for v in re.findall(r"-?\d+", body)]
    assert len(vals) == rows * cols, (path, len(vals))
    return np.array(vals, dtype=np.int32).reshape(rows, cols)


def parse_overview(path):
    text = open(path).read()
    start = text.index("```", text.index("802.11 base matrix"))
    block = text[start + 3: text.index("```", start + 3)]
    vals = [int(v) for v in re.findall(r"-?\d+", block)]
    assert len(vals) == 12 * 24
    return np.array(vals, dtype=np.int32).reshape(12, 24)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    tables = {}
    for bg, (rows, cols) in ((1, (46, 68)), (2, (42, 52))):
        for ils in range(8):
            a = parse_header(os.path.join(ref, "src", "h_matrices", f"h_bg_{bg}_i_{ils}.h"), rows, cols)
            shipped, corrected = DIGESTS[(bg, ils)]
            assert digest(a) == shipped, (bg, ils, digest(a))
            if bg == 1 and ils in (2, 3, 5, 6, 7):
                for r in COL0_TYPO_ROWS:
                    assert a[r, 0] == 1
                    a[r, 0] = -1
            if bg == 1 and ils in (2, 4):
                for (r, c, s, t) in (MOD256_I2 if ils == 2 else MOD256_I4):
                    assert a[r, c] == s, (ils, r, c, a[r, c])
                    a[r, c] = t
            assert digest(a) == corrected, (bg, ils, digest(a))
            tables[(bg, ils)] = a
        # every lifting set shares one sparsity pattern
        pat = tables[(bg, 0)] >= 0
        for ils in range(8):
            assert np.array_equal(tables[(bg, ils)] >= 0, pat), (bg, ils)
    wifi = parse_overview(os.path.join(ref, "Overview.md"))
    assert int((wifi >= 0).sum()) == 86 and int(wifi.max()) == 79

    out = []
    out.append("// Generated by nr_gpu_ldpc_amd/tables/gen_bg_tables.py -- do not edit.")
    out.append("// TS 38.212 Table 5.3.2-2 (BG1) and 5.3.2-3 (BG2) in the standard's own layout:")
    out.append("// {row, col, {V for i_LS = 0..7}}, rows ascending, columns ascending within a row.")
    out.append("// The 12x24 QC example code (IEEE 802.11n rate-1/2 prototype) follows as {row, col, V}.")
    out.append("// The includer defines NRLDPC_TABLE_QUAL (static const in C, static constexpr in C++).")
    for bg in (1, 2):
        pat = tables[(bg, 0)] >= 0
        rows, cols = pat.shape
        recs = []
        for r in range(rows):
            for c in range(cols):
                if pat[r, c]:
                    vs = [int(tables[(bg, i)][r, c]) for i in range(8)]
                    recs.append(f"  {{{r}, {c}, {{{', '.join(str(v) for v in vs)}}}}},")
        out.append(f"#define NRLDPC_BG{bg}_NNZ {len(recs)}")
        out.append(f"NRLDPC_TABLE_QUAL nrldpc_bg_entry kBG{bg}Entries[NRLDPC_BG{bg}_NNZ] = {{")
        out.extend(recs)
        out.append("};")
    recs = []
    for r in range(12):
        for c in range(24):
            if wifi[r, c] >= 0:
                recs.append(f"  {{{r}, {c}, {int(wifi[r, c])}}},")
    out.append(f"#define NRLDPC_WIFI_NNZ {len(recs)}")
    out.append("NRLDPC_TABLE_QUAL nrldpc_qc_entry kWifiEntries[NRLDPC_WIFI_NNZ] = {")
    out.extend(recs)
    out.append("};")
    with open(OUT, "w") as f:
        f.write("\n".join(out) + "\n")
    print("wrote", os.path.normpath(OUT))


if __name__ == "__main__":
    main()
