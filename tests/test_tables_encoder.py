"""Base-graph tables, code parameters, encoder and generator tables (CPU only)."""
import hashlib
import json
import os

import numpy as np
import pytest

from tests.oracle_binding import OR_BG1, OR_BG2, OR_WIFI

HERE = os.path.dirname(os.path.abspath(__file__))

# sha256[:16] over the int32 little-endian row-major table after the known corrections of the
# shipped BG1 tables (lost minus sign in column 0 of i_LS 2,3,5,6,7; values stored mod 256 in
# i_LS 2 and 4); BG2 and BG1 i_LS 0/1 are as shipped.
CORRECTED = {
    (1, 0): "89db242aff60e42e", (1, 1): "814ffd07ea0110da", (1, 2): "608227bc8842e1e1",
    (1, 3): "f9afcfb70543cfe6", (1, 4): "4b33630c787db1a4", (1, 5): "6256e9e3369195ce",
    (1, 6): "00f0c248bac6a2a2", (1, 7): "784c8d906475b205",
    (2, 0): "9e8a11386b9faea9", (2, 1): "e499f93ab7e48955", (2, 2): "eb25f59479c3425d",
    (2, 3): "5060200a3efab28d", (2, 4): "3fc54220663c4836", (2, 5): "8f925ed2aac20659",
    (2, 6): "d3eaa280a0b1ff42", (2, 7): "d7d561321e03fb39",
}
LIFTING = sorted({a * 2 ** j for a in (2, 3, 5, 7, 9, 11, 13, 15) for j in range(8) if a * 2 ** j <= 384})
BG1_DEG = [19, 19, 19, 19, 3, 8, 9, 7, 10, 9, 7, 8, 7, 6, 7, 7, 6, 6, 6, 6, 6, 6, 5, 5, 6, 5, 5, 4, 5, 5,
           5, 5, 5, 5, 5, 5, 5, 4, 5, 5, 4, 5, 4, 5, 5, 4]
BG2_DEG = [8, 10, 8, 10, 4, 6, 6, 6, 4, 5, 5, 5, 4, 5, 5, 4, 5, 5, 4, 4, 4, 4, 3, 4, 4, 3, 5, 3, 4, 3,
           5, 3, 4, 4, 4, 4, 4, 3, 4, 4, 4, 4]


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()[:16]


@pytest.mark.parametrize("bg,ils", sorted(CORRECTED))
def test_table_digest(nr, oracle, bg, ils):
    t = nr.base_matrix(bg, ils)
    assert _digest(t) == CORRECTED[(bg, ils)]
    assert np.array_equal(t, oracle.base_matrix(bg, ils))
    deg = (t >= 0).sum(axis=1).tolist()
    assert deg == (BG1_DEG if bg == 1 else BG2_DEG)


def test_qc12x24_table(nr):
    t = nr.base_matrix(nr.QC12X24, 0)
    assert int((t >= 0).sum()) == 86 and int(t.max()) == 79
    assert (t >= 0).sum(axis=1).tolist() == [7, 7, 7, 7, 7, 7, 8, 7, 7, 7, 7, 8]
    # first parity column: shifts (1, 0, 1) whose circulant sum is the identity
    assert sorted(int(v) for v in t[:, 12] if v >= 0) == [0, 1, 1]


def test_lifting_sizes_count():
    assert len(LIFTING) == 51


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("Z", LIFTING)
def test_encoder_parity_all_lifting_sizes(nr, oracle, kind, Z):
    code = nr.Code(kind, Z, 0, -1)
    rng = np.random.default_rng(Z + 1000 * kind)
    info = rng.integers(0, 2, (2, code.K), dtype=np.uint8)
    cw = code.encode_host(info)
    oc = oracle.code(OR_BG1 if kind == 1 else OR_BG2, Z)
    assert np.array_equal(cw[:, :code.K], info)
    for w in range(2):
        assert oracle.syndrome(oc, cw[w]) == 0


@pytest.mark.parametrize("kind,Z,m", [(1, 384, 4), (1, 384, 17), (1, 104, 30), (2, 208, 4), (2, 208, 22),
                                      (2, 15, 41)])
def test_encoder_truncated_rows(nr, oracle, kind, Z, m):
    code = nr.Code(kind, Z, m, -1)
    assert code.cols == (22 if kind == 1 else 10) + m
    info = np.random.default_rng(m).integers(0, 2, (1, code.K), dtype=np.uint8)
    cw = code.encode_host(info)
    assert oracle.syndrome(oracle.code(OR_BG1 if kind == 1 else OR_BG2, Z, m), cw[0]) == 0


@pytest.mark.parametrize("Z", [2, 5, 13, 81, 96, 384])
def test_encoder_qc12x24(nr, oracle, Z):
    code = nr.Code(nr.QC12X24, Z, 0, -1)
    info = np.random.default_rng(Z).integers(0, 2, (3, code.K), dtype=np.uint8)
    cw = code.encode_host(info)
    for w in range(3):
        assert oracle.syndrome(oracle.code(OR_WIFI, Z), cw[w]) == 0


def test_encoder_linear(nr):
    """Encoding is linear over GF(2): enc(a ^ b) = enc(a) ^ enc(b)."""
    code = nr.Code(2, 52, 0, -1)
    rng = np.random.default_rng(9)
    a = rng.integers(0, 2, (1, code.K), dtype=np.uint8)
    b = rng.integers(0, 2, (1, code.K), dtype=np.uint8)
    assert np.array_equal(code.encode_host(a ^ b), code.encode_host(a) ^ code.encode_host(b))


def test_generator_tables_match_reference(nr):
    """Regenerated OAI-layout generator tables equal the reference's (pairs cancelled)."""
    gold = json.load(open(os.path.join(HERE, "golden", "generator_digests.json")))
    assert len(gold) == 102
    for key, ref in gold.items():
        bg, Z = int(key[2]), int(key.split("Z")[1])
        no, ptr, gen = nr.generator_shifts(bg, Z)
        kb = 22 if bg == 1 else 10
        parts = []
        for t in range(len(no)):
            vals = sorted(int(v) for v in gen[ptr[t]:ptr[t] + no[t]])
            parts.append(f"{t // kb},{t % kb}:" + ",".join(map(str, vals)))
        assert hashlib.sha256(";".join(parts).encode()).hexdigest()[:16] == ref["digest"], key


def test_generator_bg2_raw_digests(nr):
    """For BG2 Z=208 even the uncancelled OAI count tables are reproduced."""
    no, ptr, gen = nr.generator_shifts(2, 208)
    d = lambda a: hashlib.sha256(np.ascontiguousarray(a, dtype="<i2").tobytes()).hexdigest()[:12]
    assert d(no) == "b27167e7251d" and d(ptr) == "5545ccae8ad4"
    assert len(gen) == 1483 and int(gen.astype(np.int64).sum()) == 159543


@pytest.mark.parametrize("A,rate", [(8424, 1 / 3), (2064, 0.2), (292, 0.5), (500, 0.8), (3000, 0.6),
                                    (3000, 0.7), (3900, 0.9), (100, 0.3), (640, 0.25), (8000, 0.5),
                                    (3824, 0.67), (1000, 0.67)])
def test_setup_params_vs_oracle(nr, oracle, A, rate):
    p = nr.setup_params(A, rate)
    assert list(p.values()) == oracle.setup_ldpc(A, rate)


def test_setup_params_closed_forms(nr):
    p = nr.setup_params(8424, 1 / 3)     # the reference's main(): BG1, Z=384, i_LS=1
    assert (p["bg"], p["Z"], p["ils"], p["K"], p["N"], p["N_before_puncture"]) == (1, 384, 1, 8448, 25344, 26112)
    assert (p["n_nz_in_row"], p["n_nz_in_col"]) == (19, 30)
    p = nr.setup_params(2064, 0.2)
    assert (p["bg"], p["Z"], p["ils"], p["K"], p["N"]) == (2, 208, 6, 2080, 10400)
    assert (p["n_nz_in_row"], p["n_nz_in_col"]) == (10, 23)


# Known answers worked by hand from TS 38.212 (5.1 CRC length, 7.2.2 base-graph choice, 5.2.2
# K_b, Table 5.3.2-1 lifting sizes) with the reference's own two deviations kept
# (src/nr_ldpc.cpp:37-112): Z_min = floor(K'/K_b) where the standard asks for K_b * Z >= K', and
# the rate compared as a float promoted to double.  Tuple: A, rate -> bg, Z, i_LS, K.
SETUP_KNOWN = [
    # BG choice: A <= 292 takes BG2 at any rate; B = 308 -> K_b = 8, Z >= 38 -> 40 (5 * 8)
    (292, 0.9, 2, 40, 2, 400),
    # A = 293 at a high rate: BG1; B = 309, Z >= floor(309/22) = 14 -> 14 (7 * 2), so K = 308 < B
    # (the reference's floor; the standard would take Z = 15)
    (293, 0.9, 1, 14, 3, 308),
    # A < 3824 and R <= 0.67: BG2 (B = 3016, K_b = 10, Z >= 301 -> 320 = 5 * 64) ...
    (3000, 0.66, 2, 320, 2, 3200),
    # ... 0.67 as a float is 0.67000001... > 0.67: BG1 (Z >= 137 -> 144 = 9 * 16)
    (3000, 0.67, 1, 144, 4, 3168),
    (3000, 0.7, 1, 144, 4, 3168),
    # A = 3823 < 3824: 16 CRC bits, BG2, B = 3839 <= K_cb = 3840, Z >= 383 -> 384
    (3823, 0.5, 2, 384, 1, 3840),
    # A = 3824: still 16 CRC bits but not < 3824, so BG1 at R = 0.5; B = 3840, Z >= 174 -> 176
    (3824, 0.5, 1, 176, 5, 3872),
    # A = 3825: 24 CRC bits (A > 3824), B = 3849, BG1, Z >= 174 -> 176
    (3825, 0.5, 1, 176, 5, 3872),
    # R <= 0.25 takes BG2 where neither A rule does (A = 3824: B = 3840 = K_cb, Z >= 384 -> 384)
    (3824, 0.25, 2, 384, 1, 3840),
    # BG2 K_b thresholds of 5.2.2: B = 192 -> 6; 193 -> 8; 560 -> 8; 561 -> 9; 640 -> 9; 641 -> 10
    (176, 0.2, 2, 32, 0, 320),    # Z >= 192/6 = 32 -> 32 (2 * 16)
    (177, 0.2, 2, 24, 1, 240),    # Z >= floor(193/8) = 24 -> 24 (3 * 8)
    (544, 0.2, 2, 72, 4, 720),    # Z >= 560/8 = 70 -> 72 (9 * 8)
    (545, 0.2, 2, 64, 0, 640),    # Z >= floor(561/9) = 62 -> 64 (2 * 32)
    (624, 0.2, 2, 72, 4, 720),    # Z >= floor(640/9) = 71 -> 72
    (625, 0.2, 2, 64, 0, 640),    # Z >= floor(641/10) = 64 -> 64
]


@pytest.mark.parametrize("A,rate,bg,Z,ils,K", SETUP_KNOWN)
def test_setup_params_known_answers(nr, oracle, A, rate, bg, Z, ils, K):
    p = nr.setup_params(A, rate)
    assert (p["bg"], p["Z"], p["ils"], p["K"]) == (bg, Z, ils, K)
    assert p["N"] == (66 if bg == 1 else 50) * Z and p["N_before_puncture"] == p["N"] + 2 * Z
    assert (p["H_rows"], p["H_cols"]) == ((46, 68) if bg == 1 else (42, 52))
    assert list(p.values()) == oracle.setup_ldpc(A, rate)


@pytest.mark.parametrize("A,rate", [(8425, 1 / 3), (3841, 0.2), (5000, 0.25), (20000, 0.5)])
def test_setup_params_refuses_segmentation(nr, A, rate):
    """B = A + L beyond the base graph's K_cb (8448 for BG1, 3840 for BG2) needs C > 1 code
    blocks (5.2.2), which the reference's setup_ldpc does not support (src/nr_ldpc.cpp:78-81):
    the library says so instead of returning an invalid lifting size; the transport-block chain
    (ldpc_tb_create) segments such blocks."""
    with pytest.raises(nr.LdpcError) as e:
        nr.setup_params(A, rate)
    assert e.value.status == -5


def test_invalid_codes_rejected(nr):
    with pytest.raises(nr.LdpcError):
        nr.Code(1, 17, 0, -1)        # not a lifting size
    with pytest.raises(nr.LdpcError):
        nr.Code(1, 384, 3, -1)       # fewer than the 4 core rows
    with pytest.raises(nr.LdpcError):
        nr.Code(2, 384, 43, -1)
    with pytest.raises(nr.LdpcError):
        nr.Code(7, 384, 0, -1)


@pytest.mark.parametrize("ebn0,rate", [(0.0, 1 / 3), (1.5, 1 / 3), (3.0, 0.2), (-1.0, 0.5), (10.0, 0.5)])
def test_noise_stddev_matches_oracle(nr, oracle, ebn0, rate):
    s = nr.noise_stddev(ebn0, rate)
    assert s == oracle.lib.oracle_noise_stddev(ebn0, rate)
    assert abs(s - np.sqrt(1 / rate / 10 ** (ebn0 / 10) / 2)) < 1e-6


def test_committed_noms_weights_load(nr):
    """The shipped learned weights have the base graph's edge count and sane values."""
    import numpy as np
    from nr_gpu_ldpc_amd import noms
    rec = noms.load_weights(noms.weights_path(2, 208, 10))
    info = nr.Code(2, 208, 0, -1)
    assert (rec["kind"], rec["Z"], rec["n_iter"]) == (2, 208, 10)
    assert rec["weights"].shape == (10, info.nnz_full, 2)
    assert np.isfinite(rec["weights"]).all()
    assert 0.2 < rec["weights"][..., 0].mean() < 1.5
    assert noms.degree1_edges(2, 208).sum() == 38
