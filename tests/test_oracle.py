"""The CPU oracle pinned against things other than itself (CPU only).

* Philox4x32-10 against the Random123 known-answer vectors.
* The Gaussian stream by its moments.
* The decoder against hand-derived closed forms and against an independent dense-matrix
  numpy restatement of the reference kernels (first-strict-minimum index rule, float32).
* The learned decoder: alpha=0.75/beta=0 is the plain decoder bit for bit; general weights
  equal the dense restatement; the analytic gradient agrees with double-precision finite
  differences.
"""
import ctypes

import numpy as np
import pytest

from tests.oracle_binding import OR_BG1, OR_BG2, OR_WIFI

F = np.float32


def test_philox_known_answers(oracle):
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        c = (ctypes.c_uint32 * 4)(*ctr)
        k = (ctypes.c_uint32 * 2)(*key)
        o = (ctypes.c_uint32 * 4)()
        oracle.lib.oracle_philox4x32_10(c, k, o)
        assert list(o) == want


def test_gaussian_moments(oracle):
    g = np.array([oracle.lib.oracle_gauss(5, 17, i) for i in range(200000)])
    assert abs(g.mean()) < 0.01
    assert abs(g.var() - 1) < 0.01
    assert abs(((g - g.mean()) ** 4).mean() / g.var() ** 2 - 3) < 0.05
    assert abs((np.abs(g) < 1).mean() - 0.6827) < 0.005


def test_info_bits_balanced(oracle):
    words = np.array([oracle.lib.oracle_info_bit_word(3, 9, w) for w in range(20000)], dtype=np.uint32)
    bits = np.unpackbits(words.view(np.uint8))
    assert abs(bits.mean() - 0.5) < 0.005


def _dense_H(oracle, code):
    V = oracle.shifts(code)
    Z = code.Z
    rows, cols = V.shape
    H = np.zeros((rows * Z, cols * Z), dtype=np.uint8)
    for i in range(rows):
        for j in range(cols):
            if V[i, j] >= 0:
                for r in range(Z):
                    H[i * Z + r, j * Z + (r + V[i, j]) % Z] = 1
    return H


def dense_minsum(H, llr, n_iter, half=False, char=False, alpha=None, beta=None, eidx=None):
    """Flooding scaled min-sum on a dense H, written from the reference kernels' arithmetic.
    half: the HALF_PRC build (R, dt and the APP between iterations rounded to binary16).
    char: the CHAR_PRC build (R and dt truncated to signed char, saturating; the APP between
    iterations clamped to [-128, 127]; a saturated APP enters the check node as it is).
    alpha, beta [n_iter][n_base_edges], eidx [M][N] (base edge of each nonzero of H): the learned
    rule R = sign * alpha * max(m - beta, 0) of the iteration's weights.
    Returns hard decisions, APP and, for char, the number of out-of-range conversions."""
    M, N = H.shape
    st = (lambda v: np.asarray(v, F).astype(np.float16).astype(F)) if half else (lambda v: np.asarray(v, F))
    oor = [0]

    def to_char(v):
        t = np.trunc(np.asarray(v, F))
        oor[0] += int(np.count_nonzero((t > 127) | (t < -128)))
        return np.clip(t, -128, 127).astype(F)

    app = llr.astype(F).copy()
    R = np.zeros((M, N), dtype=F)
    nbr = [np.nonzero(H[m])[0] for m in range(M)]
    for t in range(n_iter):
        dt = np.zeros((M, N), dtype=F)
        for m in range(M):
            cols = nbr[m]
            Q = app[cols] if t == 0 else (app[cols] - R[m, cols]).astype(F)
            if char:
                Q = np.where((app[cols] == 127) | (app[cols] == -128), app[cols], Q).astype(F)
            absq = np.abs(Q)
            vals = np.concatenate([absq, [F(1000.0), F(1000.0)]]).astype(F)
            idx = int(np.argmin(absq)) if absq.min() < F(1000.0) else 0  # first strict minimum
            srt = np.sort(vals)
            m1, m2 = srt[0], srt[1]
            sq = np.where(Q < 0, F(-1.0), F(1.0))
            sign = F(np.prod(sq))
            mag = np.where(np.arange(len(cols)) == idx, m2, m1).astype(F)
            if alpha is None:
                Rn = ((F(0.75) * sign) * sq * mag).astype(F)
            else:
                b = eidx[m, cols]
                off = (mag - beta[t, b].astype(F)).astype(F)
                off = np.where(off > 0, off, F(0.0)).astype(F)
                Rn = (sign * sq * (alpha[t, b].astype(F) * off).astype(F)).astype(F)
            if char:
                Rc = to_char(Rn)
                dt[m, cols] = Rc if t == 0 else to_char(Rc - R[m, cols])
                R[m, cols] = Rc
                continue
            dt[m, cols] = st(Rn if t == 0 else (Rn - R[m, cols]).astype(F))
            R[m, cols] = st(Rn)
        for m in range(M):  # ascending block row per column
            cols = nbr[m]
            app[cols] = (app[cols] + dt[m, cols]).astype(F)
        if t + 1 < n_iter:
            app = np.clip(app, -128, 127).astype(F) if char else st(app)
    if char:
        return (app <= 0).astype(np.uint8), app, oor[0]
    return (app <= 0).astype(np.uint8), app


@pytest.mark.parametrize("kind,Z,m", [(OR_WIFI, 5, 0), (OR_BG2, 2, 0), (OR_BG1, 3, 0), (OR_BG2, 3, 6)])
def test_oracle_matches_dense_restatement(oracle, kind, Z, m):
    code = oracle.code(kind, Z, m)
    H = _dense_H(oracle, code)
    rng = np.random.default_rng(Z * 31 + kind)
    n = H.shape[1]
    llr = (rng.standard_normal((3, n)) * 2.5 + 1.0).astype(F)
    llr[:, :2 * Z] = 0 if kind != OR_WIFI else llr[:, :2 * Z]
    llr[2] = np.round(llr[2])  # ties
    hd, app, _ = oracle.decode(code, llr, 5)
    for w in range(3):
        dhd, dapp = dense_minsum(H, llr[w], 5)
        assert np.array_equal(hd[w], dhd)
        assert np.array_equal(app[w], dapp)


def _row_sets(oracle, code):
    H = _dense_H(oracle, code)
    return H, H.sum(axis=0)


def test_closed_form_single_negative(oracle):
    """All LLRs +5 except one column at -2, one iteration, Z=1 (the base matrix itself)."""
    code = oracle.code(OR_WIFI, 1)
    H, deg = _row_sets(oracle, code)
    n = H.shape[1]
    c0 = 4
    llr = np.full((1, n), 5.0, dtype=F)
    llr[0, c0] = -2.0
    hd, app, _ = oracle.decode(code, llr, 1)
    for c in range(n):
        rows = np.nonzero(H[:, c])[0]
        if c == c0:
            want = -2.0 + 3.75 * len(rows)                 # other signs +, m2 = 5 -> 0.75*5
        else:
            want = 5.0 + sum(-1.5 if H[m, c0] else 3.75 for m in rows)  # min 2 with a - sign
        assert app[0, c] == F(want)


def test_closed_form_clamp(oracle):
    """Every |LLR| above the 1000.0f start value: every message is 0.75 * 1000."""
    code = oracle.code(OR_WIFI, 3)
    H, deg = _row_sets(oracle, code)
    llr = np.full((1, H.shape[1]), 2000.0, dtype=F)
    _, app, _ = oracle.decode(code, llr, 1)
    assert np.array_equal(app[0], (2000.0 + 750.0 * deg).astype(F))


@pytest.mark.parametrize("kind,Z", [(OR_BG1, 8), (OR_BG2, 6), (OR_WIFI, 4)])
def test_noiseless_codeword_one_iteration(oracle, nr, kind, Z):
    code = nr.Code({OR_BG1: 1, OR_BG2: 2, OR_WIFI: 3}[kind], Z, 0, -1)
    info = np.random.default_rng(Z).integers(0, 2, (2, code.K), dtype=np.uint8)
    cw = code.encode_host(info)
    llr = (10.0 * (1.0 - 2.0 * cw)).astype(F)
    llr[:, :code.punct * Z] = 0.0
    hd, _, _ = oracle.decode(oracle.code(kind, Z), llr, 1)
    assert np.array_equal(hd, cw)


def test_early_exit_stops_on_zero_syndrome(oracle, nr):
    code = nr.Code(2, 16, 0, -1)
    info = np.random.default_rng(1).integers(0, 2, (4, code.K), dtype=np.uint8)
    cw = code.encode_host(info)
    llr = (4.0 * (1.0 - 2.0 * cw)).astype(F)
    llr[:, :2 * 16] = 0.0
    oc = oracle.code(OR_BG2, 16)
    hd, _, iters = oracle.decode(oc, llr, 30, early_exit=True)
    assert (iters < 30).all() and np.array_equal(hd, cw)


@pytest.mark.parametrize("kind,Z", [(OR_BG2, 6), (OR_WIFI, 5)])
def test_noms_plain_weights_equal_plain_decoder(oracle, kind, Z):
    code = oracle.code(kind, Z)
    nnz = int((oracle.shifts(code) >= 0).sum())
    rng = np.random.default_rng(2)
    llr = (rng.standard_normal((3, code.cols * Z)) * 2 + 0.7).astype(F)
    T = 6
    a = np.full(T * nnz, 0.75, dtype=F)
    b = np.zeros(T * nnz, dtype=F)
    hd, app = oracle.decode_noms(code, llr, T, a, b)
    phd, papp, _ = oracle.decode(code, llr, T)
    assert np.array_equal(hd, phd) and np.array_equal(app.view(np.uint32), papp.view(np.uint32))


@pytest.mark.parametrize("kind,Z,m", [(OR_WIFI, 5, 0), (OR_BG2, 2, 0), (OR_BG1, 3, 0), (OR_BG2, 3, 6)])
def test_oracle_noms_matches_dense_restatement(oracle, kind, Z, m):
    """Learned forward with general per-(iteration, edge) weights: the oracle equals the dense
    restatement bit for bit (and differs from the plain decoder)."""
    code = oracle.code(kind, Z, m)
    H = _dense_H(oracle, code)
    V = oracle.shifts(code)
    eidx = np.full(H.shape, -1, dtype=np.int64)
    e = 0
    for i in range(V.shape[0]):
        for j in range(V.shape[1]):
            if V[i, j] >= 0:
                for r in range(Z):
                    eidx[i * Z + r, j * Z + (r + V[i, j]) % Z] = e
                e += 1
    T = 5
    rng = np.random.default_rng(Z * 7 + kind)
    alpha = rng.uniform(0.5, 1.0, (T, e)).astype(F)
    beta = rng.uniform(0.0, 0.6, (T, e)).astype(F)
    n = H.shape[1]
    llr = (rng.standard_normal((3, n)) * 2.5 + 1.0).astype(F)
    if kind != OR_WIFI:
        llr[:, :2 * Z] = 0
    hd, app = oracle.decode_noms(code, llr, T, alpha.reshape(-1), beta.reshape(-1))
    for w in range(3):
        dhd, dapp = dense_minsum(H, llr[w], T, alpha=alpha, beta=beta, eidx=eidx)
        assert np.array_equal(hd[w], dhd)
        assert np.array_equal(app[w].view(np.uint32), dapp.view(np.uint32))
    _, papp, _ = oracle.decode(code, llr, T)
    assert not np.array_equal(papp, app)


@pytest.mark.parametrize("kind,Z", [(OR_BG2, 6), (OR_BG1, 5), (OR_WIFI, 7)])
def test_noms_early_exit_plain_weights_equal_plain_early_exit(oracle, kind, Z):
    """alpha = 0.75, beta = 0 with early exit is the plain decoder with early exit, bit for bit,
    iteration counts included (on codewords that stop at different iterations)."""
    code = oracle.code(kind, Z)
    nnz = int((oracle.shifts(code) >= 0).sum())
    rng = np.random.default_rng(11 * Z + kind)
    n = code.cols * Z
    # all-zero codeword (valid for every code) through noise of several strengths
    llr = (2.0 + rng.standard_normal((24, n)) * np.linspace(0.4, 4.0, 24)[:, None]).astype(F)
    if kind != OR_WIFI:
        llr[:, :2 * Z] = 0.0
    T = 12
    a = np.full(T * nnz, 0.75, dtype=F)
    b = np.zeros(T * nnz, dtype=F)
    hd, app, it = oracle.decode_noms(code, llr, T, a, b, early_exit=True, with_iters=True)
    phd, papp, pit = oracle.decode(code, llr, T, early_exit=True)
    assert np.array_equal(it, pit) and np.array_equal(hd, phd)
    assert np.array_equal(app.view(np.uint32), papp.view(np.uint32))
    assert len(set(it.tolist())) >= 3 and it.min() < T and it.max() == T


@pytest.mark.parametrize("kind,Z,m", [(OR_BG2, 3, 0), (OR_WIFI, 4, 0), (OR_BG1, 2, 8)])
def test_noms_early_exit_is_first_zero_syndrome_of_fixed_schedule(oracle, kind, Z, m):
    """Learned decode with early exit and general weights = the dense restatement run for T'
    iterations, where T' is the first iteration count whose hard decision has zero syndrome by
    brute force (H x over GF(2)), or T if none: hard decisions, APP and iteration counts."""
    code = oracle.code(kind, Z, m)
    H = _dense_H(oracle, code)
    V = oracle.shifts(code)
    eidx = np.full(H.shape, -1, dtype=np.int64)
    e = 0
    for i in range(V.shape[0]):
        for j in range(V.shape[1]):
            if V[i, j] >= 0:
                for r in range(Z):
                    eidx[i * Z + r, j * Z + (r + V[i, j]) % Z] = e
                e += 1
    T = 6
    rng = np.random.default_rng(5 * Z + kind + m)
    alpha = rng.uniform(0.55, 1.0, (T, e)).astype(F)
    beta = rng.uniform(0.0, 0.4, (T, e)).astype(F)
    n = H.shape[1]
    llr = (1.5 + rng.standard_normal((6, n)) * np.linspace(0.5, 2.0, 6)[:, None]).astype(F)
    if kind != OR_WIFI:
        llr[:, :2 * Z] = 0
    hd, app, it = oracle.decode_noms(code, llr, T, alpha.reshape(-1), beta.reshape(-1), early_exit=True,
                                     with_iters=True)
    seen = set()
    for w in range(llr.shape[0]):
        stop = T
        for t in range(1, T + 1):
            dhd, dapp = dense_minsum(H, llr[w], t, alpha=alpha[:t], beta=beta[:t], eidx=eidx)
            if not (H.astype(np.int64) @ dhd.astype(np.int64) % 2).any():
                stop = t
                break
        dhd, dapp = dense_minsum(H, llr[w], stop, alpha=alpha[:stop], beta=beta[:stop], eidx=eidx)
        assert it[w] == stop
        assert np.array_equal(hd[w], dhd)
        assert np.array_equal(app[w].view(np.uint32), dapp.view(np.uint32))
        seen.add(stop)
    assert len(seen) >= 2


def test_noms_gradient_matches_finite_differences(oracle, nr):
    code = oracle.code(OR_BG2, 3)
    nnz = int((oracle.shifts(code) >= 0).sum())
    T = 3
    rng = np.random.default_rng(4)
    c = nr.Code(2, 3, 0, -1)
    info = rng.integers(0, 2, (2, c.K), dtype=np.uint8)
    bits = c.encode_host(info)
    llr = ((1.0 - 2.0 * bits) * 1.2 + rng.standard_normal(bits.shape) * 1.1).astype(F)
    llr[:, :6] = 0.0
    alpha = rng.uniform(0.6, 0.9, T * nnz)
    beta = rng.uniform(-0.2, 0.2, T * nnz)
    loss, ga, gb = oracle.noms_loss_grad(code, llr, bits, T, alpha, beta)
    h = 1e-6
    idx = rng.choice(T * nnz, 12, replace=False)
    for i in idx:
        for vec, g in ((alpha, ga), (beta, gb)):
            v0 = vec[i]
            vec[i] = v0 + h
            lp, _, _ = oracle.noms_loss_grad(code, llr, bits, T, alpha, beta, grad=False)
            vec[i] = v0 - h
            lm, _, _ = oracle.noms_loss_grad(code, llr, bits, T, alpha, beta, grad=False)
            vec[i] = v0
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-5 + 1e-4 * abs(fd), (i, fd, g[i])
    assert np.abs(ga).sum() > 0 and np.abs(gb).sum() > 0


def test_weighted_noms_loss_is_the_weighted_sum_of_per_codeword_losses(oracle, nr):
    """Per-codeword loss weights: the weighted loss and gradient of a batch equal the weighted
    sum of each codeword's own (one-codeword) loss and gradient divided by the batch size; all-one
    weights give the unweighted entry point's values; the weighted gradient matches central
    finite differences of the weighted loss."""
    code = oracle.code(OR_BG2, 3)
    nnz = int((oracle.shifts(code) >= 0).sum())
    T = 3
    rng = np.random.default_rng(9)
    c = nr.Code(2, 3, 0, -1)
    bits = c.encode_host(rng.integers(0, 2, (3, c.K), dtype=np.uint8))
    llr = ((1.0 - 2.0 * bits) * 1.2 + rng.standard_normal(bits.shape) * 1.1).astype(F)
    llr[:, :6] = 0.0
    alpha = rng.uniform(0.6, 0.9, T * nnz)
    beta = rng.uniform(-0.2, 0.2, T * nnz)
    w = np.array([0.25, 2.0, 0.0])
    loss, ga, gb = oracle.noms_loss_grad(code, llr, bits, T, alpha, beta, cw_weight=w)
    parts = [oracle.noms_loss_grad(code, llr[k:k + 1], bits[k:k + 1], T, alpha, beta) for k in range(3)]
    assert np.isclose(loss, sum(w[k] * parts[k][0] for k in range(3)) / 3, rtol=1e-12, atol=0)
    assert np.allclose(ga, sum(w[k] * parts[k][1] for k in range(3)) / 3, rtol=1e-10, atol=1e-15)
    assert np.allclose(gb, sum(w[k] * parts[k][2] for k in range(3)) / 3, rtol=1e-10, atol=1e-15)
    ones = oracle.noms_loss_grad(code, llr, bits, T, alpha, beta, cw_weight=np.ones(3))
    plain = oracle.noms_loss_grad(code, llr, bits, T, alpha, beta)
    assert ones[0] == plain[0] and np.array_equal(ones[1], plain[1]) and np.array_equal(ones[2], plain[2])
    h = 1e-6
    for i in rng.choice(T * nnz, 6, replace=False):
        for vec, g in ((alpha, ga), (beta, gb)):
            v0 = vec[i]
            vec[i] = v0 + h
            lp, _, _ = oracle.noms_loss_grad(code, llr, bits, T, alpha, beta, grad=False, cw_weight=w)
            vec[i] = v0 - h
            lm, _, _ = oracle.noms_loss_grad(code, llr, bits, T, alpha, beta, grad=False, cw_weight=w)
            vec[i] = v0
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-5 + 1e-4 * abs(fd), (i, fd, g[i])


@pytest.mark.parametrize("kind,Z,m", [(OR_WIFI, 5, 0), (OR_BG1, 3, 0), (OR_BG2, 4, 9)])
def test_syndrome_is_dense_parity_check(oracle, kind, Z, m):
    """oracle_syndrome = weight of H x over GF(2) with H expanded densely (brute force)."""
    code = oracle.code(kind, Z, m)
    H = _dense_H(oracle, code)
    rng = np.random.default_rng(Z + 7 * kind)
    for _ in range(5):
        x = rng.integers(0, 2, H.shape[1], dtype=np.uint8)
        assert oracle.syndrome(code, x) == int(((H.astype(np.int64) @ x) % 2).sum())
    assert oracle.syndrome(code, np.zeros(H.shape[1], np.uint8)) == 0


def test_llr_closed_form(oracle):
    """oracle_llr: 2z punctured zeros, then (y * 2) / (sigma * sigma) in float32."""
    import ctypes
    recv = np.array([0.5, -1.25, 2.0, 0.0, -0.75, 1.0, 3.5], dtype=F)
    out = np.full(9, np.nan, dtype=F)
    oracle.lib.oracle_llr(recv.ctypes.data, 9, ctypes.c_float(0.8), 1, out.ctypes.data)
    s = F(0.8)
    want = np.concatenate([[0, 0], (recv * F(2)) / (s * s)]).astype(F)
    assert np.array_equal(out, want)
    out2 = np.full(4, np.nan, dtype=F)
    oracle.lib.oracle_llr(recv.ctypes.data, 4, ctypes.c_float(0.5), 0, out2.ctypes.data)
    assert out2.tolist() == [4.0, -10.0, 16.0, 0.0]  # sigma = 0.5: LLR = 8 y exactly


def test_half_conversions_match_ieee(oracle):
    """binary16 rounding (ties to even, subnormals, overflow) against numpy's IEEE conversion."""
    rng = np.random.default_rng(0)
    v = (rng.standard_normal(4000) * 10.0 ** rng.integers(-9, 6, 4000)).astype(F)
    v = np.concatenate([v, np.array([0.0, -0.0, 65504, 65519.99, 65520, 1e-8, 2 ** -24, 2 ** -25, 3 * 2 ** -26,
                                     6.1e-5, np.inf, -np.inf], F)])
    with np.errstate(over="ignore"):
        want = v.astype(np.float16).view(np.uint16)
    got = np.array([oracle.lib.oracle_float_to_half(float(x)) for x in v], np.uint16)
    assert np.array_equal(got, want)
    h = np.arange(0, 65536, 97, dtype=np.uint16)
    back = np.array([oracle.lib.oracle_half_to_float(int(x)) for x in h], F)
    ref = h.view(np.float16).astype(F)
    assert np.array_equal(back.view(np.uint32)[~np.isnan(ref)], ref.view(np.uint32)[~np.isnan(ref)])


@pytest.mark.parametrize("kind,Z,m", [(OR_WIFI, 5, 0), (OR_BG2, 2, 0), (OR_BG1, 3, 0)])
def test_oracle_half_matches_dense_restatement(oracle, kind, Z, m):
    """HALF_PRC: the oracle equals the dense restatement with binary16 storage points."""
    code = oracle.code(kind, Z, m)
    H = _dense_H(oracle, code)
    rng = np.random.default_rng(Z * 13 + kind)
    n = H.shape[1]
    llr = (rng.standard_normal((3, n)) * 2.5 + 1.0).astype(np.float16)
    if kind != OR_WIFI:
        llr[:, :2 * Z] = 0
    llr[2] = np.round(llr[2])
    hd, app, _ = oracle.decode_half(code, llr.view(np.uint16), 5)
    for w in range(3):
        dhd, dapp = dense_minsum(H, llr[w].astype(F), 5, half=True)
        assert np.array_equal(hd[w], dhd)
        assert np.array_equal(app[w], dapp)
    # and it differs from the single-precision decoder somewhere (the storage points matter)
    f_hd, f_app, _ = oracle.decode(code, llr.astype(F), 5)
    assert not np.array_equal(f_app, app)


def _quantize(x, D=0.5, B=8):
    """The reference's quantize(D, x, B) (src/host_helpers.cpp:46-58): floor(x / D), clamped to
    [-2^(B-1), 2^(B-1) - 1]."""
    q = np.floor(np.asarray(x, np.float64) / D)
    return np.clip(q, -(1 << (B - 1)), (1 << (B - 1)) - 1).astype(np.int8)


@pytest.mark.parametrize("kind,Z,m,scale", [(OR_WIFI, 5, 0, 20.0), (OR_BG2, 2, 0, 40.0), (OR_BG1, 3, 0, 30.0),
                                            (OR_BG2, 3, 6, 90.0)])
def test_oracle_char_matches_dense_restatement(oracle, kind, Z, m, scale):
    """CHAR_PRC: the oracle equals the dense restatement with signed-char storage points, on
    LLRs large enough that saturated APPs and out-of-range conversions occur."""
    code = oracle.code(kind, Z, m)
    H = _dense_H(oracle, code)
    rng = np.random.default_rng(Z * 17 + kind)
    n = H.shape[1]
    llr = _quantize(rng.standard_normal((3, n)) * scale + 0.3 * scale)
    if kind != OR_WIFI:
        llr[:, :2 * Z] = 0
    hd, app, _, oor = oracle.decode_char(code, llr, 6)
    total = 0
    for w in range(3):
        dhd, dapp, door = dense_minsum(H, llr[w].astype(F), 6, char=True)
        assert np.array_equal(hd[w], dhd)
        assert np.array_equal(app[w], dapp)
        total += door
    assert oor == total
    assert (np.abs(app) >= 127).any()  # saturation is exercised
    if (kind, Z) == (OR_BG2, 2):
        assert oor > 0  # and out-of-range conversions


def test_quantize_matches_reference_rule():
    x = np.array([-100.0, -64.2, -0.4, -0.0, 0.0, 0.2, 0.5, 0.74, 63.4, 63.5, 70.0])
    assert _quantize(x).tolist() == [-128, -128, -1, 0, 0, 0, 1, 1, 126, 127, 127]


def test_char_workload_has_no_out_of_range_conversions(oracle, nr):
    """On the benchmark workload (BG1 Z=384, Eb/N0 1.0-2.0 dB, T=10, LLRs from quantize(0.5, x,
    8)) no float -> char conversion of the CHAR_PRC schedule leaves [-128, 127], so its results
    there are fully defined by C++ (truncation toward zero) and the saturating rule never acts."""
    code = nr.Code(1, 384, 0, -1)  # host-only context: encoder only
    rng = np.random.default_rng(8)
    info = rng.integers(0, 2, (3, code.K), dtype=np.uint8)
    cw = code.encode_host(info)
    oc = oracle.code(OR_BG1, 384)
    for ebn0 in (1.0, 1.5, 2.0):
        sigma = nr.noise_stddev(ebn0, 1 / 3)
        x = 1.0 - 2.0 * cw[:, 2 * 384:].astype(F)
        y = (x + sigma * rng.standard_normal(x.shape).astype(F)).astype(F)
        llr = np.zeros((3, code.n_vn), F)
        llr[:, 2 * 384:] = (y * F(2.0)) / F(sigma * sigma)
        _, _, _, oor = oracle.decode_char(oc, _quantize(llr), 10)
        assert oor == 0, ebn0
