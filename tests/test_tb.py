"""Transport blocks (TS 38.212): CRC, segmentation, rate matching -- host side."""
import numpy as np
import pytest


def _poly_crc(bits, poly, L):
    """Remainder of M(x) * x^L mod g(x) over GF(2) with Python integers (independent of the
    library's shift register)."""
    m = 0
    for b in bits:
        m = (m << 1) | int(b)
    m <<= L
    g = (1 << L) | poly
    for sh in range(m.bit_length() - 1, L - 1, -1):
        if (m >> sh) & 1:
            m ^= g << (sh - L)
    return m


def test_crc16_check_value(nr):
    bits = np.unpackbits(np.frombuffer(b"123456789", np.uint8))  # MSB first
    assert nr.crc(nr.CRC16, bits) == 0x31C3  # CRC-16 with g = 0x1021, zero register (XMODEM)


@pytest.mark.parametrize("kind,poly,L", [(1, 0x864CFB, 24), (2, 0x800063, 24), (0, 0x1021, 16)])
def test_crc_polynomial_division(nr, kind, poly, L):
    rng = np.random.default_rng(kind)
    for n in (1, 7, 24, 100, 1001):
        bits = rng.integers(0, 2, n, dtype=np.uint8)
        p = nr.crc(kind, bits)
        assert p == _poly_crc(bits, poly, L)
        tail = np.array([(p >> (L - 1 - k)) & 1 for k in range(L)], np.uint8)
        assert nr.crc(kind, np.concatenate([bits, tail])) == 0  # codeword divisible by g


def test_segmentation_cases(nr):
    t = nr.TransportBlock(8424, 1 / 3, 25344, device=-1)
    assert (t.bg, t.C, t.L_tb, t.K_prime, t.Z, t.F, t.L_cb) == (1, 1, 24, 8448, 384, 0, 0)
    t = nr.TransportBlock(25104, 1 / 3, 3 * 25344, device=-1)      # B = 25128 > 8448
    assert (t.bg, t.C, t.L_cb, t.K_prime, t.Z, t.K, t.F) == (1, 3, 24, 8400, 384, 8448, 48)
    t = nr.TransportBlock(3000, 0.3, 10000, device=-1)             # A <= 3824, R <= 0.67 -> BG2
    assert (t.bg, t.C, t.L_tb, t.K_prime, t.Z, t.K, t.F, t.N) == (2, 1, 16, 3016, 320, 3200, 184, 16000)
    t = nr.TransportBlock(100, 0.5, 400, device=-1)                 # A <= 292 -> BG2, Kb = 6
    assert (t.bg, t.K_prime, t.Z) == (2, 116, 20)
    with pytest.raises(nr.LdpcError):
        nr.TransportBlock(20000, 1 / 3, 60000, device=-1)            # B' = 20096 not a multiple of 3


def test_rate_matching_lengths_and_k0(nr):
    t = nr.TransportBlock(25104, 1 / 3, 3 * 20000 + 6, Qm=6, rv=2, device=-1)
    assert t.E.sum() == t.G and all(e % 6 == 0 for e in t.E)
    assert t.E.tolist() == [19998, 20004, 20004]                     # floor first, then ceil (5.4.2.1)
    assert t.k0 == 33 * t.Z                                          # BG1 rv2, full buffer
    for rv, k in ((0, 0), (1, 13), (2, 25), (3, 43)):
        assert nr.TransportBlock(3000, 0.3, 10000, Qm=2, rv=rv, device=-1).k0 == k * 320


def _select(d_len, k0, filler, E):
    out, j = [], 0
    while len(out) < E:
        p = (k0 + j) % d_len
        j += 1
        if filler[0] <= p < filler[1]:
            continue
        out.append(p)
    return out


@pytest.mark.parametrize("A,rate,G,Qm,rv", [(8424, 1 / 3, 25344, 2, 0), (25104, 1 / 3, 70008, 4, 1),
                                            (3000, 0.3, 7002, 6, 3), (3000, 0.3, 40000, 2, 2), (100, 0.5, 400, 1, 0)])
def test_encode_chain(nr, oracle, A, rate, G, Qm, rv):
    """CRCs, H c = 0 per code block, and the transmitted bits as bit selection + interleaving
    of the code blocks (restated here from TS 38.212 5.4.2)."""
    from tests.oracle_binding import OR_BG1, OR_BG2
    t = nr.TransportBlock(A, rate, G, Qm=Qm, rv=rv, device=-1)
    rng = np.random.default_rng(A + rv)
    a = rng.integers(0, 2, A, dtype=np.uint8)
    f, cw = t.encode_host(a)
    oc = oracle.code(OR_BG1 if t.bg == 1 else OR_BG2, t.Z)
    b = []
    for r in range(t.C):
        assert oracle.syndrome(oc, cw[r]) == 0
        assert not cw[r, t.K_prime:t.K].any()                          # filler bits encoded as 0
        if t.C > 1:
            assert nr.crc(nr.CRC24B, cw[r, :t.K_prime]) == 0
        b.append(cw[r, :t.K_prime - t.L_cb])
    b = np.concatenate(b)[:t.B]
    assert np.array_equal(b[:A], a)
    assert nr.crc(nr.CRC24A if t.L_tb == 24 else nr.CRC16, b) == 0
    off = 0
    for r in range(t.C):
        E = int(t.E[r])
        d = cw[r, 2 * t.Z:]
        pos = _select(t.N, t.k0, (t.K_prime - 2 * t.Z, t.K - 2 * t.Z), E)
        e = d[pos]
        want = e.reshape(Qm, E // Qm).T.reshape(-1)                    # f[i + j Qm] = e[i E/Qm + j]
        assert np.array_equal(f[off:off + E], want)
        off += E
