"""On-device Philox channel, syndrome and error counting against the oracle (GPU)."""
import numpy as np
import pytest

from tests.oracle_binding import OR_BG1, OR_BG2, OR_WIFI

KIND = {1: OR_BG1, 2: OR_BG2, 3: OR_WIFI}

pytestmark = pytest.mark.gpu
F = np.float32


@pytest.mark.parametrize("kind,Z,m", [(2, 52, 0), (1, 384, 0), (3, 96, 0), (1, 13, 9)])
def test_channel_bit_exact(nr, oracle, torch_gpu, kind, Z, m):
    print(f"case channel kind={kind} Z={Z} m={m}", flush=True)
    torch = torch_gpu
    code = nr.Code(kind, Z, m)
    n_cw, seed, base = 3, 123456789012345, 77
    sigma = nr.noise_stddev(1.0, 0.4)
    d_llr = torch.empty((n_cw, code.n_vn), dtype=torch.float32, device="cuda")
    d_bits = torch.empty((n_cw, code.n_vn), dtype=torch.uint8, device="cuda")
    code.channel(seed, base, n_cw, d_llr.data_ptr(), sigma=sigma, cw_bits_ptr=d_bits.data_ptr())
    torch.cuda.synchronize()
    llr, bits = d_llr.cpu().numpy(), d_bits.cpu().numpy()
    host = nr.Code(kind, Z, m, -1)
    for w in range(n_cw):
        words = np.array([oracle.lib.oracle_info_bit_word(seed, base + w, i) for i in range((code.K + 31) // 32)],
                         dtype=np.uint32)
        info = np.unpackbits(words.view(np.uint8), bitorder="little")[:code.K]
        cw = host.encode_host(info[None, :])[0]
        assert np.array_equal(bits[w], cw)
        p0 = code.punct * Z
        g = np.array([oracle.lib.oracle_gauss(seed, base + w, i) for i in range(code.n_vn - p0)], dtype=F)
        y = (F(1.0) - F(2.0) * cw[p0:].astype(F)) + g * F(sigma)
        want = np.zeros(code.n_vn, dtype=F)
        want[p0:] = (y * F(2.0)) / (F(sigma) * F(sigma))
        assert np.array_equal(llr[w].view(np.uint32), want.view(np.uint32))


# Z >= 32 runs 32 checks per word operation, Z < 32 one check at a time
@pytest.mark.parametrize("kind,Z,m", [(2, 52, 0), (1, 384, 0), (1, 36, 0), (3, 96, 0), (2, 104, 12), (2, 13, 0),
                                      (1, 2, 0)])
def test_syndrome_kernel_matches_oracle(nr, oracle, torch_gpu, kind, Z, m):
    print(f"case syndrome kind={kind} Z={Z} m={m}", flush=True)
    torch = torch_gpu
    code = nr.Code(kind, Z, m)
    rng = np.random.default_rng(Z + kind)
    host = nr.Code(kind, Z, m, -1)
    cw = host.encode_host(rng.integers(0, 2, (6, code.K), dtype=np.uint8))
    cw[1, 5] ^= 1
    cw[2, rng.integers(0, code.n_vn, 40)] ^= 1
    cw[3] = rng.integers(0, 2, code.n_vn)
    cw[4, -1] ^= 1
    packed = np.packbits(cw, axis=1, bitorder="little")
    d = torch.from_numpy(packed).cuda()
    w = torch.zeros(6, dtype=torch.int32, device="cuda")
    code.syndrome(d.data_ptr(), 6, w.data_ptr())
    torch.cuda.synchronize()
    oc = oracle.code(KIND[kind], Z, m)
    assert w.cpu().tolist() == [oracle.syndrome(oc, cw[i]) for i in range(6)]
    assert w[0] == 0 and w[1] > 0 and w[4] > 0


def test_count_errors(nr, torch_gpu):
    print("case count errors", flush=True)
    torch = torch_gpu
    code = nr.Code(1, 64)
    rng = np.random.default_rng(1)
    bits = rng.integers(0, 2, (5, code.n_vn), dtype=np.uint8)
    hd = bits.copy()
    hd[1, 3] ^= 1
    hd[2, :10] ^= 1
    hd[4, code.K:] ^= 1          # parity-only errors are not counted
    d_hd = torch.from_numpy(np.packbits(hd, axis=1, bitorder="little")).cuda()
    d_bits = torch.from_numpy(bits).cuda()
    cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
    code.count_errors(d_hd.data_ptr(), d_bits.data_ptr(), 5, cnt.data_ptr())
    torch.cuda.synchronize()
    assert cnt.tolist() == [11, 2]
