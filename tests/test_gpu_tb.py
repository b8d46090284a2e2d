"""Transport blocks on the GPU: de-rate-matching against a restatement of TS 38.212 5.4.2, and
the whole chain (CRC, segmentation, LDPC, rate matching, AWGN, de-rate-matching, decoding,
code-block CRC checks, reassembly)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _derate_ref(t, llr):
    out = np.zeros((t.C, t.n_vn), np.float32)
    out[:, t.K_prime:t.K] = np.float32(10000.0)
    off = 0
    for r in range(t.C):
        E, Q = int(t.E[r]), t.Qm
        pos, j = [], 0
        while len(pos) < E:
            p = (t.k0 + j) % t.N
            j += 1
            if t.K_prime - 2 * t.Z <= p < t.K - 2 * t.Z:
                continue
            pos.append(p)
        for k in range(E):  # ascending bit-selection order, sequential float32 sums
            m = (k // (E // Q)) + (k % (E // Q)) * Q
            c = 2 * t.Z + pos[k]
            out[r, c] = np.float32(out[r, c] + llr[off + m])
        off += E
    return out


CASES = [(8424, 1 / 3, 25344, 2, 0), (25104, 1 / 3, 70008, 4, 1), (3000, 0.3, 7002, 6, 3),
         (3000, 0.3, 40000, 2, 2), (100, 0.5, 400, 1, 0)]


@pytest.mark.parametrize("A,rate,G,Qm,rv", CASES)
def test_derate_matches_restatement(nr, torch_gpu, A, rate, G, Qm, rv):
    torch = torch_gpu
    print(f"case derate A={A} G={G} Qm={Qm} rv={rv}", flush=True)
    t = nr.TransportBlock(A, rate, G, Qm=Qm, rv=rv)
    llr = np.random.default_rng(G).standard_normal(G).astype(np.float32) * 3
    d_in = torch.from_numpy(llr).cuda()
    d_out = torch.full((t.C, t.n_vn), -7.0, dtype=torch.float32, device="cuda")
    t.derate(d_in.data_ptr(), d_out.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(d_out.cpu().numpy(), _derate_ref(t, llr))


@pytest.mark.parametrize("A,rate,G,Qm,rv", CASES)
def test_transport_block_round_trip(nr, torch_gpu, A, rate, G, Qm, rv):
    torch = torch_gpu
    print(f"case tb chain A={A} G={G} Qm={Qm} rv={rv}", flush=True)
    t = nr.TransportBlock(A, rate, G, Qm=Qm, rv=rv)
    rng = np.random.default_rng(A * 7 + rv)
    a = rng.integers(0, 2, A, dtype=np.uint8)
    f, _ = t.encode_host(a)
    sigma = np.float32(0.45)
    y = (1.0 - 2.0 * f.astype(np.float32)) + sigma * rng.standard_normal(G).astype(np.float32)
    llr = ((y * np.float32(2)) / (sigma * sigma)).astype(np.float32)
    code = nr.Code(t.bg, t.Z)
    d_in = torch.from_numpy(llr).cuda()
    d_cb = torch.empty((t.C, t.n_vn), dtype=torch.float32, device="cuda")
    d_hd = torch.empty((t.C, code.hd_bytes), dtype=torch.uint8, device="cuda")
    ok = torch.zeros(t.C, dtype=torch.int32, device="cuda")
    bits = torch.zeros(t.B, dtype=torch.uint8, device="cuda")
    t.derate(d_in.data_ptr(), d_cb.data_ptr())
    code.decode(d_cb.data_ptr(), t.C, 25, d_hd.data_ptr(), flags=nr.EARLY_EXIT)
    t.check(d_hd.data_ptr(), ok.data_ptr(), bits.data_ptr())
    torch.cuda.synchronize()
    assert ok.cpu().numpy().tolist() == [1] * t.C
    b = bits.cpu().numpy()
    assert np.array_equal(b[:A], a)
    assert nr.crc(nr.CRC24A if t.L_tb == 24 else nr.CRC16, b) == 0
    # a block decoded from no information fails its check
    if t.C > 1:
        d_cb.zero_()
        code.decode(d_cb.data_ptr(), t.C, 5, d_hd.data_ptr())
        t.check(d_hd.data_ptr(), ok.data_ptr())
        torch.cuda.synchronize()
        assert ok.cpu().numpy().tolist() == [0] * t.C
