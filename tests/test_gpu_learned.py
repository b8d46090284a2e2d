"""Learned offset/normalized min-sum on the GPU: forward parity, gradients, fallback."""
import numpy as np
import pytest

from tests.oracle_binding import OR_BG1, OR_BG2, OR_WIFI
from tests.test_gpu_decode import _gpu_decode, _noisy_llr

pytestmark = pytest.mark.gpu
KIND = {1: OR_BG1, 2: OR_BG2, 3: OR_WIFI}


def _learned(torch, code, llr, n_iter, flags=0):
    d_llr = torch.from_numpy(llr).cuda()
    n = llr.shape[0]
    d_hd = torch.zeros((n, code.hd_bytes), dtype=torch.uint8, device="cuda")
    d_app = torch.zeros((n, code.n_vn), dtype=torch.float32, device="cuda")
    d_it = torch.zeros(n, dtype=torch.int32, device="cuda")
    code.decode_learned(d_llr.data_ptr(), n, n_iter, d_hd.data_ptr(), d_app.data_ptr(), d_it.data_ptr(), flags=flags)
    torch.cuda.synchronize()
    hd = np.unpackbits(d_hd.cpu().numpy(), axis=1, bitorder="little")[:, :code.n_vn]
    return hd, d_app.cpu().numpy(), d_it.cpu().numpy()


@pytest.mark.parametrize("kind,Z", [(2, 208), (1, 384), (2, 52), (3, 96)])
def test_plain_weights_equal_plain_decoder(nr, torch_gpu, kind, Z):
    print(f"case learned-as-plain kind={kind} Z={Z}", flush=True)
    code = nr.Code(kind, Z)
    T = 8
    code.set_weights(code.plain_weights(T))
    llr, _ = _noisy_llr(nr, code, 6, 0.5, 0.25, seed=Z)
    hd, app, _ = _learned(torch_gpu, code, llr, T)
    phd, papp, _ = _gpu_decode(torch_gpu, code, llr, T)
    assert np.array_equal(hd, phd)
    assert np.array_equal(app.view(np.uint32), papp.view(np.uint32))


@pytest.mark.parametrize("kind,Z", [(2, 208), (2, 13), (1, 64), (3, 24)])
def test_learned_matches_oracle(nr, oracle, torch_gpu, kind, Z):
    print(f"case learned vs oracle kind={kind} Z={Z}", flush=True)
    code = nr.Code(kind, Z)
    T = 6
    rng = np.random.default_rng(kind * 100 + Z)
    w = np.stack([rng.uniform(0.5, 1.0, (T, code.nnz_full)), rng.uniform(-0.3, 0.5, (T, code.nnz_full))], -1)
    w = w.astype(np.float32)
    code.set_weights(w)
    assert np.array_equal(code.get_weights(), w)
    llr, _ = _noisy_llr(nr, code, 5, 1.0, 0.25, seed=Z + 1)
    hd, app, _ = _learned(torch_gpu, code, llr, T)
    oc = oracle.code(KIND[kind], Z)
    ohd, oapp = oracle.decode_noms(oc, llr, T, w[..., 0].reshape(-1), w[..., 1].reshape(-1))
    assert np.array_equal(hd, ohd)
    assert np.array_equal(app, oapp)


@pytest.mark.parametrize("kind,Z,T,weighted", [(2, 13, 4, False), (3, 8, 3, False), (1, 16, 3, False),
                                                (2, 13, 4, True)])
def test_gradient_matches_oracle(nr, oracle, torch_gpu, kind, Z, T, weighted):
    """GPU backward vs the oracle's double-precision analytic gradient (itself checked
    against finite differences on the CPU), unweighted and with a loss weight per codeword."""
    print(f"case gradient kind={kind} Z={Z} T={T} weighted={weighted}", flush=True)
    torch = torch_gpu
    code = nr.Code(kind, Z)
    rng = np.random.default_rng(7 + Z)
    w = np.stack([rng.uniform(0.6, 0.9, (T, code.nnz_full)), rng.uniform(-0.2, 0.2, (T, code.nnz_full))], -1)
    w = w.astype(np.float32)
    code.set_weights(w)
    n_cw = 6
    llr, bits = _noisy_llr(nr, code, n_cw, 1.0, 0.3, seed=Z)
    d_llr = torch.from_numpy(llr).cuda()
    d_bits = torch.from_numpy(bits).cuda()
    grad = torch.zeros((T, code.nnz_full, 2), dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    ws = torch.empty(code.workspace_bytes(n_cw, T), dtype=torch.uint8, device="cuda")
    cw_w = np.array([0.5, 2.0, 0.0, 1.0, 3.0, 0.25], dtype=np.float32) if weighted else None
    d_w = torch.from_numpy(cw_w).cuda() if weighted else None
    code.train_grad(d_llr.data_ptr(), d_bits.data_ptr(), n_cw, T, grad.data_ptr(), loss.data_ptr(),
                    ws.data_ptr(), ws.numel(), weight_ptr=d_w.data_ptr() if weighted else 0)
    torch.cuda.synchronize()
    oc = oracle.code(KIND[kind], Z)
    oloss, ga, gb = oracle.noms_loss_grad(oc, llr, bits, T, w[..., 0].reshape(-1).astype(np.float64),
                                          w[..., 1].reshape(-1).astype(np.float64),
                                          cw_weight=None if cw_w is None else cw_w.astype(np.float64))
    g = grad.cpu().numpy()
    assert abs(float(loss.item()) - oloss) <= 1e-4 * abs(oloss) + 1e-6
    for got, want in ((g[..., 0].reshape(-1), ga), (g[..., 1].reshape(-1), gb)):
        scale = np.abs(want).max()
        assert scale > 0
        assert np.abs(got - want).max() <= 2e-3 * scale, (np.abs(got - want).max(), scale)


def test_fallback_calls_on_two_streams_share_the_scratch_safely(nr, torch_gpu):
    """Three fallback decodes on one context issued back to back on two streams, the second one
    larger (its scratch grows while the first may still run): every output equals the same decode
    run alone."""
    torch = torch_gpu
    print("case fallback on two streams BG2 Z=52", flush=True)
    code = nr.Code(2, 52)
    T = 10
    w = code.plain_weights(T)
    w[..., 0] = 0.0  # every noisy codeword fails, so the re-decode path does real work
    code.set_weights(w)
    sizes = [96, 400, 160]
    llrs = [_noisy_llr(nr, code, n, 1.0, 0.2, seed=40 + i)[0] for i, n in enumerate(sizes)]
    lone = [_learned(torch, code, l, T, flags=nr.FALLBACK_PLAIN) for l in llrs]
    code2 = nr.Code(2, 52)  # a fresh context: its scratch starts empty and grows at call 2
    code2.set_weights(w)
    streams = [torch.cuda.Stream(), torch.cuda.Stream(), None]
    streams[2] = streams[0]
    outs = []
    for l, s in zip(llrs, streams):
        n = l.shape[0]
        d_llr = torch.from_numpy(l).cuda()
        d_hd = torch.zeros((n, code2.hd_bytes), dtype=torch.uint8, device="cuda")
        d_app = torch.zeros((n, code2.n_vn), dtype=torch.float32, device="cuda")
        d_it = torch.zeros(n, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()  # inputs in place before the calls are issued back to back
        outs.append((d_llr, d_hd, d_app, d_it, s))
    for d_llr, d_hd, d_app, d_it, s in outs:
        code2.decode_learned(d_llr.data_ptr(), d_llr.shape[0], T, d_hd.data_ptr(), d_app.data_ptr(),
                             d_it.data_ptr(), flags=nr.FALLBACK_PLAIN, stream=s.cuda_stream)
    torch.cuda.synchronize()
    for (d_llr, d_hd, d_app, d_it, _), (lhd, lapp, lit) in zip(outs, lone):
        hd = np.unpackbits(d_hd.cpu().numpy(), axis=1, bitorder="little")[:, :code2.n_vn]
        assert np.array_equal(hd, lhd)
        assert np.array_equal(d_app.cpu().numpy().view(np.uint32), lapp.view(np.uint32))
        assert np.array_equal(d_it.cpu().numpy(), lit)


def test_fallback_to_plain(nr, oracle, torch_gpu):
    """alpha = 0 freezes the learned decoder at the channel decision, so every noisy codeword
    fails its syndrome; the plain re-decode replaces it wherever plain min-sum converges to a
    zero syndrome, and the learned output stays elsewhere."""
    print("case fallback BG2 Z=52", flush=True)
    code = nr.Code(2, 52)
    T = 10
    w = code.plain_weights(T)
    w[..., 0] = 0.0
    code.set_weights(w)
    llr, cw = _noisy_llr(nr, code, 48, 1.0, 0.2, seed=3)
    lhd, lapp, lit = _learned(torch_gpu, code, llr, T)
    fhd, fapp, fit = _learned(torch_gpu, code, llr, T, flags=nr.FALLBACK_PLAIN)
    phd, papp, pit = _gpu_decode(torch_gpu, code, llr, T)
    oc = oracle.code(KIND[2], 52)
    ls = np.array([oracle.syndrome(oc, h) for h in lhd])
    ps = np.array([oracle.syndrome(oc, h) for h in phd])
    take = (ls != 0) & (ps == 0)
    assert take.any() and (~take).any(), (ls, ps)
    want_hd = np.where(take[:, None], phd, lhd)
    want_app = np.where(take[:, None], papp, lapp)
    assert np.array_equal(fhd, want_hd)
    assert np.array_equal(fapp.view(np.uint32), want_app.view(np.uint32))
    assert np.array_equal(fit, np.where(take, pit, lit))
