"""The GPU trainer of the learned decoder (BG2 Z=208, T=10): its Adam update equals a numpy Adam
step on the same gradient, and a short training run from plain weights lowers the loss and the
frame error rate against plain min-sum (the learned rule replaces the fixed 0.75 of
src/ldpc_kernels.cu:164,295)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
T = 10


def test_first_adam_step_matches_numpy(nr, torch_gpu):
    torch = torch_gpu
    print("case trainer Adam step BG2 Z=208", flush=True)
    code = nr.Code(2, 208)
    w0 = code.plain_weights(T)
    code.set_weights(w0)
    B = 256
    rng = np.random.default_rng(3)
    sig = np.array([nr.noise_stddev(float(e), 0.2) for e in rng.uniform(1.0, 2.5, B)], dtype=np.float32)
    d_sig = torch.from_numpy(sig).cuda()
    llr = torch.empty((B, code.n_vn), dtype=torch.float32, device="cuda")
    bits = torch.empty((B, code.n_vn), dtype=torch.uint8, device="cuda")
    code.channel(7, 0, B, llr.data_ptr(), sigma_ptr=d_sig.data_ptr(), cw_bits_ptr=bits.data_ptr())
    shape = (T, code.nnz_full, 2)
    grad = torch.zeros(shape, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    ws = torch.empty(code.workspace_bytes(B, T), dtype=torch.uint8, device="cuda")
    code.train_grad(llr.data_ptr(), bits.data_ptr(), B, T, grad.data_ptr(), loss.data_ptr(), ws.data_ptr(), ws.numel())
    mom = torch.zeros(shape, dtype=torch.float32, device="cuda")
    vel = torch.zeros(shape, dtype=torch.float32, device="cuda")
    lr = 3e-3
    code.adam(grad.data_ptr(), mom.data_ptr(), vel.data_ptr(), 1, lr)
    torch.cuda.synchronize()
    g = grad.cpu().numpy()
    assert np.isfinite(g).all() and (g != 0).sum() > g.size // 2
    assert 0 < float(loss.item()) < 1
    # Adam (Kingma and Ba) in float32, in the kernel's operation order, from zero moments
    f = np.float32
    b1, b2, eps = f(0.9), f(0.999), f(1e-8)
    m = b1 * f(0) + (f(1) - b1) * g
    v = b2 * f(0) + (f(1) - b2) * g * g
    mh = m / (f(1) - f(0.9))
    vh = v / (f(1) - f(0.999))
    want = (w0 - f(lr) * mh / (np.sqrt(vh) + eps)).astype(f)
    assert np.array_equal(mom.cpu().numpy(), m) and np.array_equal(vel.cpu().numpy(), v)
    got = code.get_weights()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # a first Adam step moves each weight by lr * |g| / (|g| + eps'): about lr where |g| >> eps
    moved = np.abs(got - w0)[np.abs(g) > 1e-6]
    assert moved.size > 0 and np.all(moved > 0.9 * lr) and np.all(moved < 1.01 * lr)


def test_short_training_lowers_loss_and_fer(nr, torch_gpu):
    """32 Adam steps of 1024 codewords at Eb/N0 U[1.5, 2.5] dB from plain weights: the loss
    falls, and at 1.75 dB (200k fresh codewords each) the trained decoder's FER is below half of
    plain min-sum's at the same iterations."""
    from nr_gpu_ldpc_amd import noms
    print("case trainer short schedule BG2 Z=208", flush=True)
    code, w, hist, _ = noms.train(2, 208, T, 1 << 15, 1024, 3e-3, (1.5, 2.5), seed=3, log=lambda s: None)
    assert w.shape == (T, code.nnz_full, 2) and np.isfinite(w).all()
    first, last = hist[0][1], hist[-1][1]
    assert last < 0.7 * first, hist
    learned = noms.fer_curve(code, T, [1.75], 200_000, learned=True)[0]
    code.set_weights(code.plain_weights(T))
    plain = noms.fer_curve(code, T, [1.75], 200_000)[0]
    print(plain, learned, flush=True)
    assert learned["fer"] * 2 < plain["fer"]
    assert learned["frame_errors"] > 100  # enough errors for the comparison to mean something
