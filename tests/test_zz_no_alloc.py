"""The allocation-free learned decode (ldpc_decode_learned_noalloc_async) against the allocating
one (ldpc_decode_learned_async) on the same inputs: hard decisions, APP and iteration counts bit
for bit, on two calls in a row, the second with fewer codewords (the fixed scratch still holds the
first call's contents), for BG2 Z=208 with the committed weights at 10 kept parity rows (rate 5/9,
where plain min-sum rescues some learned failures), plain fallback and early exit.
A batch beyond the fixed scratch is refused before anything runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
T = 10


def _code(nr):
    from nr_gpu_ldpc_amd import noms
    code = nr.Code(2, 208, 10)
    code.set_weights(noms.load_weights(noms.weights_path(2, 208, T))["weights"])
    return code


def _run(torch, code, fn, llr, n, flags):
    hd = torch.full((n, code.hd_bytes), 0xAB, dtype=torch.uint8, device="cuda")
    app = torch.full((n, code.n_vn), float("nan"), dtype=torch.float32, device="cuda")
    it = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    fn(llr.data_ptr(), n, T, hd.data_ptr(), app.data_ptr(), it.data_ptr(), flags=flags)
    torch.cuda.synchronize()
    return hd.cpu().numpy(), app.cpu().numpy(), it.cpu().numpy()


def test_noalloc_learned_decode_matches_allocating_one(nr, torch_gpu):
    torch = torch_gpu
    code = _code(nr)
    n_max = 600
    print(f"case no-alloc learned BG2 Z=208 m=10 n={n_max}, then 150", flush=True)
    llr = torch.empty((n_max, code.n_vn), dtype=torch.float32, device="cuda")
    eb = np.linspace(2.2, 2.8, n_max)
    rate = code.K / code.N
    sig = torch.from_numpy(np.array([nr.noise_stddev(float(e), rate) for e in eb], dtype=np.float32)).cuda()
    code.channel(21, 0, n_max, llr.data_ptr(), sigma_ptr=sig.data_ptr())
    torch.cuda.synchronize()
    flags = nr.FALLBACK_PLAIN | nr.EARLY_EXIT
    for n in (n_max, 150):
        want = _run(torch, code, code.decode_learned, llr, n, flags)
        got = _run(torch, code, code.decode_learned_noalloc, llr, n, flags)
        assert np.array_equal(got[0], want[0]), f"hard decisions, n={n}"
        assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32)), f"APP, n={n}"
        assert np.array_equal(got[2], want[2]), f"iterations, n={n}"
        # the fallback changed something: plain results replaced learned ones somewhere
        alone = _run(torch, code, code.decode_learned, llr, n, nr.EARLY_EXIT)
        assert not np.array_equal(alone[1], want[1]), f"fallback not exercised, n={n}"
    # without the fallback the call needs no temporaries
    got = _run(torch, code, code.decode_learned_noalloc, llr, 150, nr.EARLY_EXIT)
    want = _run(torch, code, code.decode_learned, llr, 150, nr.EARLY_EXIT)
    assert all(np.array_equal(g.view(np.uint8), w.view(np.uint8)) for g, w in zip(got, want))


def test_noalloc_refuses_a_batch_beyond_its_scratch(nr, torch_gpu):
    torch = torch_gpu
    code = _code(nr)
    # APP floats alone exceed the fixed array for this many codewords; every buffer is sized
    # for them, so nothing would run out of bounds even if the call went ahead
    n = nr.NOALLOC_SCRATCH_BYTES // (code.n_vn * 4) + 1
    print(f"case no-alloc refusal n={n}", flush=True)
    llr = torch.zeros((n, code.n_vn), dtype=torch.float32, device="cuda")
    hd = torch.zeros((n, code.hd_bytes), dtype=torch.uint8, device="cuda")
    app = torch.zeros((n, code.n_vn), dtype=torch.float32, device="cuda")
    with pytest.raises(nr.LdpcError, match="beyond the fixed") as e:
        code.decode_learned_noalloc(llr.data_ptr(), n, T, hd.data_ptr(), app.data_ptr(), 0,
                                    flags=nr.FALLBACK_PLAIN)
    assert e.value.status == -2
    torch.cuda.synchronize()
    assert int(hd.sum()) == 0  # nothing was enqueued
