"""The committed learned weights (BG2 Z=208, T=10) beat plain min-sum on the GPU.

Frame error rates come from fresh Philox codewords decoded on the GPU.  The full evaluation
(400k codewords per point, FER 1e-3 crossing 1.745 dB learned against 1.945 dB plain) is
stored beside the weights; this test repeats the decisive comparison at 120k codewords.
"""
import os

import pytest

pytestmark = pytest.mark.gpu


def test_learned_gain_bg2_z208(nr, torch_gpu):
    from nr_gpu_ldpc_amd import noms
    path = noms.weights_path(2, 208, 10)
    assert os.path.exists(path)
    rec = noms.load_weights(path)
    code = nr.Code(2, 208)
    code.set_weights(rec["weights"])
    n = 120_000
    print("case learned gain BG2 Z=208 T=10", flush=True)
    plain = noms.fer_curve(code, 10, [1.75, 1.875], n)
    learned = noms.fer_curve(code, 10, [1.75], n, learned=True)
    fallback = noms.fer_curve(code, 10, [1.75], n, learned=True, flags=nr.FALLBACK_PLAIN)
    print(plain, learned, fallback, flush=True)
    # expected about 1.5e-2 / 2.8e-3 (plain at 1.75 / 1.875 dB) against 9e-4 (learned at 1.75 dB)
    assert learned[0]["fer"] * 5 < plain[0]["fer"]
    assert learned[0]["fer"] * 1.5 < plain[1]["fer"]
    assert fallback[0]["fer"] <= learned[0]["fer"]
    assert learned[0]["syndrome_fail"] < plain[0]["syndrome_fail"]
