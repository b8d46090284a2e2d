"""The persistent mixed-batch decoder does not depend on which workgroup decodes which tile.

Each workgroup of the persistent grid pulls tiles from a device counter, so the grid size decides
which tile follows which in a workgroup's LDS and device slots.  A decoder that read anything a
previous tile left behind (stale LDS words, a flag not cleared, a degree-1 slot not reloaded) or
raced between wavefronts would change its output with the grid.  Here the same batch (every
lifting size of both base graphs, mixed rates and channel qualities, so tiles converge at
different iterations and lanes of finished codewords step aside) is decoded with persistent
grids of 97, 160 and the full device in both LDS layouts, with and without early exit: every
hard decision, posterior and iteration count must be identical.  Also: special channel values
(signed zeros, exact ties, large magnitudes) in a mixed batch against the oracle in both layouts.
Mixed batches have no reference counterpart beyond its one-code decode loop
(src/gpu_decoder_top.cu:16-52)."""
import numpy as np
import pytest

from nr_gpu_ldpc_amd import mixed as mx
from tests.test_gpu_mixed import _mixed

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("early", [False, True])
def test_mixed_output_does_not_depend_on_the_grid(nr, torch_gpu, early):
    torch = torch_gpu
    n, T = 4096, 12
    b = mx.draw_batch(n, seed=31, ebn0=(0.0, 4.0))
    llr, _ = mx.make_channel(torch, b, seed=8)
    flags = nr.EARLY_EXIT if early else 0
    ref = None
    for lay_name, lay in (("dup", nr.LAYOUT_DUP), ("slim", nr.LAYOUT_SLIM)):
        for grid in (0, 97, 160):
            print(f"case mixed grid early={early} layout={lay_name} grid={grid}", flush=True)
            m = nr.Mixed(n)
            m.set_layout(lay)
            m.set_grid(grid)
            got = _mixed(nr, torch, b, llr, T, flags, m)
            m.close()
            if ref is None:
                ref = got
                continue
            assert np.array_equal(got[0], ref[0]), f"hard decisions ({lay_name}, grid {grid})"
            assert np.array_equal(got[1].view(np.uint32), ref[1].view(np.uint32)), f"APP ({lay_name}, grid {grid})"
            assert np.array_equal(got[2], ref[2]), f"iterations ({lay_name}, grid {grid})"
    if early:
        assert ref[2].min() < T and ref[2].max() == T


@pytest.mark.parametrize("layout_name", ["dup", "slim"])
def test_mixed_special_values_match_oracle(nr, oracle, torch_gpu, layout_name):
    """Signed zeros, exact ties and magnitudes above the 1000.0f start value in a mixed batch of
    both base graphs, compiled-in and per-edge-table lifting sizes and truncated parity rows: every
    codeword equals the oracle's decode of its own code (hard decisions, posteriors up to the sign
    of an exactly-zero posterior, iteration counts with early exit)."""
    from tests.oracle_binding import OR_BG1, OR_BG2
    torch = torch_gpu
    print(f"case mixed special values layout={layout_name}", flush=True)
    shapes = [(1, 384, 46), (1, 384, 7), (2, 208, 42), (2, 208, 9), (1, 13, 20), (2, 6, 42), (1, 52, 4), (2, 104, 15)]
    rng = np.random.default_rng(44)
    kinds, Zs, ms, rows = [], [], [], []
    for k, z, m in shapes:
        for v in range(5):
            n = (22 if k == 1 else 10) * z + m * z
            x = [np.full(n, 2000.0, np.float32),
                 np.where(rng.random(n) < 0.5, -1.0, 1.0).astype(np.float32),
                 np.where(rng.random(n) < 0.5, -0.0, 0.0).astype(np.float32),
                 (rng.standard_normal(n) * 3000).astype(np.float32),
                 np.round(rng.standard_normal(n) * 3).astype(np.float32)][v]
            x[:2 * z] = 0.0
            kinds.append(k); Zs.append(z); ms.append(m); rows.append(x)
    b = mx.describe(np.array(kinds), np.array(Zs), np.array(ms))
    llr = torch.from_numpy(np.concatenate(rows)).cuda()
    m_obj = nr.Mixed(len(rows))
    m_obj.set_layout(nr.LAYOUT_DUP if layout_name == "dup" else nr.LAYOUT_SLIM)
    hd, app, its = _mixed(nr, torch, b, llr, 8, nr.EARLY_EXIT, m_obj)
    m_obj.close()
    d = b["desc"]
    for i, x in enumerate(rows):
        oc = oracle.code(OR_BG1 if kinds[i] == 1 else OR_BG2, Zs[i], ms[i])
        ohd, oapp, oit = oracle.decode(oc, x[None, :], 8, early_exit=True)
        h0, nb = int(d["hd_offset"][i]), int(b["hd_bytes"][i])
        a0, nv = int(d["llr_offset"][i]), int(b["n_vn"][i])
        want = np.packbits(ohd[0], bitorder="little")
        assert np.array_equal(hd[h0:h0 + nb], want), (i, kinds[i], Zs[i], ms[i])
        g = app[a0:a0 + nv]
        assert np.array_equal(g, oapp[0]), (i, kinds[i], Zs[i], ms[i])
        nz = oapp[0] != 0
        assert np.array_equal(g.view(np.uint32)[nz], oapp[0].view(np.uint32)[nz])
        assert its[i] == oit[0], (i, its[i], oit[0])
