"""The persistent mixed-batch decoder does not depend on which workgroup decodes which tile.

Each workgroup of the persistent grid pulls tiles from a device counter, so the grid size decides
which tile follows which in a workgroup's LDS and device slots.  A decoder that read anything a
previous tile left behind (stale LDS words, a flag not cleared, a degree-1 slot not reloaded) or
raced between wavefronts would change its output with the grid.  Here the same batch (every
lifting size of both base graphs, mixed rates and channel qualities, so tiles converge at
different iterations and lanes of finished codewords step aside) is decoded with persistent
grids of 97, 160 and the full device in both LDS layouts, with and without early exit: every
hard decision, posterior and iteration count must be identical.  Mixed batches have no reference
counterpart beyond its one-code decode loop (src/gpu_decoder_top.cu:16-52)."""
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
