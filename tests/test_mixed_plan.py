"""Host planner of the persistent mixed-batch decoder (no GPU: host-only planner object)."""
import numpy as np
import pytest

from nr_gpu_ldpc_amd import mixed as mx


def _planner(nr, max_cw=20000):
    return nr.Mixed(max_cw, device=-1)


def test_lifting_sizes_and_rate_policy(nr):
    assert mx.LIFTING_SIZES[:8] == [2, 3, 4, 5, 6, 7, 8, 9] and mx.LIFTING_SIZES[-1] == 384
    # BG1 reaches 1/3 .. 22/24, BG2 reaches 1/5 .. 10/12; outside rates clamp to the ends
    assert mx.rows_for_rate(nr.BG1, [0.2, 1 / 3, 0.5, 8 / 9]).tolist() == [46, 46, 24, 5]
    assert mx.rows_for_rate(nr.BG2, [0.2, 0.5, 8 / 9]).tolist() == [42, 12, 4]
    assert abs(mx.realised_rate(nr.BG1, 46) - 1 / 3) < 1e-12
    assert abs(mx.realised_rate(nr.BG2, 42) - 0.2) < 1e-12


def test_plan_covers_every_codeword_once(nr):
    b = mx.draw_batch(16384, seed=7)
    p = _planner(nr)
    p.plan(b["desc"])
    tiles, order = p.tiles()
    info = p.plan_info()
    assert info["n_tiles"] == len(tiles) and info["threads"] == 512
    assert sorted(order.tolist()) == list(range(16384))
    covered = np.zeros(16384, dtype=int)
    edges = {k: np.cumsum((nr.base_matrix(k, 0) >= 0).sum(axis=1)) for k in (1, 2)}
    costs = []
    for kind, Z, m, cpb, first, count in tiles.tolist():
        costs.append(((cpb * Z + 63) // 64 + 3) // 4 * edges[kind][m - 1])
        ids = order[first:first + count]
        covered[ids] += 1
        assert (b["kind"][ids] == kind).all() and (b["Z"][ids] == Z).all()
        assert b["m_rows"][ids].max() == m          # a tile runs with its largest row count
        assert 1 <= count == cpb and cpb * Z <= 512
    assert (covered == 1).all()
    # the queue runs longest tiles first (SIMD rounds x edges, non-increasing)
    assert all(a >= b for a, b in zip(costs, costs[1:]))
    assert len({(k, z) for k, z in zip(b["kind"].tolist(), b["Z"].tolist())}) > 90


def test_plan_orders_longest_first(nr):
    b = mx.describe([1, 1, 2, 1], [384, 8, 208, 384], [46, 46, 42, 4])
    p = _planner(nr)
    p.plan(b["desc"])
    tiles, order = p.tiles()
    assert tiles[0].tolist()[:3] == [1, 384, 46]
    assert tiles[-1].tolist()[:2] in ([1, 8], [1, 384])


def test_tile_size_follows_cpb(nr):
    b = mx.describe([2] * 10, [16] * 10, [42, 4, 30, 42, 7, 8, 9, 10, 11, 12])
    p = _planner(nr)
    p.set_cpb(2, 16, 3)
    p.plan(b["desc"])
    tiles, order = p.tiles()
    assert [t[5] for t in tiles.tolist()] == [3, 3, 3, 1]
    # rows sorted largest first inside the group
    assert b["m_rows"][order].tolist() == sorted(b["m_rows"].tolist(), reverse=True)
    assert p.get_cpb(2, 16) == 3
    p.set_cpb(2, 16, 0)
    assert p.get_cpb(2, 16) == 512 // 16


@pytest.mark.parametrize("field,value", [("kind", 3), ("Z", 17), ("m_rows", 3), ("m_rows", 47),
                                          ("llr_offset", -1), ("reserved", 1)])
def test_plan_rejects_bad_descriptors(nr, field, value):
    b = mx.describe([1, 1], [64, 64], [10, 46])
    d = b["desc"].copy()
    d[field][1] = value
    with pytest.raises(nr.LdpcError):
        _planner(nr).plan(d)


def test_plan_rejects_oversized_batch_and_cpb(nr):
    b = mx.describe([1] * 5, [64] * 5, [10] * 5)
    with pytest.raises(nr.LdpcError):
        _planner(nr, max_cw=4).plan(b["desc"])
    p = _planner(nr)
    with pytest.raises(nr.LdpcError):
        p.set_cpb(1, 384, 2)           # two BG1 Z=384 codewords exceed 512 lanes
    with pytest.raises(nr.LdpcError):
        p.decode(0, 10, 0)             # host-only object cannot decode


def test_tuning_file_roundtrip(nr, tmp_path):
    p = _planner(nr)
    p.set_cpb(1, 64, 5)
    p.set_cpb(2, 208, 1)
    path = tmp_path / "tune.txt"
    p.save_tuning(str(path))
    q = _planner(nr)
    q.load_tuning(str(path))
    assert q.get_cpb(1, 64) == 5 and q.get_cpb(2, 208) == 1
    path.write_text("1 64\n")
    with pytest.raises(nr.LdpcError):
        q.load_tuning(str(path))


def test_tuning_file_keeps_both_layouts(nr, tmp_path):
    """Codewords per workgroup are kept per layout; the tuning file holds both ("slim" lines)."""
    p = _planner(nr)
    p.set_cpb(1, 64, 5)
    p.set_layout(nr.LAYOUT_SLIM)
    assert p.plan_info()["threads"] == 768
    assert p.get_cpb(1, 64) == 768 // 64     # the slim layout's own default
    p.set_cpb(1, 64, 9)
    p.set_cpb(1, 384, 2)                     # two BG1 Z=384 codewords fit 768 lanes
    path = tmp_path / "tune.txt"
    p.save_tuning(str(path))
    assert "1 64 9 slim" in path.read_text() and "1 64 5\n" in path.read_text()
    q = _planner(nr)
    q.load_tuning(str(path))
    assert q.get_cpb(1, 64) == 5
    q.set_layout(nr.LAYOUT_SLIM)
    assert q.get_cpb(1, 64) == 9 and q.get_cpb(1, 384) == 2
    q.set_layout(nr.LAYOUT_AUTO)
    assert q.get_layout() == nr.LAYOUT_DUP and q.get_cpb(1, 64) == 5
    path.write_text("1 64 3 fast\n")
    with pytest.raises(nr.LdpcError):
        q.load_tuning(str(path))
    with pytest.raises(nr.LdpcError):
        q.set_layout(7)
