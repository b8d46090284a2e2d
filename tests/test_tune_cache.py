"""The JSON cache of launch cost models (nr_gpu_ldpc_amd.tune.save_cache / apply_cache) on host-only
contexts: a model written for one (kind, Z) is installed into a fresh context exactly, a code the
cache does not hold is left alone, a cache without a slim entry drops the slim layout, and a file
of another format is refused.  No reference counterpart: the reference fixes its launch grid at
build time (src/setup_exp.cu:131-166)."""
import json

import pytest


def test_cache_round_trip(nr, tmp_path):
    from nr_gpu_ldpc_amd import tune
    src = nr.Code(2, 13, 0, -1)
    src.set_layout_model(nr.LAYOUT_DUP, 3, 9, 88.5)
    src.set_layout_model(nr.LAYOUT_SLIM, 1, 4, 40.25)
    path = tmp_path / "models.json"
    tune.save_cache(str(path), {"2/13": tune.layout_models(src)}, "host")
    doc = json.loads(path.read_text())
    assert doc["codes"]["2/13"]["dup"] == {"cpb": 3, "cw_per_cu": 9, "round_us": 88.5}

    dst = nr.Code(2, 13, 0, -1)
    assert tune.apply_cache(dst, str(path))
    assert tune.layout_models(dst) == tune.layout_models(src)
    assert dst.launch_layout(10) == (nr.LAYOUT_DUP, 3, 4)

    other = nr.Code(2, 15, 0, -1)
    before = tune.layout_models(other)
    assert not tune.apply_cache(other, str(path))
    assert tune.layout_models(other) == before

    no_slim = {"2/13": {"dup": {"cpb": 2, "cw_per_cu": 6, "round_us": 70.0},
                        "slim": {"cpb": 0, "cw_per_cu": 0, "round_us": 0.0}}}
    tune.save_cache(str(path), no_slim)
    assert tune.apply_cache(dst, str(path))
    assert tune.layout_models(dst) == no_slim["2/13"]


def test_cache_of_another_format_is_refused(nr, tmp_path):
    from nr_gpu_ldpc_amd import tune
    path = tmp_path / "other.json"
    path.write_text(json.dumps({"codes": {}}))
    with pytest.raises(ValueError):
        tune.apply_cache(nr.Code(2, 13, 0, -1), str(path))
