"""bench.py's command line on the CPU: more than one GPU is refused with a message before any GPU
work, the layout option takes only the known layouts, and the per-layout PMC summaries the
roofline reads exist and name their kernels."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_unknown_layout_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--layout", "wide"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2 and "--layout" in r.stderr


def test_roofline_summaries_exist_for_both_layouts():
    sys.path.insert(0, ROOT)
    import bench
    for layout, kernel in (("slim", "k_decode_slim<1, 384"), ("dup", "k_decode<1, 384")):
        for kind in ("traffic", "sq"):
            with open(bench.profile_file(layout, kind)) as f:
                doc = json.load(f)
            assert doc["kernel"].startswith(kernel), (layout, kind, doc["kernel"])
            assert doc["source"].startswith("profiles/") and os.path.exists(
                os.path.join(ROOT, doc["source"].split(" ")[0]))
        with open(bench.profile_file(layout, "traffic")) as f:
            t = json.load(f)
        assert t["bytes_per_launch"] == t["fetch_bytes"] + t["write_bytes"]
