"""bench.py host logic that needs no GPU: the multi-GPU refusal and the
per-plan-shape HBM traffic lookup."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused_before_any_gpu_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout


def test_traffic_lookup_matches_the_plan_shape():
    import bench
    with open(os.path.join(ROOT, "profiles", "traffic_headline.json")) as f:
        plans = json.load(f)["plans"]
    assert plans, "no traffic measurement committed"
    for key, t in plans.items():
        k, s = (int(x) for x in key.split("x"))
        assert t["launches"] == s
        got = bench.traffic_per_launch({"tile_bits": k, "n_sweeps": s, "S": 2 * t["states"]})
        assert abs(got - 2 * t["bytes_per_launch"]) < 1e-6 * got
        # measured traffic within 2% of one read and one write pass per launch
        assert abs(t["bytes_per_launch"] / t["algorithmic_bytes_per_launch"] - 1) < 0.02
    assert bench.traffic_per_launch({"tile_bits": 9, "n_sweeps": 99, "S": 10}) is None
