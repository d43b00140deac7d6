"""bench.py's contract pieces that run without a GPU: the one-GPU refusal,
the CPU baseline record (the oracle on the host's cores and on one thread)
and the roofline record built from the committed counter summaries."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode != 0
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_cpu_baseline_record():
    import bench
    out = bench.cpu_baseline(32, budget_s=0.2)
    assert out["unit"] == "cell-updates/s" and out["kind"] == "port"
    assert out["value"] > 0 and out["cores"] >= 1
    assert out["single_thread"]["cores"] == 1 and out["single_thread"]["value"] > 0
    assert "slab" in out["sample"]
    cal = out.get("reference_calibration")
    if cal is not None and "error" not in cal:
        assert cal["bit_identical"] is True


def test_roofline_record_from_committed_profiles():
    """roofline: achieved = 80 B x N^3 per launch / the dominant kernel's mean
    launch time; traffic and the VALU issue floor come from the committed
    counter summaries (profiles/traffic.json, profiles/valu.json) of the same
    512^3 workload, for both precisions."""
    import bench
    r = {"ms": [300 * 8.0, 300 * 3.5, 300 * 3.4], "launches": [300, 300, 300]}
    for prec in ("fp64", "mixed"):
        out = bench.roofline_of(r, 512, prec)
        assert out["kernel"] == "sweep_lane_x" and out["bound"] == "hbm"
        assert abs(out["achieved"] - 80 * 512 ** 3 / 8.0e-3 / 1e9) < 0.1
        assert abs(out["frac"] - out["achieved"] / 8000.0) < 1e-4
        assert 1.0 < out["traffic"] / out["alg_bytes_per_launch"] < 1.3
        v = out["valu_issue"]
        assert abs(v["issue_bound_ms"] - v["insts_per_launch"] * 4 / (1024 * 2.4e9) * 1e3) < 1e-3
        assert abs(v["frac"] - v["issue_bound_ms"] / 8.0) < 1e-3
        assert 0.5 < v["frac"] <= 1.0
