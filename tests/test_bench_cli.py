"""bench.py's contract pieces that run without a GPU: the one-GPU refusal and
the CPU baseline record (the oracle on the host's cores and on one thread)."""
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
