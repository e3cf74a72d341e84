"""Host side of bench.py, no GPU: the one-GPU refusal and the calibration table that the CPU
baseline carries (the flang-built reference's step time over the oracle's on the same state)."""
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
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_calibration_rows_for_every_timed_workload():
    import bench
    with open(os.path.join(ROOT, "oracle", "calibration.json")) as f:
        table = json.load(f)
    for key, config in bench.CALIBRATION.items():
        cal = bench.calibration(key)
        assert cal is not None, key
        assert cal["config"] == config and config in table
        ratios = cal["reference_s_over_port_s"]
        assert "1" in ratios and all(v > 0 for v in ratios.values())
        assert cal["all_cores_threads"] == max(int(k) for k in ratios)
        assert cal["all_cores_ratio"] == table[config][str(cal["all_cores_threads"])]["ratio"]
    assert bench.calibration("c5") is None
