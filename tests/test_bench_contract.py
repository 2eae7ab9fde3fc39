"""bench.py contract pieces that run without a GPU."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_traffic_profile_is_calibrated_and_complete():
    import bench
    text = open(bench.TRAFFIC_PROFILE).read()
    doc = json.loads(text.split("```json", 1)[1].split("```", 1)[0])
    assert abs(doc["fetch_factor"] - 0.5) < 0.01 and abs(doc["write_factor"] - 1.0) < 0.01
    for kind in ("z_r2c", "y_fwd", "x_spectral", "y_inv", "z_c2r"):
        assert "0" in doc["passes"][kind]
        for row in doc["passes"][kind].values():
            # measured HBM traffic never undercuts the algorithmic bytes by more than rounding
            assert row["fetch_bytes"] + row["write_bytes"] >= 0.99 * row["alg_bytes"]
    t, info = bench.measured_traffic("x_spectral", 0, (2560, 2560, 1536))
    assert t == info["fetch_bytes"] + info["write_bytes"] and t > 1e11
    assert bench.measured_traffic("x_spectral", 0, (256, 256, 256)) == (None, None)
