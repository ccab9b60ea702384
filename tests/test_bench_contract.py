"""bench.py's contract on the CPU: --gpus above 1 is refused with a message,
and the committed traffic profile covers the default schedules of the
headline (so roofline.traffic is a measured figure, not null)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 2
    assert "one GPU" in res.stderr
    assert res.stdout == ""


@pytest.mark.parametrize("dtype_code,tune", [
    (0, {"rows_per_wave": 3, "waves": 16, "steps_per_pass": 2}),            # FP32 two-step pass
    (0, {"rows_per_wave": 4, "waves": 8, "vec": 2, "steps_per_pass": 1}),   # FP32 one-step
    (1, {"rows_per_wave": 2, "waves": 8, "vec": 2, "steps_per_pass": 0}),   # FP16 one-step
])
def test_traffic_profile_covers_default_schedules(dtype_code, tune):
    import bench
    traffic, src = bench.measured_traffic(1040, dtype_code, tune)
    assert traffic is not None and src.startswith("profiles/")
    # at least the bytes a launch must move per point: 20 B (two-step FP32
    # pass: read u, v, s, write two fields), 16 B (FP32) or 8 B (16-bit)
    floor = (20 if tune.get("steps_per_pass") == 2 else 16 if dtype_code == 0 else 8) * 1040 ** 3
    assert floor <= traffic < 2 * floor
