"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a message (the path is built for one GPU), and the on-chip roofline
figure follows from the plan's shape."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_onchip_coefficient_bytes_per_point():
    sys.path.insert(0, ROOT)
    import bench
    # 31 taps, tap-symmetric, one conjugate table, 4 shots per group: 16 taps x 8 B / 4
    r = bench.onchip_coefficients(15, 31, True, True, 4, 2147483648, 10.0)
    assert r["bytes_per_point"] == 32.0
    assert abs(r["achieved"] - 32.0 * 2147483648 / 1e-2 / 1e12) < 1e-9
    # no pairing, two tables, one shot per group: 31 x 8 x 2
    assert bench.onchip_coefficients(15, 31, False, False, 1, 1, 1.0)["bytes_per_point"] == 496.0
