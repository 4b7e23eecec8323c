"""bench.py on the CPU: more than one GPU is refused with a message, and the
workload table and traffic lookup behave (no GPU needed)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_kernel_table_and_traffic_lookup():
    sys.path.insert(0, ROOT)
    import bench
    assert len(bench.KERNELS) == 19          # one entry per kernel variant 0..18
    assert bench.KERNELS[12][0] == "stage_interior_phased"
    assert bench.KERNELS[16][0] == "stage_rows_phased"
    assert bench.KERNELS[18][0] == "stage_rows_lds"
    t = bench.measured_traffic("stage_interior_phased", "iso600x600x300")
    assert t is not None and 2.4e10 < t < 4.0e10     # above the 24.0 GB algorithmic bytes
    assert bench.measured_traffic("no_such_kernel", "iso600x600x300") is None
