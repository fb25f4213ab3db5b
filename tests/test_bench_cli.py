"""bench.py's command line and its reading of the committed HBM traffic tables (CPU only)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""  # no JSON line


def test_traffic_tables_are_read():
    import bench
    for workload in ("pileup_1000", "pileup_1000_retina"):
        b = bench.measured_traffic(workload, "search")
        # The search's measured bytes per launch sit above its algorithmic 35.3 MB on pileup_1000.
        assert b is not None and 35_000_000 < b < 80_000_000, (workload, b)
    assert bench.measured_traffic("pileup_1000", "no_such_stage") is None
    assert bench.measured_traffic("no_such_workload", "search") is None
