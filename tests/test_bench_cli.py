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
    # Clustering's row as well (the line's clustering_hbm.traffic): at least the 29.2 MB it must move.
    c = bench.measured_traffic("pileup_1000", "clustering")
    assert c is not None and 29_000_000 < c < 80_000_000, c
    assert bench.measured_traffic("pileup_1000", "no_such_stage") is None
    assert bench.measured_traffic("no_such_workload", "search") is None


def test_search_instruction_count_is_read():
    import bench
    # Wave-level VALU instructions of one 1000-event search launch (roofline.search_alu_lds.valu_issue):
    # at least one per event and wave, and far below what one launch could issue in a millisecond.
    n = bench.measured_valu_instructions("pileup_1000")
    assert n is not None and 2_000 < n < 614_400_000, n
    assert bench.measured_valu_instructions("no_such_workload") is None
