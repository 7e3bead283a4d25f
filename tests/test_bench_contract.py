"""bench.py's contract pieces that need no GPU: one GPU only (a request for
more exits with a message), and the measured HBM traffic of the roofline
kernels is read from the committed rocprofv3 summaries."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_traffic_figures_parse_from_profiles():
    sys.path.insert(0, ROOT)
    import bench
    for cls, members in bench.TRAFFIC_PROFILES:
        t = bench.measured_traffic(cls, members)
        assert isinstance(t, int) and t > 10_000_000, (cls, members, t)
