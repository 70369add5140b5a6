"""bench.py's contract pieces that need no GPU: the byte model of the
roofline and the refusal of more than one GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_algorithmic_bytes_of_a_stage():
    # 12 nt + 24 ne + 408 np per RK stage (gradient + flux + update passes)
    npt, nt, ne = 66048, 131072, 197120
    assert bench.algorithmic_bytes("k_patch", npt, nt, ne) == 12 * nt + 24 * ne + 408 * npt
    assert bench.algorithmic_bytes("k_patch", npt, nt, ne) == 33251328
    assert bench.algorithmic_bytes("nothing", npt, nt, ne) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_traffic_summary_is_committed():
    t = bench.traffic_from_profiles("k_patch")
    assert isinstance(t, int) and 0 < t < 2 * 33251328
