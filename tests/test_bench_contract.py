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


def test_viscous_and_sa_byte_models():
    # per stage, viscous flow adds 8 nt + 136 np; per iteration SA adds
    # 20 nt + 24 ne + 304 np (the SA step's passes)
    npt, nt, ne = 1050624, 2097152, 3147776
    e = 12 * nt + 24 * ne + 408 * npt
    assert bench.stage_bytes(1, npt, nt, ne) == e
    assert bench.stage_bytes(2, npt, nt, ne) == e + 8 * nt + 136 * npt
    lam = 3 * (e + 8 * nt + 136 * npt) + 176 * npt + 8 * npt
    assert bench.iteration_bytes(2, npt, nt, ne) == lam
    assert bench.iteration_bytes(3, npt, nt, ne) == lam + 20 * nt + 24 * ne + 304 * npt
    # about 2.57 kB per vertex-iteration for SA on this grid
    assert 2560 < bench.iteration_bytes(3, npt, nt, ne) / npt < 2580
