"""The C-ABI library loads and exports every entry point include/fvm2d_gpu.h
declares (no GPU needed: nothing here launches a kernel)."""
import ctypes
import os

import pytest

from fvm2d_amd import gpu


def test_library_exports_declared_symbols():
    L = gpu.lib()
    names = gpu.declared_symbols()
    assert "fvm_gpu_iterate" in names and "fvm_geometric" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match_header():
    # fvm_params: 8 ints, 7 doubles, 4 ints, 1 double; fvm_mesh: 10 ints + 13 pointers
    assert ctypes.sizeof(gpu.Params) == 8 * 4 + 7 * 8 + 4 * 4 + 8
    assert ctypes.sizeof(gpu.Mesh) == 10 * 4 + 13 * 8


def test_status_strings():
    assert "negative" in gpu.status_string(gpu.FVM_E_NONPOSITIVE)
    assert "Sonic" in gpu.status_string(gpu.FVM_E_NEG_SOUND)


def test_footprint_estimate_scales():
    L = gpu.lib()
    a = L.fvm_gpu_footprint_estimate(66048, 131072, 197120, 1, 0)
    b = L.fvm_gpu_footprint_estimate(2 * 66048, 2 * 131072, 2 * 197120, 1, 0)
    assert 1.9 < b / a < 2.1
    assert 200 < a / 66048 < 2000


def test_host_program_built():
    """The Fortran host program links against the library and runs its
    argument check without a GPU."""
    import subprocess
    from fvm2d_amd import build
    if not os.path.exists(build.FLANG):
        pytest.skip("no flang")
    assert os.path.exists(build.HOST_EXE)
    r = subprocess.run([build.HOST_EXE], capture_output=True, text=True, timeout=60)
    assert "You must specify an input file." in r.stdout
