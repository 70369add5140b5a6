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


def test_struct_layouts_match_header(tmp_path):
    """The ctypes structures against the C compiler's layout of the header:
    size and the offset of every field."""
    import subprocess
    fields = {"fvm_params": gpu.Params, "fvm_mesh": gpu.Mesh}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{gpu.HEADER}"', "int main(void) {"]
    for cname, ct in fields.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for f, _ in ct._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    (tmp_path / "layout.c").write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-o", str(tmp_path / "layout"), str(tmp_path / "layout.c")], check=True)
    got = [int(v) for v in subprocess.run([str(tmp_path / "layout")], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = []
    for ct in fields.values():
        want.append(ctypes.sizeof(ct))
        want += [getattr(ct, f).offset for f, _ in ct._fields_]
    assert got == want


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


def test_512_thread_workgroups_only_for_limited_euler():
    """fvm_gpu_create refuses block_threads other than 0/256/512, and 512 for
    kinds without that instance, before touching the device."""
    from fvm2d_amd import mesh
    from fvm2d_amd.gpu import FvmError
    from fvm2d_amd.solver import Inputs, Solver
    g = mesh.o_grid(33, 9, seed=1)
    for kw in (dict(block_threads=384), dict(block_threads=512, iflow=2, rey=100.0),
               dict(block_threads=512, ilimit=0), dict(block_threads=512, iflux=2, ilimit=0)):
        with pytest.raises(FvmError):
            Solver(g, Inputs.from_kwargs(**kw))
