"""Host-side lifetimes of the library under AddressSanitizer, on the CPU.

The library's host code (fvm2d_gpu.hip's host pass, geometry.cpp,
partition.cpp) is built with AddressSanitizer against a CPU stand-in for the
HIP runtime (tests/host_asan/hip_stand_in.cpp: device memory is host memory,
kernels do nothing, asynchronous copies run only when their stream is
synchronised).  tests/host_asan/drive_host.cpp then runs bench.py's call
sequences on the same meshes: the 513x129 leg (iterate, kernel and stage
timing, state and residual calls, variant switches), the 8-angle batches,
the 48 requests created, iterated and destroyed one after another and then
batched by flow kind, a serving engine's join/leave pattern, error paths and
destroyed handles.  Any use of freed or out-of-scope memory, a leak, a
stream, allocation or queued operation left behind fails the test.  A
planted mistake (an asynchronous copy from a buffer whose scope has ended)
checks that the build reports what it is meant to.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fvm2d_amd", "csrc")
HERE = os.path.join(ROOT, "tests", "host_asan")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"
SAN = ["-fsanitize=address", "-fno-omit-frame-pointer", "-g"]
ASAN_ENV = "detect_stack_use_after_return=1:detect_leaks=1"

pytestmark = pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)),
                                reason="needs hipcc and ROCm's clang++ (host build only, no GPU)")


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout[-3000:] + r.stderr[-3000:]
    return r


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("asan")
    objs = []
    for src in ("geometry.cpp", "partition.cpp"):
        o = str(d / src.replace(".cpp", ".o"))
        _run([CLANG, "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", *SAN, "-c",
              os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    # host pass only: the kernels' host stubs call hipLaunchKernel
    o = str(d / "fvm2d_gpu.o")
    _run([HIPCC, "-O1", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "--offload-host-only",
          "-ffp-contract=off", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
          "-fno-omit-frame-pointer", "-g", "-c", os.path.join(CSRC, "fvm2d_gpu.hip"), "-o", o])
    objs.append(o)
    # the host pass refers to the device code object by a symbol the device
    # pass would define; the stand-in registers nothing, so any bytes do
    nm = _run(["nm", "-u", o]).stdout
    fat = re.findall(r"\b(__hip_fatbin_[0-9a-f]+)\b", nm)
    fb = d / "fatbin.cpp"
    fb.write_text("".join(f'extern "C" const unsigned char {n}[64] = {{0}};\n' for n in fat))
    for src, extra in ((os.path.join(HERE, "hip_stand_in.cpp"), []), (str(fb), []),
                       (os.path.join(HERE, "drive_host.cpp"), ["-I", os.path.join(ROOT, "include")])):
        o = str(d / (os.path.basename(src) + ".o"))
        _run([CLANG, "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", *extra,
              *SAN, "-c", src, "-o", o])
        objs.append(o)
    exe = str(d / "drive_host")
    _run([CLANG, *SAN, "-o", exe, *objs, "-pthread"])
    return exe


def _write(path, g, inp):
    par = inp.params()
    with open(path, "wb") as f:
        coord = np.ascontiguousarray(g.coord, np.float64)
        elem = np.ascontiguousarray(g.elem, np.int32)
        ptype = np.ascontiguousarray(g.ptype, np.int32)
        f.write(np.array([coord.shape[0], elem.shape[0]], np.int32).tobytes())
        f.write(bytes(par))
        f.write(coord.tobytes())
        f.write(elem.tobytes())
        f.write(ptype.tobytes())


@pytest.fixture(scope="module")
def meshes(tmp_path_factory):
    """bench.py's meshes: the 513x129 workload and its 48 served requests."""
    from fvm2d_amd import mesh, serving
    from fvm2d_amd.solver import Inputs
    d = tmp_path_factory.mktemp("meshes")
    kw = dict(iflow=1, mach_inf=0.8, aoa_deg=1.25, iflux=1, ilimit=1, cfl=1.0, maxiter=10**6)
    _write(d / "main_513x129.bin", mesh.naca0012_513x129(), Inputs.from_kwargs(**kw))
    cfg = serving.ServeConfig(max_iters=10**6, minres=0.0)
    for q in serving.workload(48, 2024, np_min=1e3, np_max=2e4):
        g = mesh.o_grid(q.ni, q.nj, m=q.m, p=q.p, t=q.t, first_spacing=cfg.first_spacing, seed=q.seed)
        _write(d / f"req_{q.rid:02d}.bin", g, serving.inputs_for(q, cfg))
    return str(d)


def test_bench_sequences_are_clean_under_asan(driver, meshes):
    env = dict(os.environ, ASAN_OPTIONS=ASAN_ENV)
    r = subprocess.run([driver, meshes, "1"], capture_output=True, text=True, env=env, timeout=600)
    err = r.stderr
    assert "AddressSanitizer" not in err and "LeakSanitizer" not in err, err[-4000:]
    assert r.returncode == 0, err[-4000:]
    m = re.search(r"(\d+) launches, (\d+) graph launches, (\d+) copies; (\d+) live streams, "
                  r"(\d+) live device allocations, (\d+) queued ops", err)
    assert m, err[-2000:]
    launches, graphs, copies, streams, allocs, queued = map(int, m.groups())
    assert launches > 1000 and graphs > 100 and copies > 1000
    assert (streams, allocs, queued) == (0, 0, 0)


def test_planted_use_after_scope_is_reported(driver):
    env = dict(os.environ, ASAN_OPTIONS=ASAN_ENV)
    r = subprocess.run([driver, "--plant-scope"], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode != 0
    assert "stack-use-after-scope" in r.stderr
