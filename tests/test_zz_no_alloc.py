"""fvm_gpu_wall_dist_noalloc (device pointers, a static device scratch, no
allocation) against fvm_gpu_wall_dist: the same wd bit for bit, on two calls
in a row (the second on a smaller mesh, so the scratch holds the first call's
list), and a clear error past the scratch's 65536 solid edges."""
import ctypes
import os

import numpy as np
import pytest

from fvm2d_amd import gpu, mesh
from fvm2d_amd.solver import geometric

pytestmark = pytest.mark.gpu

FVM_E_SIZE = -8


class DeviceArrays:
    """Device copies of host arrays through the HIP runtime the library
    itself links (/opt/rocm's libamdhip64), so that one runtime serves the
    process."""

    def __init__(self):
        self.hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so.7"))
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 8))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes), 1) == 0
        return p

    def get(self, p, a):
        assert self.hip.hipMemcpy(ctypes.c_void_p(a.ctypes.data), p, ctypes.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def noalloc(geo, compat):
    L = gpu.lib()                        # loads the library (and its runtime) first
    d = geo.dims
    dv = DeviceArrays()
    try:
        coord = dv.put(np.asarray(geo.coord, dtype=np.float64))
        ptype = dv.put(np.asarray(geo.ptype, dtype=np.int32))
        edge = dv.put(np.asarray(geo.edge, dtype=np.int32))
        wd_d = dv.put(np.full(d["np"], -1.0))
        rc = L.fvm_gpu_wall_dist_noalloc(d["np"], coord, ptype, d["netot"], edge, compat, wd_d, None)
        wd = dv.get(wd_d, np.zeros(d["np"]))
    finally:
        dv.free()
    return rc, wd


@pytest.mark.parametrize("compat", [0, 1])
def test_noalloc_wall_distance_equals_allocating_one(compat):
    grids = [mesh.o_grid(257, 65, first_spacing=1e-4, jitter=0.1, seed=3),
             mesh.o_grid(65, 17, first_spacing=2e-3, seed=4)]
    for g in grids:
        ref = geometric(g.coord, g.elem, g.ptype, iflow=3, sa_compat=compat, wall_distance="gpu")
        rc, wd = noalloc(ref, compat)
        assert rc == 0
        np.testing.assert_array_equal(wd, ref.wd)


def test_noalloc_rejects_more_solid_edges_than_its_scratch():
    g = mesh.o_grid(65538, 3, first_spacing=1e-3, seed=1)      # 65537 solid edges
    geo = geometric(g.coord, g.elem, g.ptype, iflow=1)
    rc, _ = noalloc(geo, 0)
    assert rc == FVM_E_SIZE
    # and the next call on a mesh that fits is whole again
    g = mesh.o_grid(65, 17, first_spacing=2e-3, seed=4)
    ref = geometric(g.coord, g.elem, g.ptype, iflow=3, sa_compat=0, wall_distance="gpu")
    rc, wd = noalloc(ref, 0)
    assert rc == 0
    np.testing.assert_array_equal(wd, ref.wd)
