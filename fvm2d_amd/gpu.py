"""ctypes binding of libfvm2d_gpu.so (include/fvm2d_gpu.h).

The product path: every call goes to the HIP library.  There is no CPU
fallback; if the library is missing, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libfvm2d_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "fvm2d_gpu.h")

FVM_OK, FVM_CONVERGED, FVM_ITER_LIMIT = 0, 1, 2
FVM_E_NEG_SOUND, FVM_E_NONPOSITIVE, FVM_E_ZERO_RES1 = -4, -5, -6

vp = ctypes.c_void_p
ci = ctypes.c_int
cd = ctypes.c_double


class Params(ctypes.Structure):
    """fvm_params: values of the reference's COMMON blocks + options."""
    _fields_ = [("iflow", ci), ("iflux", ci), ("ilimit", ci), ("vortex", ci),
                ("cell_type", ci), ("iterlast", ci), ("maxiter", ci), ("saveinterval", ci),
                ("mach_inf", cd), ("aoa_deg", cd), ("rey", cd), ("cfl", cd), ("minres", cd),
                ("xref", cd), ("yref", cd),
                ("sa_compat", ci), ("hllc_rotate", ci), ("precision", ci), ("patch_size", ci),
                ("nut_factor", cd), ("block_threads", ci), ("reserved_params", ci)]


class Mesh(ctypes.Structure):
    """fvm_mesh: host pointers to the arrays geometric() leaves."""
    _fields_ = [("np", ci), ("nt", ci), ("ne", ci), ("netot", ci), ("nsp", ci),
                ("nswe1", ci), ("nswe2", ci), ("nffe1", ci), ("nffe2", ci), ("reserved", ci),
                ("coord", vp), ("elem", vp), ("edge", vp), ("iedge", vp), ("beindx", vp),
                ("ptype", vp), ("spts", vp), ("ds", vp), ("elarea", vp), ("cvarea", vp),
                ("mcarea", vp), ("drmin", vp), ("wd", vp)]


class FvmError(RuntimeError):
    def __init__(self, code: int, index: int = 0, what: str = ""):
        self.code = code
        self.index = index
        msg = status_string(code)
        if index:
            msg += f" (index {index})"
        if what:
            msg = f"{what}: {msg}"
        super().__init__(msg)


_lib = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m fvm2d_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.fvm_geometric.restype = ci
    L.fvm_geometric.argtypes = [ci, ci, vp, vp, vp, ci, ci, ci, ci] + [vp] * 14
    L.fvm_gpu_create.restype = ci
    L.fvm_gpu_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(Mesh), ctypes.POINTER(vp)]
    L.fvm_gpu_destroy.restype = ci
    L.fvm_gpu_destroy.argtypes = [vp]
    L.fvm_gpu_set_state.restype = ci
    L.fvm_gpu_set_state.argtypes = [vp, vp, vp]
    L.fvm_gpu_wall_dist.restype = ci
    L.fvm_gpu_wall_dist.argtypes = [ci, vp, vp, ci, vp, ci, vp, vp]
    L.fvm_gpu_wall_dist_noalloc.restype = ci
    L.fvm_gpu_wall_dist_noalloc.argtypes = [ci, vp, vp, ci, vp, ci, vp, vp]
    L.fvm_gpu_tune.restype = ci
    L.fvm_gpu_tune.argtypes = [vp, ci, ci, ci]
    L.fvm_gpu_set_cl.restype = ci
    L.fvm_gpu_set_cl.argtypes = [vp, cd]
    L.fvm_gpu_get_state.restype = ci
    L.fvm_gpu_get_state.argtypes = [vp, vp, vp, vp, vp, vp]
    L.fvm_gpu_residu.restype = ci
    L.fvm_gpu_residu.argtypes = [vp, vp, vp, vp]
    L.fvm_gpu_iterate.restype = ci
    L.fvm_gpu_iterate.argtypes = [vp, ci, vp, ctypes.POINTER(ci), ctypes.POINTER(ci)]
    L.fvm_gpu_time_kernels.restype = ci
    L.fvm_gpu_time_kernels.argtypes = [vp, ci, vp, ctypes.POINTER(ctypes.c_char_p), ci]
    L.fvm_gpu_time_stage.restype = ci
    L.fvm_gpu_time_stage.argtypes = [vp, ci, ctypes.POINTER(ctypes.c_double)]
    L.fvm_gpu_stream_copy.restype = ci
    L.fvm_gpu_stream_copy.argtypes = [ctypes.c_longlong, ci, ctypes.POINTER(ctypes.c_double)]
    L.fvm_patch_plan.restype = ci
    L.fvm_patch_plan.argtypes = [ctypes.POINTER(Mesh), ci, ci, ci, ci, vp]
    L.fvm_gpu_layout_info.argtypes = [vp, vp]
    L.fvm_gpu_workgroup_threads.restype = ci
    L.fvm_gpu_workgroup_threads.argtypes = [vp]
    L.fvm_gpu_monitor.restype = ci
    L.fvm_gpu_monitor.argtypes = [vp, vp]
    L.fvm_batch_create.restype = ci
    L.fvm_batch_create.argtypes = [ctypes.POINTER(vp)]
    for f in ("fvm_batch_add", "fvm_batch_remove"):
        getattr(L, f).restype = ci
        getattr(L, f).argtypes = [vp, vp]
    L.fvm_batch_size.restype = ci
    L.fvm_batch_size.argtypes = [vp]
    L.fvm_batch_iterate.restype = ci
    L.fvm_batch_iterate.argtypes = [vp, ci, vp]
    L.fvm_batch_launch.restype = ci
    L.fvm_batch_launch.argtypes = [vp, ci]
    L.fvm_batch_query.restype = ci
    L.fvm_batch_query.argtypes = [vp]
    L.fvm_batch_wait.restype = ci
    L.fvm_batch_wait.argtypes = [vp, vp]
    L.fvm_batch_destroy.restype = ci
    L.fvm_batch_destroy.argtypes = [vp]
    L.fvm_gpu_footprint.restype = ctypes.c_longlong
    L.fvm_gpu_footprint.argtypes = [vp]
    L.fvm_gpu_footprint_estimate.restype = ctypes.c_longlong
    L.fvm_gpu_footprint_estimate.argtypes = [ci, ci, ci, ci, ci]
    L.fvm_status_string.restype = ctypes.c_char_p
    L.fvm_status_string.argtypes = [ci]
    _lib = L
    return L


def status_string(code: int) -> str:
    try:
        return lib().fvm_status_string(code).decode()
    except Exception:  # noqa: BLE001 - message only
        return f"status {code}"


def declared_symbols(header: str = HEADER) -> list[str]:
    """Function names declared in include/fvm2d_gpu.h."""
    import re
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fvm_\w+)\s*\(", txt)))
