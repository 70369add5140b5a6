"""ctypes binding of the C oracle (liboracle.so built from fvm_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU
baseline leg of bench.py as the checker, never by the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcInput(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("istart", "iflow", "iterlast", "maxiter", "saveinterval", "iflux",
                 "ilimit", "vortex", "cell_type", "sa_compat")] + \
               [(n, ctypes.c_double) for n in
                ("mach_inf", "aoa_deg", "rey", "cfl", "minres", "xref", "yref", "nut_factor")] + \
               [("hllc_rotate", ctypes.c_int)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp,
                                 ctypes.POINTER(OrcInput), ctypes.POINTER(ctypes.c_int)]
        L.orc_destroy.argtypes = [vp]
        L.orc_dims.argtypes = [vp, vp]
        L.orc_get.restype = ctypes.c_long
        L.orc_get.argtypes = [vp, ctypes.c_char_p, vp]
        L.orc_set.argtypes = [vp, ctypes.c_char_p, vp]
        L.orc_scalar.restype = ctypes.c_double
        L.orc_scalar.argtypes = [vp, ctypes.c_char_p]
        for f in ("orc_residu", "orc_check_positivity", "orc_conres"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = ctypes.c_int
        for f in ("orc_save_old", "orc_time_step", "orc_sa_model", "orc_clcd"):
            getattr(L, f).argtypes = [vp]
        L.orc_update.argtypes = [vp, ctypes.c_int]
        L.orc_iterate.restype = ctypes.c_int
        L.orc_iterate.argtypes = [vp, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_int)]
        L.orc_error_index.argtypes = [vp]
        L.orc_flux.restype = ctypes.c_int
        L.orc_flux.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        _lib = L
    return _lib


INT_FIELDS = {"elem", "ptype", "spts", "edge", "edneigh", "iedge", "beindx"}
DIMS = ("np", "nt", "ne", "netot", "nsp", "nfp", "nbe", "nswe1", "nswe2", "nffe1", "nffe2")


def make_input(**kw) -> OrcInput:
    d = dict(istart=1, iflow=1, iterlast=0, maxiter=1000, saveinterval=10**9, iflux=1,
             ilimit=1, vortex=0, cell_type=1, sa_compat=1, mach_inf=0.5, aoa_deg=2.0,
             rey=0.0, cfl=1.0, minres=0.0, xref=0.25, yref=0.0, nut_factor=0.1, hllc_rotate=0)
    d.update(kw)
    return OrcInput(**d)


class Oracle:
    """One oracle solver instance on a grid (coord (np,2), elem (nt,3) 1-based)."""

    def __init__(self, coord, elem, ptype, **inp):
        self.L = lib()
        self.inp = make_input(**inp)
        coord = np.ascontiguousarray(coord, dtype=np.float64)
        elem = np.ascontiguousarray(elem, dtype=np.int32)
        ptype = np.ascontiguousarray(ptype, dtype=np.int32)
        err = ctypes.c_int(0)
        self.h = self.L.orc_create(coord.shape[0], elem.shape[0], coord.ctypes.data,
                                   elem.ctypes.data, ptype.ctypes.data,
                                   ctypes.byref(self.inp), ctypes.byref(err))
        if not self.h:
            raise RuntimeError(f"oracle setup failed: code {err.value}")
        d = np.zeros(11, dtype=np.int32)
        self.L.orc_dims(self.h, d.ctypes.data)
        self.dims = dict(zip(DIMS, (int(v) for v in d)))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def _count(self, name):
        P, T, E = self.dims["np"], self.dims["nt"], self.dims["netot"]
        return {"coord": 2 * P, "elem": 3 * T, "ptype": P, "spts": self.dims["nsp"],
                "edge": 2 * E, "edneigh": 2 * E, "iedge": self.dims["ne"],
                "beindx": self.dims["nbe"], "ds": 2 * E, "dsb": 2 * P, "elarea": T,
                "cvarea": P, "mcarea": P, "drmin": P, "wd": P, "wd1": self.dims["nsp"],
                "tcoord": 2 * T, "prim": 4 * P, "prim_old": 4 * P, "nut": P, "qx": 4 * P,
                "qy": 4 * P, "mu": P, "mul": P, "res": 4 * P, "dt": P, "divf": P}[name]

    def get(self, name):
        n = self._count(name)
        a = np.zeros(max(n, 1), dtype=np.int32 if name in INT_FIELDS else np.float64)
        got = self.L.orc_get(self.h, name.encode(), a.ctypes.data)
        if got < 0:
            raise KeyError(name)
        return a[:n]

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.int32 if name in INT_FIELDS else np.float64).ravel()
        assert a.size == self._count(name), (name, a.size)
        self.L.orc_set(self.h, name.encode(), a.ctypes.data)

    def scalar(self, name):
        return self.L.orc_scalar(self.h, name.encode())

    def residu(self):
        rc = self.L.orc_residu(self.h)
        if rc != 0:
            raise RuntimeError(f"residu failed: {rc} at {self.L.orc_error_index(self.h)}")
        return self.get("res")

    def iterate(self, n):
        hist = np.zeros((max(n, 1), 6))
        st = ctypes.c_int(0)
        done = self.L.orc_iterate(self.h, n, hist.ctypes.data, ctypes.byref(st))
        return hist[:done], st.value

    def error_index(self):
        return self.L.orc_error_index(self.h)

    def flux(self, iflux, ds, pl, pr):
        """One numerical flux (orc_flux) on primitive states pl, pr across ds."""
        ds = np.ascontiguousarray(ds, dtype=np.float64)
        pl = np.ascontiguousarray(pl, dtype=np.float64)
        pr = np.ascontiguousarray(pr, dtype=np.float64)
        f = np.zeros(4)
        rc = self.L.orc_flux(self.h, iflux, ds.ctypes.data, pl.ctypes.data, pr.ctypes.data, f.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"flux failed: {rc}")
        return f
