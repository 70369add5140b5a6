"""Host-side mirror of fvm2d's driver (src/main.f) over the HIP library.

    g = mesh.naca0012_129x33()
    s = Solver(g, Inputs(mach_inf=0.5, aoa_deg=2.0))
    hist, status = s.iterate(1000)          # rows: iter fres iresi fresi cl cd

Names follow the reference: :func:`geometric` (src/geometric.f),
:meth:`Solver.residu` (src/residu.f), :meth:`Solver.iterate` (the loop of
src/main.f:42-76).  Arrays use the reference layout: prim (np, 4) rows are
(rho, u, v, p) per vertex, connectivity 1-based.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import gpu
from .gpu import FvmError, Mesh, Params

DIMS = ("np", "nt", "ne", "netot", "nsp", "nfp", "nbe", "nswe1", "nswe2", "nffe1", "nffe2")


@dataclasses.dataclass
class Inputs:
    """The input deck (src/read_input.f:20-66) plus options it does not hold."""
    iflow: int = 1            # FlowType: 1 Euler, 2 laminar, 3 SA-RANS
    mach_inf: float = 0.5
    aoa_deg: float = 2.0
    rey: float = 0.0
    cfl: float = 1.0
    iterlast: int = 0
    maxiter: int = 1000
    minres: float = 0.0
    saveinterval: int = 10**9
    iflux: int = 1            # 1 Roe, 2 KFVS, 3 HLLC
    ilimit: int = 1
    vortex: int = 0
    xref: float = 0.25
    yref: float = 0.0
    cell_type: int = 1        # 1 median, 2 Barth
    sa_compat: int = 1
    hllc_rotate: int = 0
    precision: int = 0        # 0 FP64, 1 split FP32 flux / FP64 accumulation
    nut_factor: float = 0.1
    istart: int = 1
    patch_size: int = 0       # vertices per patch of the stage kernel (0 = automatic)
    block_threads: int = 0    # workgroup size of the stage kernel (0 = 256; 512 for Euler/Roe/limited)

    def params(self) -> Params:
        return Params(iflow=self.iflow, iflux=self.iflux, ilimit=self.ilimit, vortex=self.vortex,
                      cell_type=self.cell_type, iterlast=self.iterlast, maxiter=self.maxiter,
                      saveinterval=self.saveinterval, mach_inf=self.mach_inf,
                      aoa_deg=self.aoa_deg, rey=self.rey, cfl=self.cfl, minres=self.minres,
                      xref=self.xref, yref=self.yref, sa_compat=self.sa_compat,
                      hllc_rotate=self.hllc_rotate, precision=self.precision,
                      patch_size=self.patch_size,
                      nut_factor=self.nut_factor, block_threads=self.block_threads)

    @classmethod
    def from_kwargs(cls, **kw) -> "Inputs":
        names = {f.name for f in dataclasses.fields(cls)}
        return cls(**{k: v for k, v in kw.items() if k in names})


@dataclasses.dataclass
class Geometry:
    """What geometric() leaves in COMMON /dims/ and /gdata/ and its outputs."""
    dims: dict
    coord: np.ndarray
    elem: np.ndarray
    ptype: np.ndarray
    edge: np.ndarray
    edneigh: np.ndarray
    iedge: np.ndarray
    ds: np.ndarray
    dsb: np.ndarray
    elarea: np.ndarray
    cvarea: np.ndarray
    mcarea: np.ndarray
    drmin: np.ndarray
    wd: np.ndarray
    wd1: np.ndarray
    beindx: np.ndarray
    spts: np.ndarray


def _p(a: np.ndarray):
    return a.ctypes.data


def geometric(coord, elem, ptype, cell_type: int = 1, iflow: int = 1,
              sa_compat: int = 1, wall_distance: str = "host") -> Geometry:
    """Dual-mesh preprocessing (src/geometric.f:2-63) through fvm_geometric.
    For SA-RANS (iflow 3) the wall distance comes from the host loop or, with
    ``wall_distance="gpu"``, from fvm_gpu_wall_dist (same values)."""
    if iflow == 3 and wall_distance == "gpu":
        geo = geometric(coord, elem, ptype, cell_type, 1, sa_compat)
        gpu_wall_distance(geo, sa_compat)
        return geo
    L = gpu.lib()
    coord = np.ascontiguousarray(coord, dtype=np.float64).reshape(-1, 2)
    elem = np.ascontiguousarray(elem, dtype=np.int32).reshape(-1, 3).copy()
    ptype = np.ascontiguousarray(ptype, dtype=np.int32).ravel()
    npt, nt = coord.shape[0], elem.shape[0]
    me = 3 * nt + 3
    edge = np.zeros((me, 2), np.int32)
    edneigh = np.zeros((me, 2), np.int32)
    iedge = np.zeros(me, np.int32)
    ds = np.zeros((me, 2))
    dsb = np.zeros((npt, 2))
    elarea = np.zeros(nt)
    cvarea = np.zeros(npt)
    mcarea = np.zeros(npt)
    drmin = np.zeros(npt)
    wd = np.zeros(npt)
    wd1 = np.zeros(npt)
    beindx = np.zeros(me, np.int32)
    spts = np.zeros(npt, np.int32)
    dims = np.zeros(12, np.int32)
    rc = L.fvm_geometric(npt, nt, _p(coord), _p(elem), _p(ptype), cell_type, iflow, sa_compat, me,
                         _p(edge), _p(edneigh), _p(iedge), _p(ds), _p(dsb), _p(elarea),
                         _p(cvarea), _p(mcarea), _p(drmin), _p(wd), _p(wd1), _p(beindx),
                         _p(spts), _p(dims))
    if rc != 0:
        raise FvmError(rc, int(dims[11]), "fvm_geometric")
    d = dict(zip(DIMS, (int(v) for v in dims[:11])))
    E = d["netot"]
    return Geometry(dims=d, coord=coord, elem=elem, ptype=ptype, edge=edge[:E].copy(),
                    edneigh=edneigh[:E].copy(), iedge=iedge[:d["ne"]].copy(), ds=ds[:E].copy(),
                    dsb=dsb, elarea=elarea, cvarea=cvarea, mcarea=mcarea, drmin=drmin, wd=wd,
                    wd1=wd1[:d["nsp"]].copy(), beindx=beindx[:d["nbe"]].copy(),
                    spts=spts[:d["nsp"]].copy())


def gpu_wall_distance(geo: Geometry, sa_compat: int = 1) -> None:
    """geo.wd, geo.wd1 from fvm_gpu_wall_dist (src/wall_dist.f)."""
    wd = np.zeros(geo.dims["np"])
    wd1 = np.zeros(max(geo.dims["nsp"], 1))
    rc = gpu.lib().fvm_gpu_wall_dist(geo.dims["np"], _p(geo.coord), _p(geo.ptype),
                                     geo.dims["netot"], _p(geo.edge), sa_compat, _p(wd), _p(wd1))
    if rc != 0:
        raise FvmError(rc, 0, "fvm_gpu_wall_dist")
    geo.wd = wd
    geo.wd1 = wd1[:geo.dims["nsp"]].copy()


def mesh_struct(g: Geometry, keep: list) -> Mesh:
    """fvm_mesh pointing into g's arrays (``keep`` holds references)."""
    arrs = [g.coord, g.elem, g.edge, g.iedge, g.beindx, g.ptype, g.spts, g.ds, g.elarea,
            g.cvarea, g.mcarea, g.drmin, g.wd]
    keep.extend(arrs)
    d = g.dims
    return Mesh(np=d["np"], nt=d["nt"], ne=d["ne"], netot=d["netot"], nsp=d["nsp"],
                nswe1=d["nswe1"], nswe2=d["nswe2"], nffe1=d["nffe1"], nffe2=d["nffe2"],
                reserved=0, coord=_p(g.coord), elem=_p(g.elem), edge=_p(g.edge),
                iedge=_p(g.iedge), beindx=_p(g.beindx), ptype=_p(g.ptype), spts=_p(g.spts),
                ds=_p(g.ds), elarea=_p(g.elarea), cvarea=_p(g.cvarea), mcarea=_p(g.mcarea),
                drmin=_p(g.drmin), wd=_p(g.wd))


def patch_plan(geo: Geometry, iflow: int = 1, patch_size: int = 0, lds_limit: int = 0,
               precision: int = 0) -> dict:
    """Layout of the LDS-tiled stage kernel for a mesh (host only)."""
    keep: list = []
    m = mesh_struct(geo, keep)
    st = np.zeros(8, dtype=np.int64)
    rc = gpu.lib().fvm_patch_plan(ctypes.byref(m), iflow, precision, patch_size, lds_limit, _p(st))
    names = ("patches", "max_patch", "max_local", "max_triangles", "max_edges", "lds_bytes",
             "halo_total", "failed_check")
    out = dict(zip(names, (int(v) for v in st)))
    if rc != 0:
        raise FvmError(rc, out["failed_check"], "fvm_patch_plan")
    return out


class Solver:
    """One steady solve resident on the GPU (one fvm_ctx)."""

    def __init__(self, grid, inputs: Inputs | None = None, geometry: Geometry | None = None):
        self.inputs = inputs or Inputs()
        self.L = gpu.lib()
        if geometry is None:
            geometry = geometric(grid.coord, grid.elem, grid.ptype, self.inputs.cell_type,
                                 self.inputs.iflow, self.inputs.sa_compat)
        self.geo = geometry
        self.np = geometry.dims["np"]
        keep: list = []
        m = mesh_struct(geometry, keep)
        par = self.inputs.params()
        h = ctypes.c_void_p()
        rc = self.L.fvm_gpu_create(ctypes.byref(par), ctypes.byref(m), ctypes.byref(h))
        if rc != 0:
            raise FvmError(rc, 0, "fvm_gpu_create")
        self.h = h
        self.status = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.fvm_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # state ------------------------------------------------------------------
    def set_state(self, prim, nut=None):
        prim = np.ascontiguousarray(prim, dtype=np.float64).reshape(-1)
        nutp = None
        if nut is not None:
            nut = np.ascontiguousarray(nut, dtype=np.float64).reshape(-1)
            nutp = _p(nut)
        rc = self.L.fvm_gpu_set_state(self.h, _p(prim), nutp)
        if rc != 0:
            raise FvmError(rc, 0, "fvm_gpu_set_state")
        self.status = 0

    def tune(self, variant: int = 2, vertices_per_block: int = 64, xcd_remap: int = 1):
        """Kernel variant and launch shape (see fvm_gpu_tune)."""
        rc = self.L.fvm_gpu_tune(self.h, variant, vertices_per_block, xcd_remap)
        if rc != 0:
            raise FvmError(rc, 0, "fvm_gpu_tune")

    def set_cl(self, cl: float):
        """COMMON /liftdrag/ cl, read by the farfield vortex correction."""
        rc = self.L.fvm_gpu_set_cl(self.h, float(cl))
        if rc != 0:
            raise FvmError(rc, 0, "fvm_gpu_set_cl")

    def get_state(self) -> dict:
        n = self.np
        prim = np.zeros(4 * n); nut = np.zeros(n); qx = np.zeros(4 * n); qy = np.zeros(4 * n)
        mu = np.zeros(n)
        rc = self.L.fvm_gpu_get_state(self.h, _p(prim), _p(nut), _p(qx), _p(qy), _p(mu))
        if rc != 0:
            raise FvmError(rc, 0, "fvm_gpu_get_state")
        return dict(prim=prim, nut=nut, qx=qx, qy=qy, mu=mu)

    # hot path ---------------------------------------------------------------
    def residu(self, prim, nut=None) -> np.ndarray:
        """res(4,np) of src/residu.f on the given state (flattened)."""
        prim = np.ascontiguousarray(prim, dtype=np.float64).reshape(-1)
        res = np.zeros(4 * self.np)
        nutp = None
        if nut is not None:
            nut = np.ascontiguousarray(nut, dtype=np.float64).reshape(-1)
            nutp = _p(nut)
        rc = self.L.fvm_gpu_residu(self.h, _p(prim), nutp, _p(res))
        if rc != 0:
            raise FvmError(rc, 0, "fvm_gpu_residu")
        return res

    def iterate(self, n: int):
        """Up to n iterations; returns (hist rows, status).  Raises FvmError
        on a reference STOP condition, with the rows done before it on
        ``err.hist``."""
        hist = np.zeros((max(n, 1), 6))
        nd = ctypes.c_int(0)
        ei = ctypes.c_int(0)
        rc = self.L.fvm_gpu_iterate(self.h, n, _p(hist), ctypes.byref(nd), ctypes.byref(ei))
        hist = hist[: nd.value]
        self.status = rc
        if rc < 0:
            err = FvmError(rc, ei.value, "fvm_gpu_iterate")
            err.hist = hist
            raise err
        return hist, rc

    def time_kernels(self, n: int) -> dict:
        ms = np.zeros(16)
        names = (ctypes.c_char_p * 16)()
        k = self.L.fvm_gpu_time_kernels(self.h, n, _p(ms), names, 16)
        return {names[i].decode(): float(ms[i]) for i in range(min(k, 16))}

    def time_stage(self, n: int = 50) -> float:
        """Milliseconds per launch of the patch stage kernel (first RK stage,
        n launches back to back between two HIP events)."""
        ms = ctypes.c_double(0.0)
        rc = self.L.fvm_gpu_time_stage(self.h, n, ctypes.byref(ms))
        if rc != 0:
            raise FvmError(rc, 0, "fvm_gpu_time_stage")
        return ms.value

    def footprint(self) -> int:
        return int(self.L.fvm_gpu_footprint(self.h))

    def monitor(self) -> tuple[dict, int]:
        """Monitors of the last iteration and the solve's status."""
        m = np.zeros(6)
        st = self.L.fvm_gpu_monitor(self.h, _p(m))
        return dict(iter=int(m[0]), fres=m[1], iresi=int(m[2]), fresi=m[3], cl=m[4], cd=m[5]), int(st)

    def layout_info(self) -> dict:
        """Launch layout of the stage kernel (patches, LDS, occupancy)."""
        a = np.zeros(8, dtype=np.int64)
        self.L.fvm_gpu_layout_info(self.h, _p(a))
        keys = ("patches", "max_patch", "lds_bytes", "workgroups_per_cu", "cus", "patch_size",
                "halo_total", "variant")
        out = {k: int(v) for k, v in zip(keys, a)}
        out["threads"] = int(self.L.fvm_gpu_workgroup_threads(self.h))
        return out


class Batch:
    """Solves of one kind advanced together, one launch per RK stage over all
    their patches (fvm_batch_*).  Members compute bit for bit what they
    would compute alone."""

    def __init__(self):
        self.L = gpu.lib()
        h = ctypes.c_void_p()
        rc = self.L.fvm_batch_create(ctypes.byref(h))
        if rc != 0:
            raise FvmError(rc, 0, "fvm_batch_create")
        self.h = h
        self.members: list[Solver] = []

    def add(self, s: Solver):
        rc = self.L.fvm_batch_add(self.h, s.h)
        if rc != 0:
            raise FvmError(rc, 0, "fvm_batch_add")
        self.members.append(s)

    def remove(self, s: Solver):
        rc = self.L.fvm_batch_remove(self.h, s.h)
        if rc != 0:
            raise FvmError(rc, 0, "fvm_batch_remove")
        self.members.remove(s)

    def iterate(self, n: int) -> list[int]:
        """Up to n iterations of every running member; statuses in member order."""
        st = np.zeros(max(len(self.members), 1), dtype=np.int32)
        rc = self.L.fvm_batch_iterate(self.h, n, _p(st))
        if rc != 0:
            raise FvmError(rc, 0, "fvm_batch_iterate")
        return [int(v) for v in st[: len(self.members)]]

    def launch(self, n: int):
        """Enqueue up to n (<= 4096) iterations and return at once."""
        rc = self.L.fvm_batch_launch(self.h, n)
        if rc != 0:
            raise FvmError(rc, 0, "fvm_batch_launch")

    def done(self) -> bool:
        """True when the launched iterations have finished (never blocks)."""
        rc = self.L.fvm_batch_query(self.h)
        if rc < 0:
            raise FvmError(rc, 0, "fvm_batch_query")
        return rc == 1

    def wait(self) -> list[int]:
        """Wait for the launched iterations; statuses in member order."""
        st = np.zeros(max(len(self.members), 1), dtype=np.int32)
        rc = self.L.fvm_batch_wait(self.h, _p(st))
        if rc != 0:
            raise FvmError(rc, 0, "fvm_batch_wait")
        return [int(v) for v in st[: len(self.members)]]

    def close(self):
        if getattr(self, "h", None):
            self.L.fvm_batch_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
