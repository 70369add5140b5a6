"""Serving loop: steady solves of different meshes, Mach numbers and angles of
attack, admitted by HBM footprint and advanced by continuous batching.

A request is a NACA 4-digit airfoil O-grid with its flow conditions.  Its life:

1. arrival (Poisson process) -> host preparation: mesh and preprocessing
   (fvm_geometric);
2. admission, first come first served, while the HBM budget holds its
   footprint (fvm_gpu_footprint_estimate): RANS wall distance on the GPU,
   context creation (device layout, upload);
3. it joins the batch of its class (flow kind x mesh-size class) at the
   next engine step.  Each step launches a chunk of iterations on every
   class's batch (one stream each, so they run at the same time), prepares
   and admits new requests while the chunks run, waits, and retires the
   members that stopped (residual drop reached or iteration cap);
4. the monitors of the last iteration (fres, cl, cd) are the result; the
   context is destroyed and its bytes go back to the budget.

Every member computes bit for bit what it would compute alone (the batch
kernels run the same operations in the same order), so a result does not
depend on what else was in flight.

The reference has no serving layer (one solve per process, src/main.f); this
module sits above the C-ABI and changes nothing in the solve itself.
"""
from __future__ import annotations

import dataclasses
import math
import sys
import time

import numpy as np

from . import gpu, mesh
from .solver import Batch, Inputs, Solver, geometric, gpu_wall_distance


@dataclasses.dataclass
class Request:
    rid: int
    flow: str                 # "euler" or "rans"
    m: float                  # max camber
    p: float                  # camber position
    t: float                  # thickness
    nj: int                   # radial points; wrap-around points ni - 1 = 4 (nj - 1)
    mach: float
    aoa: float                # degrees
    rey: float                # Reynolds number (RANS)
    seed: int                 # diagonal coin flips of the O-grid
    arrival: float = 0.0      # seconds after the start of the run

    @property
    def ni(self) -> int:
        return 4 * (self.nj - 1) + 1

    @property
    def np(self) -> int:
        return (self.ni - 1) * self.nj


@dataclasses.dataclass
class Result:
    rid: int
    flow: str
    np: int
    arrival: float
    prepared: float
    admitted: float
    joined: float
    finished: float
    iters: int
    status: int
    fres: float
    cl: float
    cd: float

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def workload(n: int, seed: int, np_min: float = 1e3, np_max: float = 1e6,
             rans_fraction: float = 0.3) -> list[Request]:
    """Seeded requests: camber U[0, 0.04], camber position U[0.2, 0.6],
    thickness U[0.09, 0.18], vertices log-uniform in [np_min, np_max] with
    (ni - 1) = 4 (nj - 1), M U[0.3, 0.85], alpha U[-4, 8] deg, RANS with
    Re log-uniform in [1e6, 1e7] for a fraction of them."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        npv = math.exp(rng.uniform(math.log(np_min), math.log(np_max)))
        nj = max(9, int(round(0.5 + math.sqrt(npv / 4.0))))
        rans = rng.uniform() < rans_fraction
        out.append(Request(
            rid=i, flow="rans" if rans else "euler",
            m=float(rng.uniform(0.0, 0.04)), p=float(rng.uniform(0.2, 0.6)),
            t=float(rng.uniform(0.09, 0.18)), nj=nj,
            mach=float(rng.uniform(0.3, 0.85)), aoa=float(rng.uniform(-4.0, 8.0)),
            rey=float(math.exp(rng.uniform(math.log(1e6), math.log(1e7)))),
            seed=int(rng.integers(1, 2**31 - 1))))
    return out


def poisson_arrivals(reqs: list[Request], rate: float, seed: int) -> list[Request]:
    """Arrival times of a Poisson process of the given rate (requests/s);
    rate <= 0 puts every request at t = 0 (a closed batch, for saturation)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    for r in reqs:
        if rate > 0:
            t += float(rng.exponential(1.0 / rate))
        r.arrival = t
    return reqs


@dataclasses.dataclass
class ServeConfig:
    hbm_budget: int = 64 << 30        # bytes of context state admitted at once
    precision: int = 1                # Euler requests: 1 split precision (FP32 increments and flux
                                      # dissipation, FP64 states, consistent flux and update), 0 FP64
    # SA-RANS requests: split precision as well; its perturbation form follows
    # FP64 on the 2049x513 RANS batch to 20 000 iterations (CL within 2e-8)
    rans_precision: int = 1
    minres: float = 1e-6              # residual drop that ends a solve
    max_iters: int = 2000             # iteration cap of a solve
    # work of one engine step per class (sets its chunk): long enough that the
    # host work done meanwhile (meshing, preprocessing, context creation, up
    # to about a second for a million-vertex request) finds the GPU busy
    step_vertex_iters: float = 1e8
    size_classes: tuple = (16384, 131072)   # vertex bounds of the mesh-size classes
    cfl: float = 1.0
    first_spacing: float = 1e-3       # wall spacing of the O-grids (config 0 construction)
    # vertices per patch of a served solve: a batch fills the GPU with many
    # solves, so the patches are sized for little halo work rather than for
    # filling the GPU with one small mesh (0: the library's automatic size)
    patch_size: int = 64


def precision_for(req: Request, cfg: ServeConfig) -> int:
    return cfg.rans_precision if req.flow == "rans" else cfg.precision


def inputs_for(req: Request, cfg: ServeConfig) -> Inputs:
    return Inputs(iflow=3 if req.flow == "rans" else 1, iflux=1, ilimit=1,
                  mach_inf=req.mach, aoa_deg=req.aoa, rey=req.rey, cfl=cfg.cfl,
                  iterlast=0, maxiter=cfg.max_iters, minres=cfg.minres,
                  saveinterval=0, precision=precision_for(req, cfg), nut_factor=3.0, sa_compat=0,
                  patch_size=cfg.patch_size)


def prepare(req: Request, cfg: ServeConfig):
    """Host side of a request (no GPU calls, so it can run on worker
    threads): the O-grid and its preprocessing without the wall distance."""
    g = mesh.o_grid(req.ni, req.nj, m=req.m, p=req.p, t=req.t,
                    first_spacing=cfg.first_spacing, seed=req.seed)
    geo = geometric(g.coord, g.elem, g.ptype, iflow=1, sa_compat=0)
    return g, geo


def create(req: Request, g, geo, cfg: ServeConfig) -> Solver:
    """Device side of admission: the wall distance of a RANS request on the
    GPU, then the context.  RANS runs the SA model in its corrected form
    (sa_compat 0: nu_tilde at the triangle's own corners, exact wall
    distance); the freestream nu_tilde is 3 nu."""
    if req.flow == "rans":
        gpu_wall_distance(geo, 0)
    return Solver(g, inputs_for(req, cfg), geometry=geo)


def footprint(geo, req: Request, cfg: ServeConfig) -> int:
    d = geo.dims
    return int(gpu.lib().fvm_gpu_footprint_estimate(d["np"], d["nt"], d["netot"],
                                                      3 if req.flow == "rans" else 1,
                                                      precision_for(req, cfg)))


class Server:
    """Runs a list of requests to completion (arrivals in wall-clock time)
    and returns their results.

    One thread does everything, in steps: launch a chunk of iterations on
    every class's batch (each batch on its own stream, so the chunks run at
    the same time); while any chunk still runs, do host work one request at a
    time: admit the oldest waiting request if the HBM budget holds it (device
    wall distance, context creation), else prepare the oldest arrived one
    (host meshing and preprocessing); then wait, retire the members that
    stopped and let the new ones join."""

    def __init__(self, cfg: ServeConfig | None = None):
        self.cfg = cfg or ServeConfig()
        self.budget = self.cfg.hbm_budget
        self.waiting: list = []            # prepared, not admitted (FIFO)
        self.results: list[Result] = []
        self.errors: list = []
        self.t0 = 0.0
        self.peak_in_flight = 0
        self.in_flight = 0
        self.peak_reserved = 0             # most bytes admitted at once
        self.steps = 0
        self.wall = 0.0
        # wall seconds of the engine thread by phase
        self.phase = dict(launch=0.0, prepare=0.0, admit=0.0, wait=0.0, retire=0.0, join=0.0, idle=0.0)

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def size_class(self, npv: int) -> int:
        for k, bound in enumerate(self.cfg.size_classes):
            if npv < bound:
                return k
        return len(self.cfg.size_classes)

    def _chunk(self, members) -> int:
        """Iterations of one step for a class: about step_vertex_iters
        vertex-iterations, a multiple of 8 in [8, 512]."""
        nv = sum(m[1].np for m in members)
        k = int(round(self.cfg.step_vertex_iters / max(nv, 1) / 8.0)) * 8
        return max(8, min(512, k))

    def _host_unit(self, arrivals: list):
        """One unit of host work: admit the oldest waiting request if the
        budget holds its footprint (returns its entry, to join its batch
        after the running chunk), else prepare the oldest arrived request.
        Returns (did_work, entry or None)."""
        t0 = time.perf_counter()
        if self.waiting and self.waiting[0][3] <= self.budget:
            r, g, geo, fp, t_prep = self.waiting.pop(0)
            self.budget -= fp
            self.in_flight += 1
            self.peak_in_flight = max(self.peak_in_flight, self.in_flight)
            self.peak_reserved = max(self.peak_reserved, self.cfg.hbm_budget - self.budget)
            t_adm = self.now()
            try:
                s = create(r, g, geo, self.cfg)
            except Exception as e:  # noqa: BLE001 - recorded, the run goes on
                self.errors.append((r.rid, repr(e)))
                self.budget += fp
                self.in_flight -= 1
                self.phase["admit"] += time.perf_counter() - t0
                return True, None
            self.phase["admit"] += time.perf_counter() - t0
            return True, ((r.flow, self.size_class(s.np)), [r, s, fp, t_prep, t_adm, 0.0])
        if arrivals and arrivals[0].arrival <= self.now():
            r = arrivals.pop(0)
            try:
                g, geo = prepare(r, self.cfg)
                self.waiting.append((r, g, geo, footprint(geo, r, self.cfg), self.now()))
            except Exception as e:  # noqa: BLE001
                self.errors.append((r.rid, repr(e)))
            self.phase["prepare"] += time.perf_counter() - t0
            return True, None
        return False, None

    def _join(self, classes, ready):
        for key, entry in ready:
            if key not in classes:
                classes[key] = (Batch(), [])
            batch, members = classes[key]
            batch.add(entry[1])
            entry[5] = self.now()
            members.append(tuple(entry))

    def _retire(self, batch, members, statuses):
        keep = []
        for m, code in zip(members, statuses):
            if code == 0:
                keep.append(m)
                continue
            r, s, fp, t_prep, t_adm, t_join = m
            mon, status = s.monitor()
            self.results.append(Result(
                rid=r.rid, flow=r.flow, np=s.np, arrival=r.arrival, prepared=t_prep,
                admitted=t_adm, joined=t_join, finished=self.now(), iters=mon["iter"],
                status=status, fres=mon["fres"], cl=mon["cl"], cd=mon["cd"]))
            batch.remove(s)
            s.close()
            self.budget += fp
            self.in_flight -= 1
        members[:] = keep

    def run(self, reqs: list[Request], timeout: float | None = None,
            progress: float | None = None) -> list[Result]:
        """Serves reqs to completion; with `progress` (seconds), prints a line
        on stderr that often (requests done, in flight, waiting)."""
        n = len(reqs)
        arrivals = sorted(reqs, key=lambda r: r.arrival)
        classes: dict = {}
        self.t0 = time.perf_counter()
        t_print = progress
        try:
            while len(self.results) + len(self.errors) < n:
                if timeout is not None and self.now() > timeout:
                    raise TimeoutError(f"serving run over {timeout} s: {len(self.results)} of {n} done")
                if progress is not None and self.now() > t_print:
                    t_print += progress
                    print(f"serving: {self.now():.0f} s, {len(self.results)} done, {self.in_flight} in flight, "
                          f"{len(self.waiting)} waiting, {len(arrivals)} not arrived, {len(self.errors)} errors",
                          file=sys.stderr, flush=True)
                live = [(b, m) for b, m in classes.values() if m]
                tp = time.perf_counter()
                for b, m in live:
                    b.launch(self._chunk(m))
                self.phase["launch"] += time.perf_counter() - tp
                # host work while the chunks run (at least one unit a step);
                # with nothing running, until a request is admitted
                ready = []
                while True:
                    did, entry = self._host_unit(arrivals)
                    if entry is not None:
                        ready.append(entry)
                    if not did or (live and all(b.done() for b, _ in live)) or (not live and ready):
                        break
                for b, m in live:
                    tp = time.perf_counter()
                    st = b.wait()
                    tq = time.perf_counter()
                    self._retire(b, m, st)
                    self.phase["wait"] += tq - tp
                    self.phase["retire"] += time.perf_counter() - tq
                tp = time.perf_counter()
                self._join(classes, ready)
                self.phase["join"] += time.perf_counter() - tp
                if live:
                    self.steps += 1
                elif not ready and not self.waiting and arrivals:
                    tp = time.perf_counter()
                    time.sleep(max(0.0, min(arrivals[0].arrival - self.now(), 0.01)))
                    self.phase["idle"] += time.perf_counter() - tp
        finally:
            for b, m in classes.values():
                for entry in m:
                    entry[1].close()
                b.close()
        self.wall = self.now()
        return sorted(self.results, key=lambda r: r.rid)


def summary(results: list[Result], wall: float) -> dict:
    lat = np.array([r.latency for r in results]) if results else np.zeros(1)
    vi = float(sum(r.np * r.iters for r in results))
    return dict(solves=len(results), wall_s=round(wall, 3),
                vertex_iters_per_s=vi / wall if wall > 0 else 0.0,
                solves_per_s=len(results) / wall if wall > 0 else 0.0,
                p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                converged=sum(1 for r in results if r.status == 1),
                capped=sum(1 for r in results if r.status == 2),
                failed=sum(1 for r in results if r.status < 0))


def solve_alone(req: Request, cfg: ServeConfig, precision: int | None = None):
    """The same request solved by itself with the plain path (fvm_gpu_iterate)."""
    c = cfg if precision is None else dataclasses.replace(cfg, precision=precision,
                                                          rans_precision=precision)
    g, geo = prepare(req, c)
    with create(req, g, geo, c) as s:
        try:
            s.iterate(c.max_iters)
        except gpu.FvmError:
            pass
        mon, status = s.monitor()
        return mon, status, s.np
