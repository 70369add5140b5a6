"""Throughput of the FP64 residual + RK3 path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--flux 1|2|3]

A step is one pseudo-time iteration (save_old, time_step, 3 x [residual,
RK stage], CL/CD, residual norms) of the transonic Euler workload on the
seeded 513 x 129 NACA0012 O-grid (66 048 vertices, 131 072 triangles,
197 120 edges; M = 0.8, alpha = 1.25 deg, Roe flux, limited MUSCL, CFL 1).
The grid and state are resident in HBM before timing starts.  Prints one
JSON line with the throughput in vertex-iterations/s, the roofline of the
dominant kernel (algorithmic bytes per launch / its HIP-event duration on
the solver's stream) and the CPU baseline (the oracle, single core, on a
bounded sample of the same workload); beside them split precision, batches
of solves, the serving loop's p50/p99 latency on a bounded stream of
requests solved to a 1e-6 residual drop (with converged / capped counts), the
viscous workloads (laminar 1025x257; SA-RANS 2049x513, a batch of 8) in FP64
and split precision with their own rooflines, and the same transonic workload
with the other Riemann solvers (KFVS, HLLC as shipped, HLLC rotated) and on
O-grids of 16k to 1M vertices.
"""
from __future__ import annotations

import argparse
import faulthandler
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, vendor figure


def algorithmic_bytes(kernel: str, npt: int, nt: int, ne: int) -> float | None:
    """Compulsory traffic of one launch (each array touched once, gathers
    counted once per vertex; FP64 8 B, int32 4 B):
      gradient pass  G = 12 nt + (16 + 32 + 8 + 64) np
      edge-flux pass F = 24 ne + (16 + 32 + 64 + 32) np
      update pass    U = (32 + 32 + 32 + 8 + 8 + 32) np
      save_old + time_step = (64 + 48) np"""
    G, F, U = 12 * nt + 120 * npt, 24 * ne + 144 * npt, 144 * npt
    return {"k_grad": G, "k_grad2": G, "k_flux": F, "k_update": U, "k_fu2": F + U,
            "k_patch": G + F + U, "k_tstep": 112 * npt}.get(kernel)


def stage_bytes(flow: int, npt: int, nt: int, ne: int) -> float:
    """Algorithmic bytes of one RK stage (gradient + flux + update passes):
    Euler 12 nt + 24 ne + 408 np; viscous flow adds the Sutherland / mu pass
    (48 np), the Galerkin viscous terms in the gradient pass (8 nt + 56 np)
    and the residual read-modify-write in the flux pass (32 np)."""
    b = 12 * nt + 24 * ne + 408 * npt
    return b + (8 * nt + 136 * npt if flow != 1 else 0)


def iteration_bytes(flow: int, npt: int, nt: int, ne: int) -> float:
    """Algorithmic bytes of one iteration: three stages, save_old + time_step
    + conres (176 np; the viscous time step also reads mu, 8 np) and, for SA,
    the SA step (its triangle, edge and vertex passes: 20 nt + 24 ne +
    304 np)."""
    b = 3 * stage_bytes(flow, npt, nt, ne) + 176 * npt + (8 * npt if flow != 1 else 0)
    return b + (20 * nt + 24 * ne + 304 * npt if flow == 3 else 0)


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), if a summary for
    this workload is committed under profiles/."""
    path = os.path.join(ROOT, "profiles", "traffic_513x129_euler_roe.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


def stream_copy_gbs(nbytes: int = 1 << 30, reps: int = 20):
    """Achievable HBM bandwidth: a plain 16-byte copy kernel between two
    buffers of nbytes (fvm_gpu_stream_copy), read + write bytes per second."""
    import ctypes
    from fvm2d_amd import gpu
    v = ctypes.c_double(0.0)
    rc = gpu.lib().fvm_gpu_stream_copy(nbytes, reps, ctypes.byref(v))
    return round(v.value, 1) if rc == 0 else None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def _cpu_sample(args) -> float:
    """One process of the CPU aggregate: pinned to one host CPU, the same
    bounded sample as the single-core leg."""
    grid, kw, n, cpu = args
    if cpu is not None and hasattr(os, "sched_setaffinity"):
        os.sched_setaffinity(0, {cpu})
    from oracle.oracle import Oracle
    o = Oracle(grid.coord, grid.elem, grid.ptype, **kw)
    t0 = time.perf_counter()
    h, _ = o.iterate(n)
    return grid.np * h.shape[0] / (time.perf_counter() - t0)


def cpu_baseline(grid, kw, target_s: float = 12.0, procs: int = 16) -> dict:
    """The oracle (a serial C restatement of the CPU solver, gcc -O3) on the
    host: iterations of the same workload from the free stream, on one core,
    then as independent processes pinned one per host CPU (up to procs, within
    the CPUs this process may use), run at the same time.  Runs before the
    GPU is touched, so the processes are plain forks."""
    import multiprocessing
    from oracle.oracle import Oracle
    o = Oracle(grid.coord, grid.elem, grid.ptype, **kw)
    t0 = time.perf_counter()
    o.iterate(2)
    t_it = (time.perf_counter() - t0) / 2
    n = max(3, min(2000, int(target_s / max(t_it, 1e-6))))
    o2 = Oracle(grid.coord, grid.elem, grid.ptype, **kw)
    t0 = time.perf_counter()
    h, _ = o2.iterate(n)
    dt = time.perf_counter() - t0
    model = cpu_model()
    out = {"value": grid.np * h.shape[0] / dt, "unit": "vertex-iterations/s", "cores": 1,
           "kind": "port",
           "sample": f"{h.shape[0]} iterations of the same 513x129 Euler/Roe workload from the "
                     f"free stream, serial C oracle (gcc -O3 -ffp-contract=off) on {model}, "
                     f"{dt:.1f} s"}
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else []
    k = min(procs, len(cpus))
    if k > 1:
        with multiprocessing.get_context("fork").Pool(k) as pool:
            t0 = time.perf_counter()
            rates = pool.map(_cpu_sample, [(grid, kw, n, c) for c in cpus[:k]])
            wall = time.perf_counter() - t0
        out["aggregate"] = {
            "value": sum(rates), "unit": "vertex-iterations/s", "processes": k,
            "slowest_process": min(rates),
            "note": f"{k} independent processes of the same sample at once, one per host CPU "
                    f"({wall:.1f} s)"}
    return out


def batch_legs(g, geo, kw, plain_value: float, warmup: int) -> dict:
    """Batches: (a) eight solves of the same grid at alpha = 0, 2, ..., 14 deg
    in one launch per RK stage, FP64 and split precision; (b) 48 requests of
    the serving workload's distribution (NACA 4-digit airfoils, 1k-20k
    vertices, 70 % Euler / 30 % SA-RANS, seeded) as two batches (one per flow
    kind, one stream each) with the serving precisions (split precision for
    both kinds), against the same requests solved one after another by the
    plain FP64 path."""
    from fvm2d_amd import serving
    from fvm2d_amd.solver import Batch, Inputs, Solver
    out = {}
    steps = 200
    alphas = [2.0 * k for k in range(8)]
    for prec, name in ((0, "fp64"), (1, "split")):
        # 256-thread workgroups: with eight solves in a launch the smaller
        # patches balance better (8 x 513x129: 1.99e9 against 1.95e9 FP64,
        # 2.46e9 against 2.37e9 split with the 512 threads a single solve of
        # this size gets)
        sol = [Solver(g, Inputs.from_kwargs(**dict(kw, aoa_deg=al, precision=prec, block_threads=256)),
                      geometry=geo)
               for al in alphas]
        with Batch() as b:
            for x in sol:
                b.add(x)
            b.iterate(warmup)
            t0 = time.perf_counter()
            b.iterate(steps)
            dt = time.perf_counter() - t0
        out[f"same_grid_8_alphas_{name}"] = round(len(sol) * geo.dims["np"] * steps / dt, 1)
        for x in sol:
            x.close()
    out["same_grid_8_alphas_split_vs_plain_fp64"] = round(out["same_grid_8_alphas_split"] / plain_value, 2)

    reqs = serving.workload(48, 2024, np_min=1e3, np_max=2e4)
    cfg = serving.ServeConfig(max_iters=10**6, minres=0.0)
    prepared = [(q, *serving.prepare(q, cfg)) for q in reqs]
    # plain FP64 path, one request at a time, with the library's automatic
    # patch size (iteration time only)
    c64 = serving.ServeConfig(max_iters=10**6, minres=0.0, precision=0, patch_size=0)
    vi = 0.0
    tp = 0.0
    from fvm2d_amd.gpu import FvmError
    for q, g1, geo1 in prepared:
        with serving.create(q, g1, geo1, c64) as x:
            x.iterate(8)
            i0 = x.monitor()[0]["iter"]
            t0 = time.perf_counter()
            try:
                x.iterate(steps)
            except FvmError:            # a solve that stops on a reference STOP condition
                pass
            tp += time.perf_counter() - t0
            vi += x.np * (x.monitor()[0]["iter"] - i0)
    plain = vi / tp
    # batched with the serving precisions, one batch per flow kind, chunks
    # launched together
    sol = {"euler": [], "rans": []}
    for q, g1, geo1 in prepared:
        sol[q.flow].append(serving.create(q, g1, geo1, cfg))
    batches = []
    for xs in sol.values():
        if xs:
            b = Batch()
            for x in xs:
                b.add(x)
            batches.append(b)
    for b in batches:
        b.iterate(8)
    i0 = {id(x): x.monitor()[0]["iter"] for xs in sol.values() for x in xs}
    t0 = time.perf_counter()
    for _ in range(steps // 50):
        for b in batches:
            b.launch(50)
        for b in batches:
            b.wait()
    dt = time.perf_counter() - t0
    nv = sum(x.np for xs in sol.values() for x in xs)
    vb = sum(x.np * (x.monitor()[0]["iter"] - i0[id(x)]) for xs in sol.values() for x in xs)
    for b in batches:
        b.close()
    for xs in sol.values():
        for x in xs:
            x.close()
    out["mixed_48_requests"] = {
        "workload": "48 NACA 4-digit requests, 1k-20k vertices log-uniform, 70/30 Euler/SA-RANS, "
                    f"{nv} vertices in all, {steps} iterations each",
        "precision": "split (FP32 increments and dissipation, FP64 states, consistent flux and update)",
        "batched": round(vb / dt, 1), "plain_fp64_one_at_a_time": round(plain, 1),
        "speedup": round(vb / dt / plain, 2)}
    return out


def serving_leg(n: int = 32, seed: int = 11, max_iters: int = 20000) -> dict:
    """Solve latency of the serving loop (fvm2d_amd/serving.py) on a bounded
    stream: n seeded requests of 1000-4000 vertices (NACA 4-digit airfoils,
    70/30 Euler/SA-RANS), each solved to a 1e-6 residual drop, with an
    iteration cap of max_iters, in split precision.
    Offered all at once, the run gives the saturation rate in solves/s; then
    the same requests arrive as Poisson processes at 50 % and 90 % of that
    rate.  A latency runs from a request's arrival to its result: host
    meshing and preprocessing, admission, the solve.  Latencies are given over
    all requests and over the converged ones, with the counts."""
    import numpy as np
    from fvm2d_amd import serving
    cfg = serving.ServeConfig(max_iters=max_iters, minres=1e-6, size_classes=(2500,))

    def run(rate: float) -> dict:
        reqs = serving.poisson_arrivals(serving.workload(n, seed, np_min=1000, np_max=4000),
                                        rate, seed)
        srv = serving.Server(cfg)
        res = srv.run(reqs, timeout=240.0)
        s = serving.summary(res, srv.wall)
        conv = [r.latency for r in res if r.status == 1]
        out = {"offered_per_s": round(rate, 3), "solves": s["solves"], "wall_s": s["wall_s"],
               "solves_per_s": round(s["solves_per_s"], 3),
               "vertex_iters_per_s": round(s["vertex_iters_per_s"], 1),
               "p50_latency_s": round(s["p50_latency_s"], 4),
               "p99_latency_s": round(s["p99_latency_s"], 4),
               "converged": s["converged"], "capped": s["capped"], "failed": s["failed"],
               "p50_converged_latency_s": round(float(np.percentile(conv, 50)), 4) if conv else None,
               "p99_converged_latency_s": round(float(np.percentile(conv, 99)), 4) if conv else None,
               "mean_iterations": round(float(np.mean([r.iters for r in res])), 1),
               "peak_in_flight": srv.peak_in_flight}
        if srv.errors:
            out["errors"] = srv.errors
        return out

    sat = run(0.0)                      # every request at t = 0: the saturation rate
    rate = sat["solves_per_s"]
    return {"workload": f"{n} requests, NACA 4-digit O-grids of 1000-4000 vertices, 70/30 "
                        f"Euler/SA-RANS (split precision), each to a 1e-6 "
                        f"residual drop, iteration cap {max_iters}",
            "saturation": sat, "load_50": run(0.5 * rate), "load_90": run(0.9 * rate)}


def serving_full(n: int, seed: int, cap: int, rate: float, plain_every: int = 0,
                 step_vertex_iters: float = 3e8) -> dict:
    """The serving workload as specified: a seeded stream of n requests
    (NACA 4-digit airfoils, max camber 0-4 %, camber position 0.2-0.6,
    thickness 9-18 %; O-grids of 1k-1M vertices, log-uniform; M 0.3-0.85,
    alpha -4..8 deg; 70 % Euler, 30 % SA-RANS with Re 1e6-1e7 log-uniform),
    each solved to a 1e-6 residual drop with an iteration cap of `cap`.
    rate <= 0: every request at t = 0, which gives the saturation rate; else
    Poisson arrivals at that rate.  With plain_every = k, every k-th request
    is also solved alone by the plain FP64 path (one context, fvm_gpu_iterate,
    same cap) for the throughput comparison on the same requests."""
    import numpy as np
    from fvm2d_amd import serving
    from fvm2d_amd.gpu import FvmError
    cfg = serving.ServeConfig(max_iters=cap, minres=1e-6, step_vertex_iters=step_vertex_iters)
    reqs = serving.poisson_arrivals(serving.workload(n, seed), rate, seed)
    out = {"workload": f"{n} requests (seed {seed}), NACA 4-digit O-grids of 1k-1M vertices, 70/30 "
                       f"Euler/SA-RANS (split precision), each to a 1e-6 residual "
                       f"drop, iteration cap {cap}",
           "offered_per_s": rate, "vertices_total": int(sum(q.np for q in reqs))}
    if plain_every:
        vi, tp, cnt = 0.0, 0.0, 0
        c64 = serving.ServeConfig(max_iters=cap, minres=1e-6, precision=0, patch_size=0)
        for q in reqs[::plain_every]:
            g1, geo1 = serving.prepare(q, c64)
            with serving.create(q, g1, geo1, c64) as x:
                t0 = time.perf_counter()
                try:
                    x.iterate(cap)
                except FvmError:
                    pass
                tp += time.perf_counter() - t0
                vi += x.np * x.monitor()[0]["iter"]
                cnt += 1
            print(f"plain FP64: {cnt} requests, {vi / tp:.3e} vertex-iterations/s", file=sys.stderr, flush=True)
        out["plain_fp64_one_at_a_time"] = {"requests": cnt, "every": plain_every,
                                           "vertex_iters_per_s": round(vi / tp, 1),
                                           "solve_time_s": round(tp, 1)}
        return out
    srv = serving.Server(cfg)
    res = srv.run(reqs, progress=30.0)
    s = serving.summary(res, srv.wall)
    lat = np.array([r.latency for r in res])
    conv = np.array([r.latency for r in res if r.status == 1])
    by_flow = {}
    for flow in ("euler", "rans"):
        rs = [r for r in res if r.flow == flow]
        by_flow[flow] = {"requests": len(rs), "converged": sum(r.status == 1 for r in rs),
                         "capped": sum(r.status == 2 for r in rs), "failed": sum(r.status < 0 for r in rs)}
    out.update({"solves": s["solves"], "wall_s": s["wall_s"], "solves_per_s": round(s["solves_per_s"], 4),
                "vertex_iters_per_s": round(s["vertex_iters_per_s"], 1),
                "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
                "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
                "p50_converged_latency_s": round(float(np.percentile(conv, 50)), 3) if conv.size else None,
                "p99_converged_latency_s": round(float(np.percentile(conv, 99)), 3) if conv.size else None,
                "converged": s["converged"], "capped": s["capped"], "failed": s["failed"], "by_flow": by_flow,
                "peak_in_flight": srv.peak_in_flight, "peak_reserved_bytes": srv.peak_reserved,
                "engine_steps": srv.steps, "engine_phase_s": {k: round(v, 2) for k, v in srv.phase.items()},
                "errors": srv.errors[:10]})
    return out


def size_legs(kw) -> list:
    """The same Euler/Roe workload on seeded O-grids of 16 640 to 1 050 624
    vertices (±10 % jitter), FP64 and split precision: vertex-iterations/s
    over about half a second of iterations after as many untimed, and the
    stage kernel's roofline (50 launches between HIP events, best of 3)."""
    from fvm2d_amd import mesh
    from fvm2d_amd.solver import Inputs, Solver
    out = []
    for ni, nj in ((257, 65), (513, 129), (1025, 257), (2049, 513)):
        g = mesh.o_grid(ni, nj, jitter=0.1, seed=ni * 1000 + nj)
        for prec in (0, 1):
            with Solver(g, Inputs.from_kwargs(**dict(kw, precision=prec))) as x:
                d = x.geo.dims
                n = max(200, int(1e9 / d["np"]))
                x.iterate(n)
                t0 = time.perf_counter()
                h, _ = x.iterate(n)
                dt = time.perf_counter() - t0
                t_ms = min(x.time_stage(50) for _ in range(3))
                B = algorithmic_bytes("k_patch", d["np"], d["nt"], d["ne"])
                info = x.layout_info()
                out.append({"grid": f"{ni}x{nj}", "np": d["np"], "precision": "split" if prec else "fp64",
                            "iterations": int(h.shape[0]),
                            "vertex_iters_per_s": round(d["np"] * h.shape[0] / dt, 1),
                            "stage_ms": round(t_ms, 5),
                            "frac": round(B / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                            "patches": info["patches"], "patch_size": info["patch_size"],
                            "threads": info["threads"]})
    return out


def flux_legs(g, geo, kw, steps: int, warmup: int) -> dict:
    """The transonic 513x129 workload with the other Riemann solvers, FP64,
    `steps` iterations each after `warmup`: KFVS, HLLC as the reference ships
    it (flux left in the edge frame; on this case it stops in its first
    iteration on the reference's positivity check, which is reported with
    the stage time on the free-stream state) and HLLC rotated to x-y
    (`hllc_rotate` 1).  Each with its stage kernel's roofline (100 launches
    between HIP events; the byte model does not depend on the flux)."""
    from fvm2d_amd.gpu import FvmError
    from fvm2d_amd.solver import Inputs, Solver
    dims = geo.dims
    npt, nt, ne = dims["np"], dims["nt"], dims["ne"]
    B = algorithmic_bytes("k_patch", npt, nt, ne)
    out = {}
    for name, iflux, rot in (("kfvs", 2, 0), ("hllc", 3, 0), ("hllc_rotated", 3, 1)):
        inp = Inputs.from_kwargs(**dict(kw, iflux=iflux, hllc_rotate=rot))
        r = {}
        with Solver(g, inp, geometry=geo) as x:
            try:
                x.iterate(warmup)
                t0 = time.perf_counter()
                h, _ = x.iterate(steps)
                dt = time.perf_counter() - t0
                r.update(vertex_iters_per_s=round(npt * h.shape[0] / dt, 1), iterations=int(h.shape[0]),
                         cl=float(h[-1, 4]), cd=float(h[-1, 5]), fres_drop=float(h[-1, 1]))
                t_ms = x.time_stage(100)
            except FvmError as e:
                r["stopped"] = {"status": e.code, "index": e.index,
                                "iteration": int(x.monitor()[0]["iter"]) + 1}
                t_ms = None
        if t_ms is None:                # stage time on the free-stream state
            with Solver(g, inp, geometry=geo) as x:
                t_ms = x.time_stage(100)
            r["stage_state"] = "free stream"
        r["stage_ms"] = round(t_ms, 5)
        r["frac"] = round(B / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        out[name] = r
    return out


def config_legs(iters_lam: int, iters_rans: int) -> dict:
    """The viscous workloads, FP64 and split precision.

    Laminar: NACA0012 1025x257 O-grid (wall spacing 1e-4, ratio 1.1), M 0.5,
    alpha 0, Re 5000, Roe, limited.  RANS: NACA0012 2049x513 O-grid (wall
    spacing 4e-6, ratio 1.15), M 0.15, Re 6e6, SA in its corrected form
    (sa_compat 0: nu_tilde gradient at each triangle's corners, exact wall
    distance) with freestream nu_tilde = 3 nu, a batch of 8 solves at alpha =
    0, 2, ..., 14 deg in one launch per RK stage.  Both run a bounded number
    of iterations from the free stream (stated); the split-minus-FP64
    differences of CL, CD and of the L2 density-residual history are taken at
    the same iteration.  Rooflines: the laminar stage kernel over 100
    back-to-back launches (HIP events on the solver's stream) against the
    viscous stage bytes; the RANS batch over whole iterations (its kernels and
    the gaps between them) against the SA iteration bytes of its 8 solves."""
    import numpy as np
    from fvm2d_amd import mesh
    from fvm2d_amd.solver import Batch, Inputs, Solver, geometric
    out = {}
    g = mesh.naca0012_1025x257()
    geo = geometric(g.coord, g.elem, g.ptype, iflow=2)
    d = geo.dims
    kw = dict(iflow=2, iflux=1, ilimit=1, mach_inf=0.5, aoa_deg=0.0, rey=5000.0, cfl=1.0,
              maxiter=10**9, minres=0.0)
    lam = {}
    for prec in (0, 1):
        print(f"laminar 1025x257, precision {prec}", file=sys.stderr, flush=True)
        with Solver(g, Inputs.from_kwargs(precision=prec, **kw), geometry=geo) as x:
            x.iterate(2)
            t0 = time.perf_counter()
            h, _ = x.iterate(iters_lam)
            rate = d["np"] * iters_lam / (time.perf_counter() - t0)
            stage_ms = x.time_stage(100)
            lam[prec] = (h, rate, stage_ms)
    h0, h1 = lam[0][0], lam[1][0]
    B = stage_bytes(2, d["np"], d["nt"], d["ne"])
    out["laminar_1025x257"] = {
        "iterations": iters_lam, "np": d["np"],
        "fp64_vertex_iters_per_s": round(float(lam[0][1]), 1),
        "split_vertex_iters_per_s": round(float(lam[1][1]), 1),
        "roofline": {"bound": "hbm", "kernel": "stage kernel (laminar, FP64)",
                     "achieved": round(B / (lam[0][2] * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(B / (lam[0][2] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_launch": B, "kernel_ms_per_launch": round(lam[0][2], 5),
                     "split_kernel_ms_per_launch": round(lam[1][2], 5)},
        "fres_drop_fp64": float(h0[-1, 1]), "cl_fp64": float(h0[-1, 4]), "cd_fp64": float(h0[-1, 5]),
        "dcl": abs(float(h1[-1, 4] - h0[-1, 4])), "dcd": abs(float(h1[-1, 5] - h0[-1, 5])),
        "max_rel_dfres": float(np.max(np.abs(h1[:, 1] - h0[:, 1]) / np.abs(h0[:, 1])))}
    del geo
    g = mesh.naca0012_2049x513()
    t0 = time.perf_counter()
    geo = geometric(g.coord, g.elem, g.ptype, iflow=3, sa_compat=0, wall_distance="gpu")
    prep_s = time.perf_counter() - t0
    d = geo.dims
    kw = dict(iflow=3, iflux=1, ilimit=1, mach_inf=0.15, rey=6.0e6, cfl=1.0, maxiter=10**9,
              minres=0.0, nut_factor=3.0, sa_compat=0)
    alphas = [2.0 * k for k in range(8)]
    rans = {}
    for prec in (0, 1):
        print(f"RANS 2049x513 batch of 8, precision {prec}", file=sys.stderr, flush=True)
        sol = [Solver(g, Inputs.from_kwargs(precision=prec, aoa_deg=al, **kw), geometry=geo)
               for al in alphas]
        with Batch() as b:
            for x in sol:
                b.add(x)
            b.iterate(2)
            t0 = time.perf_counter()
            st = b.iterate(iters_rans)
            dt = time.perf_counter() - t0
            mons = [x.monitor()[0] for x in sol]
        for x in sol:
            x.close()
        rans[prec] = (mons, len(sol) * d["np"] * iters_rans / dt, st, dt)
    Bi = len(alphas) * iteration_bytes(3, d["np"], d["nt"], d["ne"])
    it_ms = rans[0][3] / iters_rans * 1e3
    out["rans_2049x513_batch8"] = {
        "iterations": iters_rans, "np": d["np"], "alphas_deg": alphas, "sa": "corrected (sa_compat 0)",
        "fp64_vertex_iters_per_s": round(rans[0][1], 1),
        "split_vertex_iters_per_s": round(rans[1][1], 1),
        "roofline": {"bound": "hbm", "kernel": "whole iteration of the batch (FP64)",
                     "achieved": round(Bi / (it_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(Bi / (it_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_iteration": Bi, "ms_per_iteration": round(it_ms, 4)},
        "preprocessing_s": round(prep_s, 2),
        "status_fp64": rans[0][2], "status_split": rans[1][2],
        "fres_fp64": [m["fres"] for m in rans[0][0]],
        "cl_fp64": [m["cl"] for m in rans[0][0]], "cd_fp64": [m["cd"] for m in rans[0][0]],
        "dcl": [abs(a["cl"] - b["cl"]) for a, b in zip(rans[0][0], rans[1][0])],
        "dcd": [abs(a["cd"] - b["cd"]) for a, b in zip(rans[0][0], rans[1][0])]}
    return out


def rans_converge(max_iters: int, chunk: int = 1000) -> dict:
    """The RANS workload as specified: the batch of 8 SA solves on the
    2049x513 grid (alpha 0..14 deg) run to a 1e-8 drop of the density
    residual or max_iters iterations, in FP64 and in split precision; per
    solve the iterations, status, final residual drop, CL and CD, and the
    split-minus-FP64 differences of CL and CD at the end."""
    from fvm2d_amd import mesh
    from fvm2d_amd.solver import Batch, Inputs, Solver, geometric
    g = mesh.naca0012_2049x513()
    geo = geometric(g.coord, g.elem, g.ptype, iflow=3, sa_compat=0, wall_distance="gpu")
    kw = dict(iflow=3, iflux=1, ilimit=1, mach_inf=0.15, rey=6.0e6, cfl=1.0, maxiter=max_iters,
              minres=1e-8, nut_factor=3.0, sa_compat=0)
    alphas = [2.0 * k for k in range(8)]
    res = {}
    for prec in (0, 1):
        sol = [Solver(g, Inputs.from_kwargs(precision=prec, aoa_deg=al, **kw), geometry=geo)
               for al in alphas]
        t0 = time.perf_counter()
        with Batch() as b:
            for x in sol:
                b.add(x)
            done = 0
            while done < max_iters:
                st = b.iterate(min(chunk, max_iters - done))
                done += chunk
                mons = [x.monitor()[0] for x in sol]
                print(f"RANS batch, precision {prec}: {min(done, max_iters)} iterations, "
                      f"fres {max(m['fres'] for m in mons):.3e}, status {st}",
                      file=sys.stderr, flush=True)
                if all(v != 0 for v in st):
                    break
            mons = [x.monitor()[0] for x in sol]
        wall = time.perf_counter() - t0
        for x in sol:
            x.close()
        res[prec] = (mons, st, wall)
    m0, m1 = res[0][0], res[1][0]
    return {"workload": "NACA0012 2049x513 (1 050 624 vertices), SA, M 0.15, Re 6e6, "
                        "alpha 0..14 deg, a batch of 8, to a 1e-8 drop or %d iterations" % max_iters,
            "alphas_deg": alphas,
            "iterations_fp64": [m["iter"] for m in m0], "iterations_split": [m["iter"] for m in m1],
            "status_fp64": res[0][1], "status_split": res[1][1],
            "fres_fp64": [m["fres"] for m in m0], "fres_split": [m["fres"] for m in m1],
            "cl_fp64": [m["cl"] for m in m0], "cd_fp64": [m["cd"] for m in m0],
            "dcl": [abs(a["cl"] - b["cl"]) for a, b in zip(m0, m1)],
            "dcd": [abs(a["cd"] - b["cd"]) for a, b in zip(m0, m1)],
            "wall_s_fp64": round(res[0][2], 1), "wall_s_split": round(res[1][2], 1)}


def lam_converge(iters: int, chunk: int = 1000) -> dict:
    """The laminar workload (1025x257, M 0.5, Re 5000) run for `iters`
    iterations in FP64 and in split precision; residual drop, CL, CD and the
    split-minus-FP64 differences every `chunk` iterations."""
    import numpy as np
    from fvm2d_amd import mesh
    from fvm2d_amd.solver import Inputs, Solver, geometric
    g = mesh.naca0012_1025x257()
    geo = geometric(g.coord, g.elem, g.ptype, iflow=2)
    kw = dict(iflow=2, iflux=1, ilimit=1, mach_inf=0.5, aoa_deg=0.0, rey=5000.0, cfl=1.0,
              maxiter=10**9, minres=0.0)
    hist = {}
    for prec in (0, 1):
        with Solver(g, Inputs.from_kwargs(precision=prec, **kw), geometry=geo) as x:
            parts = []
            for k in range(0, iters, chunk):
                h, _ = x.iterate(min(chunk, iters - k))
                parts.append(h)
                print(f"laminar, precision {prec}: {k + h.shape[0]} iterations, fres {h[-1, 1]:.3e}",
                      file=sys.stderr, flush=True)
            hist[prec] = np.concatenate(parts)
    h0, h1 = hist[0], hist[1]
    rows = []
    for it in range(chunk - 1, min(h0.shape[0], h1.shape[0]), chunk):
        rows.append({"iter": int(h0[it, 0]), "fres_fp64": float(h0[it, 1]), "fres_split": float(h1[it, 1]),
                     "cl_fp64": float(h0[it, 4]), "cd_fp64": float(h0[it, 5]),
                     "dcl": abs(float(h1[it, 4] - h0[it, 4])), "dcd": abs(float(h1[it, 5] - h0[it, 5]))})
    return {"workload": "NACA0012 1025x257 (263 168 vertices), laminar, M 0.5, Re 5000, FP64 and split",
            "rows": rows}


def main() -> int:
    faulthandler.enable()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    # about 0.1 s of untimed iterations: after 50 the same 2000 timed
    # iterations ran at 1.36e9 against 1.47e9 after 2000 or 5000 (the GPU's
    # clocks still ramping at the start of the work)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--flux", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the split-precision and batched legs")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-latency leg")
    ap.add_argument("--no-sizes", action="store_true",
                    help="skip the Euler legs on O-grids of 16k to 1M vertices")
    ap.add_argument("--no-fluxes", action="store_true",
                    help="skip the KFVS / HLLC / rotated-HLLC legs of the transonic workload")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the laminar 1025x257 and RANS 2049x513 batch-of-8 legs")
    ap.add_argument("--lam-iters", type=int, default=5000,
                    help="iterations of the laminar 1025x257 leg (FP64 and split)")
    ap.add_argument("--rans-iters", type=int, default=500,
                    help="iterations of the RANS 2049x513 batch-of-8 leg (FP64 and split)")
    ap.add_argument("--variant", type=int, default=None, help="kernel layout 1 or 2 (default: tuned)")
    ap.add_argument("--vb", type=int, default=64, help="vertices per workgroup (variant 2)")
    ap.add_argument("--tune", action="store_true",
                    help="pick the launch shape with the autotuner (cached) instead of the "
                         "occupancy model")
    ap.add_argument("--configs", nargs=2, type=int, default=None, metavar=("LAM", "RANS"),
                    help="only run the split-vs-FP64 legs of the laminar 1025x257 and RANS "
                         "2049x513 batch-of-8 workloads for LAM and RANS iterations")
    ap.add_argument("--lam-converge", type=int, default=0, metavar="N",
                    help="only run the laminar 1025x257 workload N iterations, FP64 and split, "
                         "and report the differences every 1000 iterations")
    ap.add_argument("--rans-converge", type=int, default=0, metavar="N",
                    help="only run the RANS batch of 8 to a 1e-8 residual drop or N iterations, "
                         "FP64 and split, and report CL/CD and their differences")
    ap.add_argument("--serving-full", type=int, default=0, metavar="N",
                    help="only run the serving workload as specified with N requests (see serving_full)")
    ap.add_argument("--serving-cap", type=int, default=1500, help="iteration cap of --serving-full")
    ap.add_argument("--serving-rate", type=float, default=0.0,
                    help="arrival rate of --serving-full in requests/s (0: all at once, saturation)")
    ap.add_argument("--serving-step", type=float, default=3e8,
                    help="vertex-iterations per class and engine step of --serving-full")
    ap.add_argument("--serving-plain-every", type=int, default=0,
                    help="with --serving-full: only solve every k-th request alone by the plain FP64 path")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES, five times: the "
                         "reference reading for the FETCH_SIZE/WRITE_SIZE counters")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        n = a.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 * n, "copies": 5}))
        return 0
    if a.configs:
        print(json.dumps(config_legs(*a.configs)), flush=True)
        return 0
    if a.serving_full:
        print(json.dumps(serving_full(a.serving_full, 2024, a.serving_cap, a.serving_rate,
                                      a.serving_plain_every, a.serving_step)), flush=True)
        return 0
    if a.rans_converge:
        print(json.dumps(rans_converge(a.rans_converge)), flush=True)
        return 0
    if a.lam_converge:
        print(json.dumps(lam_converge(a.lam_converge)), flush=True)
        return 0
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the workload is a single-GPU solve and this build "
              "has no multi-GPU path (independent solves shard one per GPU).", file=sys.stderr)
        return 2

    import numpy as np
    from fvm2d_amd import mesh

    g = mesh.naca0012_513x129()
    kw = dict(iflow=1, iflux=a.flux, ilimit=1, mach_inf=0.8, aoa_deg=1.25, cfl=1.0,
              maxiter=10**9, minres=0.0)
    # the CPU baseline runs before anything touches the GPU (its processes
    # are forks of this one)
    cpu = None if a.no_cpu else cpu_baseline(g, dict(kw))
    from fvm2d_amd.solver import Inputs, Solver
    launch = {"patch_size": 0, "xcd": 1, "chosen_by": "automatic launch shape (occupancy model)"}
    if a.tune:
        from fvm2d_amd import autotune
        from fvm2d_amd.solver import geometric
        geo0 = geometric(g.coord, g.elem, g.ptype)
        ch = autotune.tune(g, geo0, Inputs.from_kwargs(**kw), device="gfx950")
        kw["patch_size"] = ch.patch_size
        kw["block_threads"] = ch.threads
        launch = {"patch_size": ch.patch_size, "xcd": ch.xcd, "threads": ch.threads, "chosen_by": "autotuner",
                  "us_per_iter_tuning": round(ch.us_per_iter, 2)}
    s = Solver(g, Inputs.from_kwargs(**kw))
    if a.tune:
        s.tune(3, 64, launch["xcd"])
    if a.variant is not None:
        s.tune(a.variant, a.vb, 1)
    info = s.layout_info()
    launch.update(patches=info["patches"], patch_size=info["patch_size"], threads=info["threads"],
                  lds_bytes=info["lds_bytes"], workgroups_per_cu=info["workgroups_per_cu"])
    dims = s.geo.dims
    npt, nt, ne = dims["np"], dims["nt"], dims["ne"]

    from fvm2d_amd.gpu import FvmError
    stopped = None
    t0 = None
    try:
        s.iterate(a.warmup)
        t0 = time.perf_counter()
        hist, status = s.iterate(a.steps)      # returns after a device synchronize
        elapsed = time.perf_counter() - t0
    except FvmError as e:
        # a reference STOP (e.g. HLLC's positivity check on this transonic
        # case): with another flux than the default Roe the rate is reported
        # over the iterations that ran, with the stop
        if a.flux == 1 or t0 is None or not hasattr(e, "hist") or e.hist.shape[0] == 0:
            raise
        elapsed = time.perf_counter() - t0
        hist, status = e.hist, e.code
        stopped = {"status": e.code, "index": e.index, "timed_iterations": int(hist.shape[0])}
    if hist.shape[0] != a.steps and stopped is None:
        print(f"solve stopped early: status {status}", file=sys.stderr)
        return 1
    value = npt * hist.shape[0] / elapsed

    roofline = None
    if not a.no_kernel_timing:
        ms = None
        if a.variant in (None, 3):
            # the dominant kernel: the patch stage kernel, 3 launches per
            # iteration; timed over 100 launches between two events on the
            # solver's stream
            dom, t_ms = "k_patch", s.time_stage(100)
        else:
            # per kernel, one event pair per launch (the pair's own overhead
            # included)
            ms = s.time_kernels(20)
            per_stage = ("k_grad", "k_flux", "k_update")
            per_iter = {k: v * (3 if k in per_stage else 1) for k, v in ms.items()}
            dom = max(per_iter, key=per_iter.get)
            t_ms = ms[dom]
        B = algorithmic_bytes(dom, npt, nt, ne)
        if B is not None:
            achieved = B / (t_ms * 1e-3) / 1e9
            traffic = traffic_from_profiles(dom)
            copy = stream_copy_gbs()
            roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": traffic,
                        "traffic_rate_gbs": (round(traffic / (t_ms * 1e-3) / 1e9, 1)
                                             if traffic else None),
                        "achievable_copy_gbs": copy,
                        "frac_of_copy": round(achieved / copy, 4) if copy else None,
                        "kernel_ms_per_launch": round(t_ms, 5),
                        "algorithmic_bytes_per_launch": B}
            if ms:
                roofline["kernel_ms_event_pairs"] = {k: round(v, 5) for k, v in ms.items()}

    # split precision on the same workload: CL/CD against this FP64 run
    split = None
    if not a.no_batch:
        with Solver(g, Inputs.from_kwargs(precision=1, **kw), geometry=s.geo) as s1:
            s1.iterate(a.warmup)
            t1 = time.perf_counter()
            h1, _ = s1.iterate(a.steps)
            e1 = time.perf_counter() - t1
            t1_ms = None if a.no_kernel_timing else s1.time_stage(100)
        split = {"value": round(npt * a.steps / e1, 1), "unit": "vertex-iterations/s",
                 "dcl": abs(float(h1[-1, 4] - hist[-1, 4])), "dcd": abs(float(h1[-1, 5] - hist[-1, 5])),
                 "note": "same workload and step count with FP32 gradients, reconstruction and flux; "
                         "dcl, dcd against this FP64 run"}
        if t1_ms:
            # the split stage kernel moves the same HBM arrays (FP64 states;
            # its FP32 copies live in LDS), so the same byte model
            B1 = algorithmic_bytes("k_patch", npt, nt, ne)
            split["roofline"] = {"kernel": "k_patch (split precision)", "kernel_ms_per_launch": round(t1_ms, 5),
                                 "achieved": round(B1 / (t1_ms * 1e-3) / 1e9, 1), "unit": "GB/s",
                                 "frac": round(B1 / (t1_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    batched = None if a.no_batch else batch_legs(g, s.geo, kw, value, a.warmup)
    served = None if a.no_serving else serving_leg()
    viscous = None if a.no_configs else config_legs(a.lam_iters, a.rans_iters)
    fluxes = None if (a.no_fluxes or a.flux != 1) else flux_legs(g, s.geo, kw, a.steps, a.warmup)
    sizes = None if (a.no_sizes or a.flux != 1) else size_legs(kw)

    line = {
        "metric": "vertex-iters/s & %HBM roofline vs plain FP64; p50/p99 solve latency",
        "value": round(value, 1), "unit": "vertex-iterations/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(1e3 * elapsed / hist.shape[0], 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "NACA0012 O-grid 513x129 (66048 vertices, 131072 triangles), "
                               "Euler, %s flux, limited MUSCL, RK3, CFL 1, M=0.8, alpha=1.25, FP64"
                               % {1: "Roe", 2: "KFVS", 3: "HLLC"}[a.flux],
                   "np": npt, "nt": nt, "ne": ne, "precision": "fp64", "launch": launch},
        "roofline": roofline, "cpu_baseline": cpu, "split_precision": split, "batched": batched,
        "serving": served, "viscous_configs": viscous, "other_fluxes": fluxes, "by_mesh_size": sizes,
        "cl": float(hist[-1, 4]), "cd": float(hist[-1, 5]),
    }
    if stopped:
        line["stopped"] = stopped
    if served:                          # the latency part of the metric, at 90% load
        l90 = served["load_90"]
        line["p50_latency_s"] = l90["p50_latency_s"]
        line["p99_latency_s"] = l90["p99_latency_s"]
        line["latency_of"] = (f"serving at 90% of saturation, arrival to result of {l90['solves']} "
                              f"requests of 1000-4000 vertices solved to a 1e-6 residual drop: "
                              f"{l90['converged']} converged, {l90['capped']} capped at "
                              f"{served['workload'].rsplit(' ', 1)[-1]} iterations, {l90['failed']} "
                              f"stopped on a reference STOP")
    if cpu:
        line["speedup_vs_cpu"] = round(value / cpu["value"], 1)
        agg = cpu.get("aggregate", {}).get("value")
        if agg:
            line["speedup_vs_cpu_aggregate"] = round(value / agg, 1)
        if batched:
            # both ratios on one workload: the batched serving precisions
            # against the plain FP64 path on the same requests (speedup) and
            # against the host CPU (one core; all cores)
            mix = batched["mixed_48_requests"]
            mix["vs_cpu_one_core"] = round(mix["batched"] / cpu["value"], 1)
            if agg:
                mix["vs_cpu_aggregate"] = round(mix["batched"] / agg, 1)
    print(json.dumps(line), flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
