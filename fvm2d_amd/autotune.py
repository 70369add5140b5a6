"""Autotuner for the launch shape of the stage kernel.

The knobs: the workgroup size of the stage kernel (256 threads, or 512 for
the kinds compiled with that instance: Euler with the limiter, any flux), the
patch size
(vertices per workgroup tile, which sets the number of workgroups, the LDS
per workgroup and the halo overhead) and the XCD-aware ordering of
workgroups.  Candidates come from an occupancy model (whole rounds of
resident workgroups over the device's CUs), each is timed on the actual mesh
and flow kind over graph-replayed iterations, and the winner is cached per
(device, mesh size, flow kind, precision).

The choice changes no result except the last bits of the L2 residual norm
(its partial sums follow the workgroups); states, CL and CD are the same
for every launch shape.
"""
from __future__ import annotations

import dataclasses
import json
import math
import os
import time

CACHE_ENV = "FVM2D_TUNE_CACHE"


def default_cache_path() -> str:
    return os.environ.get(CACHE_ENV) or os.path.join(
        os.path.expanduser("~"), ".cache", "fvm2d_amd", "tune.json")


def candidates(np_: int, cus: int = 256, per_cu: int = 4, lo: int = 24, hi: int = 128) -> list[int]:
    """Patch sizes that fill k whole rounds of cus*per_cu workgroups for the
    round counts around (lo + hi) / 2.375 vertices a patch (64 for 256
    threads), plus a few fixed sizes in [lo, hi]."""
    slots = cus * per_cu
    mid = (lo + hi) / 2.375
    k0 = max(1, round(np_ / (mid * slots)))
    out = set()
    for k in range(max(1, k0 - 1), k0 + 3):
        out.add(int(min(hi, max(lo, math.ceil(np_ / (k * slots))))))
    for ps in (2 * lo, round(mid), 4 * lo, hi):
        if lo <= ps <= hi and (ps * 4 <= max(np_, 4 * lo) or ps == lo):
            out.add(ps)
    return sorted(out)


@dataclasses.dataclass
class Choice:
    patch_size: int
    xcd: int
    us_per_iter: float
    table: list            # (threads, patch_size, xcd, us_per_iter) of every candidate
    threads: int = 256     # workgroup size of the stage kernel


def block_sizes(inputs) -> list[int]:
    """Workgroup sizes the stage kernel has an instance for, for this kind."""
    wide = inputs.iflow == 1 and inputs.ilimit == 1
    return [256, 512] if wide else [256]


def shapes(np_: int, inputs, cus: int = 256) -> list[tuple[int, int]]:
    """(threads, patch size) candidates: per_cu workgroups of 256 threads a
    CU (4 inviscid, 3 viscous), half as many of 512 with patches twice as
    large."""
    per_cu = 3 if inputs.iflow != 1 else 4
    out = [(256, ps) for ps in candidates(np_, cus=cus, per_cu=per_cu)]
    if 512 in block_sizes(inputs):
        out += [(512, ps) for ps in candidates(np_, cus=cus, per_cu=per_cu // 2, lo=48, hi=256)]
    return out


def key_of(dims: dict, inputs, device: str = "") -> str:
    return (f"{device}|np={dims['np']}|nt={dims['nt']}|flow={inputs.iflow}|flux={inputs.iflux}"
            f"|lim={inputs.ilimit}|prec={inputs.precision}")


def load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(path: str, cache: dict) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def time_shape(grid, geo, inputs, patch_size: int, xcd: int, warmup: int = 8, iters: int = 48,
               threads: int = 256) -> float:
    """Microseconds per iteration of one launch shape (graph replay)."""
    from .solver import Solver
    inp = dataclasses.replace(inputs, patch_size=patch_size, maxiter=10**9, minres=0.0,
                              block_threads=threads)
    with Solver(grid, inp, geometry=geo) as s:
        s.tune(3, 64, xcd)
        s.iterate(warmup)
        t0 = time.perf_counter()
        s.iterate(iters)
        return 1e6 * (time.perf_counter() - t0) / iters


def tune(grid, geo, inputs, cache_path: str | None = None, device: str = "", timer=None,
         iters: int = 48, use_cache: bool = True) -> Choice:
    """Best (patch size, XCD ordering) for this mesh and flow kind, timed on
    the GPU unless the cache already holds it."""
    path = cache_path or default_cache_path()
    key = key_of(geo.dims, inputs, device)
    cache = load_cache(path) if use_cache else {}
    if key in cache and "threads" in cache[key]:
        c = cache[key]
        return Choice(c["patch_size"], c["xcd"], c["us_per_iter"], c["table"], c["threads"])
    timer = timer or (lambda th, ps, x: time_shape(grid, geo, inputs, ps, x, iters=iters, threads=th))
    table = []
    for th, ps in shapes(geo.dims["np"], inputs):
        for x in (1, 0):
            table.append((th, ps, x, float(timer(th, ps, x))))
    best = min(table, key=lambda r: r[3])
    choice = Choice(best[1], best[2], best[3], table, best[0])
    cache[key] = dataclasses.asdict(choice)
    if use_cache:
        save_cache(path, cache)
    return choice
