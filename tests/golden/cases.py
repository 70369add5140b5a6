"""Golden-vector cases: grid recipe + input deck + iteration count.

Shared by make_golden.py (which runs the reference on them) and by the tests
(which rebuild the same grid and inputs for the oracle and the HIP path).
"""
from __future__ import annotations

# Deck keyword order of read_input.f:20-66
DECK_KEYS = ("Start", "FlowType", "Mach", "AOA", "Reynolds", "CFL", "LastIter", "MaxIter",
             "MinRes", "SaveInt", "NoIso", "Flux", "Limiter", "FarBC", "CellType", "GridFile")


def deck(**kw) -> dict:
    d = dict(Start=1, FlowType=1, Mach=0.5, AOA=2.0, Reynolds=0.0, CFL=1.0, LastIter=0,
             MaxIter=1000, MinRes=0.0, SaveInt=1000000, NoIso=30, Flux=1, Limiter=1,
             FarBC=(0, 0.25, 0.0), CellType=1, GridFile='"grid.amdba"')
    d.update(kw)
    return d


def deck_text(d: dict) -> str:
    lines = []
    for k in DECK_KEYS:
        v = d[k]
        if isinstance(v, tuple):
            v = "  ".join(repr(x) for x in v)
        elif isinstance(v, float):
            v = repr(v)
        lines.append(f"{k:<9} {v}")
    return "\n".join(lines) + "\n"


def oracle_kwargs(d: dict, sa_compat: int = 1, nut_factor: float = 0.1) -> dict:
    """The deck as keyword arguments of oracle.Oracle / the HIP solver."""
    return dict(istart=d["Start"], iflow=d["FlowType"], mach_inf=float(d["Mach"]),
                aoa_deg=float(d["AOA"]), rey=float(d["Reynolds"]), cfl=float(d["CFL"]),
                iterlast=d["LastIter"], maxiter=d["MaxIter"], minres=float(d["MinRes"]),
                saveinterval=d["SaveInt"], iflux=d["Flux"], ilimit=d["Limiter"],
                vortex=d["FarBC"][0], xref=float(d["FarBC"][1]), yref=float(d["FarBC"][2]),
                cell_type=d["CellType"], sa_compat=sa_compat, nut_factor=nut_factor)


G0 = dict(ni=129, nj=33, t=0.12, first_spacing=1e-3, seed=12345)        # 4 224 vertices
GS = dict(ni=65, nj=17, t=0.12, first_spacing=2e-3, seed=7, jitter=0.1)  # 1 024 vertices
GV = dict(ni=97, nj=25, t=0.12, first_spacing=2e-3, seed=11)            # 2 304 vertices

CASES = {
    # the CPU-runnable workload: Euler, Roe, limited MUSCL, CFL 1, 1000 iterations
    "euler_roe_lim_129x33": dict(grid=G0, deck=deck(Mach=0.5, AOA=2.0, Flux=1, Limiter=1),
                                 iters=1000),
    "euler_kfvs_lim_129x33": dict(grid=G0, deck=deck(Mach=0.5, AOA=2.0, Flux=2, Limiter=1),
                                  geometry=False, iters=60),
    # the reference's HLLC leaves the flux in the edge-normal frame
    # (hllc_flux.f:105-154) and stops on a negative pressure in iteration 1
    "euler_hllc_lim_129x33": dict(grid=G0, deck=deck(Mach=0.5, AOA=2.0, Flux=3, Limiter=1),
                                  geometry=False, iters=60),
    # restarts (Start 2, initialize.f:91-103) from the Roe solution: residuals
    # of every flux on a developed, non-uniform state
    "restart_hllc_129x33": dict(grid=G0, deck=deck(Start=2, Mach=0.5, AOA=2.0, Flux=3, Limiter=1),
                                geometry=False, iters=1, sol_from="euler_roe_lim_129x33"),
    "restart_kfvs_129x33": dict(grid=G0, deck=deck(Start=2, Mach=0.5, AOA=2.0, Flux=2, Limiter=1),
                                geometry=False, iters=20, sol_from="euler_roe_lim_129x33"),
    "restart_roe_nolim_129x33": dict(grid=G0, deck=deck(Start=2, Mach=0.5, AOA=2.0, Flux=1,
                                                        Limiter=0),
                                     geometry=False, iters=20, sol_from="euler_roe_lim_129x33"),
    "euler_roe_nolim_65x17": dict(grid=GS, deck=deck(Mach=0.8, AOA=1.25, Flux=1, Limiter=0),
                                  iters=60),
    "euler_roe_transonic_65x17": dict(grid=GS, deck=deck(Mach=0.8, AOA=1.25, Flux=1, Limiter=1),
                                      geometry=False, iters=100),
    "euler_roe_vortex_65x17": dict(grid=GS, deck=deck(Mach=0.6, AOA=3.0, Flux=1, Limiter=1,
                                                      FarBC=(1, 0.25, 0.0)), geometry=False, iters=60),
    "euler_roe_barth_65x17": dict(grid=GS, deck=deck(Mach=0.5, AOA=2.0, Flux=1, Limiter=1,
                                                     CellType=2), iters=40),
    "laminar_roe_97x25": dict(grid=GV, deck=deck(FlowType=2, Mach=0.5, AOA=0.0, Reynolds=5000.0,
                                                 Flux=1, Limiter=1, CFL=0.8), iters=60),
    "laminar_kfvs_97x25": dict(grid=GV, deck=deck(FlowType=2, Mach=0.5, AOA=1.0, Reynolds=5000.0,
                                                  Flux=2, Limiter=1, CFL=0.8), geometry=False, iters=30),
    "rans_roe_97x25": dict(grid=GV, deck=deck(FlowType=3, Mach=0.3, AOA=4.0, Reynolds=1.0e6,
                                              Flux=1, Limiter=1, CFL=0.8), iters=40),
}
