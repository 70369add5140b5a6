"""Write the golden vectors under tests/golden/ from the reference itself.

Runs oracle/_ref/refdump (the reference's own Fortran routines compiled with
flang, see oracle/Makefile, linked to the dump harness oracle/ref_dump.f) on
seeded NACA O-grids and stores what it writes -- dual-mesh geometry, the
residual at the initial and final states, states after the first and last
iteration, the full-precision (iter, fres, iresi, fresi, cl, cd) history,
and the text of the WALL.DAT and PARAM.DAT the reference writes at the end --
as compressed .npz fixtures.  Needs /root/reference, so it only
runs in the development container:

    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from fvm2d_amd import mesh  # noqa: E402
from tests.golden.cases import CASES, deck_text  # noqa: E402

REFDUMP = os.path.join(ROOT, "oracle", "_ref", "refdump")
OUT = os.path.dirname(os.path.abspath(__file__))

INT_NAMES = {"dims", "elem", "edge", "edneigh", "iedge", "beindx"}
# dual-mesh arrays; stored once per grid (the first case that uses it)
GEOMETRY = {"elem", "edge", "edneigh", "iedge", "beindx", "ds", "dsb", "elarea", "cvarea",
            "mcarea", "drmin", "wd"}


def write_sol(path: str, prim: np.ndarray, nut: np.ndarray) -> None:
    """SOL restart file (finalize.f:14-24): conserved variables and nu_tilde,
    one vertex per line, 17 significant digits so that the read is exact."""
    r, u, v, p = prim.reshape(-1, 4).T
    g1 = 1.4 - 1.0
    cons = np.stack([r, r * u, r * v, p / g1 + 0.5 * r * (u**2 + v**2), nut], axis=1)
    with open(path, "w") as f:
        f.writelines(" ".join(f"{x:.17e}" for x in row) + "\n" for row in cons.tolist())


def run_case(name: str, case: dict) -> dict:
    g = mesh.o_grid(**case["grid"])
    with tempfile.TemporaryDirectory() as td:
        mesh.write_amdba(os.path.join(td, "grid.amdba"), g)
        with open(os.path.join(td, "fvm.inp"), "w") as f:
            f.write(deck_text(case["deck"]))
        if "sol_from" in case:
            src = np.load(os.path.join(OUT, case["sol_from"] + ".npz"))
            write_sol(os.path.join(td, "SOL"), src["primN"], src["nutN"])
        r = subprocess.run([REFDUMP, "fvm.inp", str(case["iters"]), "."], cwd=td,
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: refdump failed\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        # a reference STOP ends the run with status 0; record where it stopped
        stop_point = 0
        m = re.search(r"negative at point\s+(\d+)", r.stdout)
        if m:
            stop_point = int(m.group(1))
        out = {"stop": np.array([stop_point], dtype=np.int64)}
        for fn in sorted(os.listdir(td)):
            if not fn.endswith(".bin"):
                continue
            key = fn[:-4]
            dt = np.int32 if key in INT_NAMES else np.float64
            out[key] = np.fromfile(os.path.join(td, fn), dtype=dt)
        out["hist"] = out.get("hist", np.zeros(0)).reshape(-1, 6)
        # the reference's WALL.DAT and PARAM.DAT text (result.f:92-118,
        # finalize.f:26-61), as bytes
        for fn, key in (("WALL.DAT", "wall_text"), ("PARAM.DAT", "param_text")):
            path = os.path.join(td, fn)
            if os.path.exists(path):
                out[key] = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
    return out


def main(names=None, out_dir=OUT):
    if not os.path.exists(REFDUMP):
        raise SystemExit("oracle/_ref/refdump is not built (make -C oracle ref)")
    for name, case in CASES.items():
        if names and name not in names:
            continue
        data = run_case(name, case)
        if not case.get("geometry", True):
            data = {k: v for k, v in data.items() if k not in GEOMETRY}
        path = os.path.join(out_dir, f"{name}.npz")
        np.savez_compressed(path, **data)
        print(f"{name}: {len(data)} arrays, {os.path.getsize(path) / 1024:.0f} KiB, "
              f"{data['hist'].shape[0]} iterations")


if __name__ == "__main__":
    main(sys.argv[1:])
