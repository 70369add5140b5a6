"""Shared helpers for the tests: fixtures, grids and oracle instances."""
from __future__ import annotations

import functools
import os

import numpy as np

from fvm2d_amd import mesh
from tests.golden.cases import CASES, oracle_kwargs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@functools.lru_cache(maxsize=None)
def fixture(name: str) -> dict:
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


@functools.lru_cache(maxsize=None)
def grid(name: str) -> mesh.Grid:
    return mesh.o_grid(**CASES[name]["grid"])


def case_kwargs(name: str, **over) -> dict:
    kw = oracle_kwargs(CASES[name]["deck"])
    kw.update(over)
    return kw


def geometry_fixture(name: str) -> dict:
    """Geometry arrays of the case's grid (stored with the first case on it)."""
    g = CASES[name]["grid"]
    for other, c in CASES.items():
        if c["grid"] == g and c.get("geometry", True):
            return fixture(other)
    raise KeyError(name)


def make_oracle(name: str, **over):
    from oracle.oracle import Oracle
    g = grid(name)
    o = Oracle(g.coord, g.elem, g.ptype, **case_kwargs(name, **over))
    fx = fixture(name)
    if CASES[name]["deck"]["Start"] == 2:
        o.set("prim", fx["prim0"])
        o.set("nut", fx["nut0"])
    return o


def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)
