"""NACA 4-digit O-grid generator and the amdba grid format.

The reference ships no grids (src/README:155-157); it reads the "amdba"
ASCII format in read_grid.f:19-39:

    np nt
    i  x  y  ptype          (np lines; ptype 0 interior, 3 solid, 5 farfield)
    i  n1 n2 n3 type        (nt lines)

This module builds the O-grids the benchmark workloads are defined on:

* surface: NACA 4-digit section, closed trailing edge
  (y_t = 5t(0.2969 sqrt(x) - 0.1260 x - 0.3516 x^2 + 0.2843 x^3 - 0.1036 x^4)),
  camber line for m > 0, chordwise cosine clustering; the wrap-around index
  i = 0..NI-2 starts at the trailing edge and runs along the lower surface;
* radial lines: straight, from each surface point to a circle of radius
  ``r_far`` chords centred at (0.5, 0);
* radial spacing: geometric growth from ``first_spacing`` (absolute, in
  chords).  With ``ratio`` given, that ratio is used from the wall outwards
  for as many cells as fit, and the remaining cells grow with a smaller
  constant ratio that makes the line end exactly on the far circle;
* optional jitter of interior vertices by seeded uniform +-``jitter`` of the
  local i- and j-spacing;
* every quad is split along a diagonal picked by a seeded coin flip.

Random numbers come from splitmix64 (Steele, Lea, Flood 2014), evaluated in
closed form so that the grid is the same on every machine.

Vertex numbering is 1-based, ``id = j*(NI-1) + i + 1`` (surface ring first),
and triangles are written counter-clockwise.
"""
from __future__ import annotations

import dataclasses
import math
import numpy as np

SOLID = 3
FARFIELD = 5
INTERIOR = 0

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Values n outputs of splitmix64 started at ``seed``, skipping ``offset``."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n doubles in [0, 1) from the top 53 bits of splitmix64."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def naca4_surface(m: float, p: float, t: float, n: int) -> np.ndarray:
    """n distinct surface points (closed loop), index 0 at the trailing edge,
    then the lower surface to the leading edge and the upper surface back."""
    s = 2.0 * np.pi * np.arange(n) / n
    xc = 0.5 * (1.0 + np.cos(s))
    yt = 5.0 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc**2
                    + 0.2843 * xc**3 - 0.1036 * xc**4)
    yt[0] = 0.0                      # closed trailing edge, exactly
    if m > 0.0:
        yc = np.where(xc < p, m / p**2 * (2 * p * xc - xc**2),
                      m / (1 - p)**2 * ((1 - 2 * p) + 2 * p * xc - xc**2))
        dyc = np.where(xc < p, 2 * m / p**2 * (p - xc), 2 * m / (1 - p)**2 * (p - xc))
    else:
        yc = np.zeros_like(xc)
        dyc = np.zeros_like(xc)
    th = np.arctan(dyc)
    lower = s <= np.pi
    sign = np.where(lower, -1.0, 1.0)
    x = xc - sign * yt * np.sin(th)
    y = yc + sign * yt * np.cos(th)
    return np.stack([x, y], axis=1)


def _geometric_total(d1, r, n):
    return np.where(np.abs(r - 1.0) < 1e-14, d1 * n, d1 * (r**n - 1.0) / (r - 1.0))


def _solve_ratio(d1: float, n: int, L: np.ndarray) -> np.ndarray:
    """Ratio r with d1*(r^n-1)/(r-1) = L, by bisection (vectorised over L)."""
    lo = np.full_like(L, 1e-6)
    hi = np.full_like(L, 4.0)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        f = _geometric_total(d1, mid, n) - L
        hi = np.where(f > 0, mid, hi)
        lo = np.where(f > 0, lo, mid)
    return 0.5 * (lo + hi)


def radial_fractions(L: np.ndarray, n: int, d1: float, ratio: float | None) -> np.ndarray:
    """Normalised distances xi[i, j], j = 0..n, of the radial points on each
    line of length L[i]: xi[:, 0] = 0, xi[:, n] = 1."""
    L = np.asarray(L, dtype=np.float64)
    if ratio is None:
        # one ratio per line, all lines at once
        r = _solve_ratio(d1, n, L)
        h = d1 * r[:, None] ** np.arange(n)[None, :]
        c = np.concatenate([np.zeros((L.size, 1)), np.cumsum(h, axis=1)], axis=1)
        return c / c[:, -1:]
    xi = np.zeros((L.size, n + 1))
    for i, Li in enumerate(L):
        # wall cells at the given ratio, as many as leave room for the rest
        # to grow (ratio >= 1) up to the far circle
        m = 0
        for k in range(1, n):
            head = d1 * (ratio**k - 1.0) / (ratio - 1.0)
            last = d1 * ratio ** (k - 1)
            if head + (n - k) * last * ratio <= Li:
                m = k
            else:
                break
        head_h = d1 * ratio ** np.arange(m)
        rest = Li - head_h.sum()
        last = head_h[-1] if m > 0 else d1 / ratio
        # spacing continues from the last wall cell with ratio r2:
        # last*r2 + last*r2^2 + ... = rest
        lo, hi = 1.0, ratio
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            tot = last * mid * (mid ** (n - m) - 1.0) / (mid - 1.0) if mid != 1.0 else last * (n - m)
            if tot > rest:
                hi = mid
            else:
                lo = mid
        r2 = 0.5 * (lo + hi)
        tail_h = last * r2 ** np.arange(1, n - m + 1)
        h = np.concatenate([head_h, tail_h])
        c = np.concatenate([[0.0], np.cumsum(h)])
        xi[i] = c / c[-1]
    return xi


@dataclasses.dataclass
class Grid:
    """A triangle grid in the reference's layout (1-based connectivity)."""
    coord: np.ndarray   # (np, 2) float64
    elem: np.ndarray    # (nt, 3) int32, 1-based, counter-clockwise
    ptype: np.ndarray   # (np,) int32
    ni: int = 0
    nj: int = 0

    @property
    def np(self) -> int:
        return self.coord.shape[0]

    @property
    def nt(self) -> int:
        return self.elem.shape[0]


def o_grid(ni: int, nj: int, *, m: float = 0.0, p: float = 0.4, t: float = 0.12,
           r_far: float = 20.0, first_spacing: float = 1e-3, ratio: float | None = None,
           jitter: float = 0.0, seed: int = 0) -> Grid:
    """NACA 4-digit O-grid with ``ni`` wrap-around points (first == last, so
    ni-1 distinct) and ``nj`` radial points, split into triangles."""
    nc = ni - 1
    surf = naca4_surface(m, p, t, nc)
    s = 2.0 * np.pi * np.arange(nc) / nc
    outer = np.stack([0.5 + r_far * np.cos(s), -r_far * np.sin(s)], axis=1)
    L = np.linalg.norm(outer - surf, axis=1)
    xi = radial_fractions(L, nj - 1, first_spacing, ratio)          # (nc, nj)
    P = surf[:, None, :] + xi[:, :, None] * (outer - surf)[:, None, :]   # (nc, nj, 2)

    if jitter > 0.0 and nj > 2:
        u = uniform01(seed ^ 0x4A17735EED, 2 * nc * (nj - 2)).reshape(nj - 2, nc, 2)
        a = (2.0 * u[..., 0] - 1.0) * jitter
        b = (2.0 * u[..., 1] - 1.0) * jitter
        Pi = P[:, 1:-1, :]
        di = 0.5 * (np.roll(P, -1, axis=0) - np.roll(P, 1, axis=0))[:, 1:-1, :]
        dj = 0.5 * (P[:, 2:, :] - P[:, :-2, :])
        P = P.copy()
        P[:, 1:-1, :] = Pi + a.T[:, :, None] * di + b.T[:, :, None] * dj

    coord = P.transpose(1, 0, 2).reshape(-1, 2).copy()   # id = j*nc + i
    ptype = np.zeros(nc * nj, dtype=np.int32)
    ptype[:nc] = SOLID
    ptype[(nj - 1) * nc:] = FARFIELD

    jj, ii = np.meshgrid(np.arange(nj - 1), np.arange(nc), indexing="ij")
    ii = ii.ravel(); jj = jj.ravel()
    A = jj * nc + ii
    B = jj * nc + (ii + 1) % nc
    C = (jj + 1) * nc + (ii + 1) % nc
    D = (jj + 1) * nc + ii
    flip = (splitmix64(seed, ii.size) >> np.uint64(63)).astype(bool)
    t1 = np.where(flip[:, None], np.stack([A, B, D], 1), np.stack([A, B, C], 1))
    t2 = np.where(flip[:, None], np.stack([B, C, D], 1), np.stack([A, C, D], 1))
    tri = np.empty((2 * ii.size, 3), dtype=np.int64)
    tri[0::2] = t1
    tri[1::2] = t2
    x = coord[tri]
    cross = ((x[:, 1, 0] - x[:, 0, 0]) * (x[:, 2, 1] - x[:, 1, 1])
             - (x[:, 2, 0] - x[:, 1, 0]) * (x[:, 1, 1] - x[:, 0, 1]))
    cw = cross < 0
    tri[cw, 1], tri[cw, 2] = tri[cw, 2].copy(), tri[cw, 1].copy()
    return Grid(coord=coord, elem=(tri + 1).astype(np.int32), ptype=ptype, ni=ni, nj=nj)


# Grids of the benchmark workloads ------------------------------------------------

def naca0012_129x33(seed: int = 12345) -> Grid:
    """4 224 vertices, 8 192 triangles, first spacing 1e-3 chord."""
    return o_grid(129, 33, t=0.12, first_spacing=1e-3, seed=seed)


def naca0012_513x129(seed: int = 513129) -> Grid:
    """66 048 vertices, 131 072 triangles, interior jitter +-10 %."""
    return o_grid(513, 129, t=0.12, first_spacing=1e-3, jitter=0.10, seed=seed)


def naca0012_1025x257(seed: int = 1025257) -> Grid:
    """263 168 vertices, wall spacing 1e-4, ratio 1.1 (laminar)."""
    return o_grid(1025, 257, t=0.12, first_spacing=1e-4, ratio=1.1, seed=seed)


def naca0012_2049x513(seed: int = 2049513) -> Grid:
    """1 050 624 vertices, wall spacing 4e-6 (y+ ~ 1 at Re 6e6), ratio 1.15."""
    return o_grid(2049, 513, t=0.12, first_spacing=4e-6, ratio=1.15, seed=seed)


# amdba I/O (read_grid.f:19-39) ---------------------------------------------------

def write_amdba(path: str, g: Grid) -> None:
    with open(path, "w") as f:
        f.write(f"{g.np} {g.nt}\n")
        xs = g.coord.tolist()
        ts = g.ptype.tolist()
        f.writelines(f"{k + 1} {x!r} {y!r} {t}\n" for k, ((x, y), t) in enumerate(zip(xs, ts)))
        f.writelines(f"{k + 1} {a} {b} {c} 1\n" for k, (a, b, c) in enumerate(g.elem.tolist()))


def read_amdba(path: str) -> Grid:
    with open(path) as f:
        npt, nt = (int(v) for v in f.readline().split()[:2])
        pts = np.loadtxt(f, max_rows=npt, ndmin=2)
        els = np.loadtxt(f, max_rows=nt, dtype=np.int64, ndmin=2)
    return Grid(coord=pts[:, 1:3].astype(np.float64).copy(),
                elem=els[:, 1:4].astype(np.int32).copy(),
                ptype=pts[:, 3].astype(np.int32).copy())


def dims_of(ni: int, nj: int) -> dict:
    """Closed-form counts for an ni x nj O-grid (annulus: V - E + F = 0)."""
    nc = ni - 1
    npt = nc * nj
    nt = 2 * nc * (nj - 1)
    return {"np": npt, "nt": nt, "netot": npt + nt, "nsp": nc, "nfp": nc}
