"""The initialisers of src/populate.jl: seeds on concentric circles, a Cartesian grid, random
points, a Vogel spiral, a hexagonal lattice, and Lloyd-relaxed random points.

Each builds the seed positions on the host in the reference's order, keeps those with
charfun(x) and isinside(boundary_rect, x) (closed rectangle, src/geometry.jl:127-129), puts them
on the device and remeshes.  populate_lloyd's relaxation runs on the device (lv_lloyd).

Two adapters stand in for Julia idioms:
  * ic(grid) is called once after the last remesh, where the reference calls ic!(p) per polygon;
    it sets the fields of all polygons (grid.set / grid.set_soa, reading grid.get("area") etc.);
  * the random points of populate_rand / populate_lloyd come from a seeded generator (rand() in
    the reference is unseeded).
"""
import math

import numpy as np

from . import _check, lib, remesh


def _everywhere(x):
    return True


def _inside(rect, x1, x2):
    return rect.xmin[0] <= x1 <= rect.xmax[0] and rect.xmin[1] <= x2 <= rect.xmax[1]


def _place(grid, pts, charfun, ic):
    keep = [(a, b) for a, b in pts if charfun((a, b)) and _inside(grid.boundary_rect, a, b)]
    x = np.array(keep, dtype=np.float64).reshape(-1, 2)
    grid.reserve(x.shape[0])
    grid.set_positions(np.concatenate([x[:, 0], x[:, 1]]))
    remesh(grid)
    if ic is not None:
        ic(grid)
    return grid


def _verts(rect):
    (a, b), (c, d) = rect.xmin, rect.xmax
    return [(a, b), (c, b), (c, d), (a, d)]


def _step_range(start, step, stop):
    """start:step:stop of Float64s (the values start + k*step that do not pass stop)"""
    n = int(math.floor((stop - start) / step + 1e-10)) + 1
    return [start + k * step for k in range(max(n, 0))]


def _julia_range(a, b, n):
    """range(a, b, n): n values from a to b inclusive"""
    if n == 1:
        return [a]
    return [a + (b - a) * k / (n - 1) for k in range(n)]


def populate_circ(grid, charfun=_everywhere, center=(0.0, 0.0), ic=None):
    """populate_circ!(grid; charfun, center, ic!), src/populate.jl:9-35"""
    r_max = max(math.hypot(v[0] - center[0], v[1] - center[1]) for v in _verts(grid.boundary_rect))
    pts = []
    for r in _step_range(0.5 * grid.dr, grid.dr, r_max):
        k_max = round(2.0 * math.pi * r / grid.dr)
        for k in range(1, k_max + 1):
            theta = 2.0 * math.pi * k / k_max
            pts.append((center[0] + r * math.cos(theta), center[1] + r * math.sin(theta)))
    return _place(grid, pts, charfun, ic)


def populate_rect(grid, charfun=_everywhere, ic=None):
    """populate_rect!(grid; charfun, ic!), src/populate.jl:37-62.  As in the reference, the N values of
    range(xmin, xmax, N) include xmax, so after the dr/2 shift the last row and column fall outside
    the rectangle: (N-1) x (M-1) seeds at spacing (xmax - xmin)/(N - 1)."""
    r = grid.boundary_rect
    N = round((r.xmax[0] - r.xmin[0]) / grid.dr)
    M = round((r.xmax[1] - r.xmin[1]) / grid.dr)
    pts = [(x1 + 0.5 * grid.dr, x2 + 0.5 * grid.dr) for x1 in _julia_range(r.xmin[0], r.xmax[0], N)
           for x2 in _julia_range(r.xmin[1], r.xmax[1], M)]
    return _place(grid, pts, charfun, ic)


def _rand_points(grid, seed):
    r = grid.boundary_rect
    N = round(abs(r.xmax[0] - r.xmin[0]) * abs(r.xmax[1] - r.xmin[1]) / grid.dr ** 2)
    rng = np.random.default_rng(seed)
    s = rng.random((N, 2))
    return [(s1 * r.xmax[0] + (1 - s1) * r.xmin[0], s2 * r.xmax[1] + (1 - s2) * r.xmin[1]) for s1, s2 in s]


def populate_rand(grid, charfun=_everywhere, ic=None, seed=0x5EED0800):
    """populate_rand!(grid; charfun, ic!), src/populate.jl:64-90"""
    return _place(grid, _rand_points(grid, seed), charfun, ic)


def populate_vogel(grid, charfun=_everywhere, center=(0.0, 0.0), ic=None):
    """populate_vogel!(grid; charfun, center, ic!), src/populate.jl:92-117"""
    r_max = max(math.hypot(v[0] - center[0], v[1] - center[1]) for v in _verts(grid.boundary_rect))
    N = round(math.pi * r_max * r_max / grid.dr ** 2)
    pts = []
    for i in range(1, N + 1):
        r = r_max * math.sqrt(i / N)
        theta = 2.39996322972865332 * i
        pts.append((center[0] + r * math.cos(theta), center[1] + r * math.sin(theta)))
    return _place(grid, pts, charfun, ic)


def populate_hex(grid, charfun=_everywhere, ic=None):
    """populate_hex!(grid; charfun, ic!), src/populate.jl:147-180"""
    a = (4.0 / 3.0) ** 0.25 * grid.dr
    b = (3.0 / 4.0) ** 0.25 * grid.dr
    r = grid.boundary_rect
    i_min = int(math.floor(r.xmin[0] / a)) - 1
    j_min = int(math.floor(r.xmin[1] / b))
    i_max = int(math.ceil(r.xmax[0] / a))
    j_max = int(math.ceil(r.xmax[1] / b))
    pts = [((i + int(math.fmod(j, 2)) / 2) * a, j * b) for i in range(i_min, i_max + 1)
           for j in range(j_min, j_max + 1)]
    return _place(grid, pts, charfun, ic)


def populate_lloyd(grid, charfun=_everywhere, niterations=100, ic=None, seed=0x5EED0800):
    """populate_lloyd!(grid; charfun, niterations, ic!), src/populate.jl:119-145: random seeds, then
    niterations times remesh! and x = centroid(p) on the device, then remesh! and ic"""
    populate_rand(grid, charfun=charfun, seed=seed)
    _check(grid.ctx, lib().lv_lloyd(grid.ctx, int(niterations)))
    if ic is not None:
        ic(grid)
    return grid
