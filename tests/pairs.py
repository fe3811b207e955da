"""Identical seeded set-ups on the GPU library and on the oracle."""
import numpy as np


def domain(kind):
    """(xmin, xmax, xperiodic, yperiodic) of the test domains"""
    if kind == "periodic":
        return (0.0, 0.0), (1.0, 1.0), True, True
    if kind == "walls":
        return (0.0, 0.0), (1.0, 1.0), False, False
    if kind == "xperiodic":
        return (0.0, 0.0), (1.0, 0.5), True, False
    if kind == "tall":
        return (0.0, 0.0), (1.0, 2.0), False, False
    raise ValueError(kind)


def lattice(lv, kind, K, seed):
    xmin, xmax, xp, yp = domain(kind)
    dr = (xmax[0] - xmin[0]) / K
    ny = int(round((xmax[1] - xmin[1]) / dr))
    x = lv.inputs.jittered_lattice(K, ny, dr, xmin=xmin, seed=seed)
    return x, dr


def make_pair(lv, oracle, kind="periodic", K=32, seed=0x5EED0001, ic="tg", remesh=True):
    xmin, xmax, xp, yp = domain(kind)
    x, dr = lattice(lv, kind, K, seed)
    rect = lv.Rectangle(xmin, xmax)
    grid = lv.VoronoiGrid(rect, dr, xperiodic=xp, yperiodic=yp, positions=x)
    og = oracle.Grid(xmin, xmax, dr, xperiodic=xp, yperiodic=yp)
    op = oracle.Particles(og, x)
    if remesh:
        lv.remesh(grid)
        op.remesh()
        area = op.area()
        if ic is not None:
            fields = initial_fields(lv, ic, x, area)
            for k, v in fields.items():
                grid.set_soa(k, v)
                op.set(k, v)
    return grid, op, x, dr


def initial_fields(lv, ic, x, area):
    n = x.size // 2
    rng = np.random.default_rng(12345)
    if ic == "tg":
        f = lv.inputs.taylor_green(x, area)
        out = {k: f[k] for k in ("v", "rho", "mass", "P", "e", "mu")}
        # a non-trivial, smooth dv and D so every operator sees generic input
        out["dv"] = 0.1 * np.concatenate([np.sin(3 * x[n:]), np.cos(5 * x[:n])])
        out["D"] = 0.01 * rng.standard_normal(4 * n)
        out["c2"] = np.full(n, 2500.0) + rng.random(n)
        return out
    if ic == "cavity":
        f = lv.inputs.cavity(x, area)
        out = {k: f[k] for k in ("v", "rho", "mass", "mu", "c2")}
        out["v"] = 0.05 * rng.standard_normal(2 * n)
        out["e"] = 1.0 + rng.random(n)
        out["P"] = rng.random(n)
        return out
    raise ValueError(ic)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    if scale == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / scale)
