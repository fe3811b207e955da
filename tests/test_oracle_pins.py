"""Pins of the CPU oracle against what the mathematics or independent libraries fix.

The oracle restates the reference (no Julia here to run it), so each restated function is checked
against something that would catch a plausible restatement error: a closed form, a brute force
over all seeds, scipy's Qhull Voronoi diagram, a conservation law of the discretisation, a dense
linear solve, and the reference test's own acceptance thresholds (tests/taylorgreen.jl:112-114).
"""
import math

import numpy as np
import pytest

import flows
from pairs import domain, lattice


def oracle_pair(lv, oracle, kind, K, seed):
    xmin, xmax, xp, yp = domain(kind)
    x, dr = lattice(lv, kind, K, seed)
    g = oracle.Grid(xmin, xmax, dr, xperiodic=xp, yperiodic=yp)
    p = oracle.Particles(g, x)
    p.remesh()
    return g, p, x, dr


# ---- magic path (src/neighborlist.jl:29-41) -------------------------------------------------------

def brute_path(n1, n2, h):
    nodes = []
    gen = 0
    for i2 in range(1 - n1, n1):          # the reference's n1/n2 swap, kept on purpose
        for i1 in range(1 - n2, n2):
            k = max(0, abs(i1) - 1) ** 2 + max(0, abs(i2) - 1) ** 2
            nodes.append(((h * h) * float(k), i1 * i1 + i2 * i2, gen, i1, i2))
            gen += 1
    nodes.sort(key=lambda t: (t[0], t[1], t[2]))
    return nodes


@pytest.mark.parametrize("xmax,dr", [((1.0, 1.0), 1.0 / 20), ((1.0, 0.25), 1.0 / 24), ((0.3, 1.0), 0.05)])
def test_magic_path_equals_full_sorted_path(oracle, xmax, dr):
    g = oracle.Grid((0.0, 0.0), xmax, dr)
    full = brute_path(g.n1, g.n2, 2.0 * dr)
    i1, i2, rr = g.magic_path()
    rr_max = (10.0 * dr) * (10.0 * dr)
    prefix = [t for t in full if not t[0] > rr_max]
    assert len(prefix) == g.npath
    assert [t[3] for t in prefix] == list(i1) and [t[4] for t in prefix] == list(i2)
    assert np.array_equal(rr, np.array([t[0] for t in prefix]))
    nxt = full[len(prefix)][0] if len(full) > len(prefix) else -1.0
    assert g.next_rr == nxt


def test_magic_path_closed_form(oracle):
    g = oracle.Grid((0.0, 0.0), (1.0, 1.0), 1.0 / 64)
    i1, i2, rr = g.magic_path()
    first9 = list(zip(i1[:9], i2[:9]))
    assert first9 == [(0, 0), (0, -1), (-1, 0), (1, 0), (0, 1), (-1, -1), (1, -1), (-1, 1), (1, 1)]
    assert np.all(rr[:9] == 0.0)
    assert g.npath == 129  # h = 2dr, r_max = 10dr: offsets with max(0,|i|-1)^2 sums <= 25
    assert np.all(np.diff(rr) >= 0)


# ---- cell list (src/neighborlist.jl:47-67, src/voronoigrid.jl:130-147) -----------------------------

@pytest.mark.parametrize("kind", ["periodic", "walls", "xperiodic"])
def test_cell_list_brute_force(lv, oracle, kind):
    xmin, xmax, xp, yp = domain(kind)
    x, dr = lattice(lv, kind, 24, 0x5EED0100)
    g = oracle.Grid(xmin, xmax, dr, xperiodic=xp, yperiodic=yp)
    cs, cl = g.cell_list(x)
    n = x.size // 2
    h = 2.0 * dr
    origin = (g.cropping_min[0] - h, g.cropping_min[1] - h)
    Px, Py = xmax[0] - xmin[0], xmax[1] - xmin[1]
    expect = {}
    for i in range(n):
        x1, x2 = x[i], x[n + i]
        imgs = [(x1, x2)]
        if xp:
            imgs += [(x1 + Px, x2), (x1 - Px, x2)]
        if yp:
            imgs += [(x1, x2 + Py), (x1, x2 - Py)]
        if xp and yp:
            imgs += [(x1 + Px, x2 + Py), (x1 + Px, x2 - Py), (x1 - Px, x2 + Py), (x1 - Px, x2 - Py)]
        for a, b in imgs:
            k1 = math.floor((a - origin[0]) / h) + 1
            k2 = math.floor((b - origin[1]) / h) + 1
            if 1 <= k1 <= g.n1 and 1 <= k2 <= g.n2:
                expect.setdefault((k1 - 1) + (k2 - 1) * g.n1, []).append(i)
    for c in range(g.n1 * g.n2):
        got = list(cl[cs[c]:cs[c + 1]])
        assert got == expect.get(c, []), c
        assert got == sorted(got)


# ---- Voronoi clipping (src/polygon.jl, src/voronoigrid.jl:53-108) ---------------------------------

def polygon_area(verts):
    v = np.asarray(verts)
    c = v.mean(axis=0)
    ang = np.arctan2(v[:, 1] - c[1], v[:, 0] - c[0])
    v = v[np.argsort(ang)]
    return 0.5 * abs(np.dot(v[:, 0], np.roll(v[:, 1], -1)) - np.dot(v[:, 1], np.roll(v[:, 0], -1)))


def test_periodic_mesh_matches_qhull(lv, oracle):
    from scipy.spatial import Voronoi

    g, p, x, dr = oracle_pair(lv, oracle, "periodic", 24, 0x5EED0101)
    n = p.n
    pts = np.stack([x[:n], x[n:]], axis=1)
    tiles = [pts + np.array([sx, sy]) for sx in (-1, 0, 1) for sy in (-1, 0, 1)]
    allp = np.concatenate(tiles)
    center = 4 * n  # the (0, 0) tile
    vor = Voronoi(allp)
    nb = [set() for _ in range(n)]
    for a, b in vor.ridge_points:
        if center <= a < center + n:
            nb[a - center].add(b % n)
        if center <= b < center + n:
            nb[b - center].add(a % n)
    ptr, lab, ver = p.mesh()
    area = p.area()
    for i in range(n):
        mine = set(int(l) for l in lab[ptr[i]:ptr[i + 1]] if l >= 0)
        assert mine == nb[i], i
        reg = vor.regions[vor.point_region[center + i]]
        assert -1 not in reg
        assert abs(polygon_area(vor.vertices[reg]) - area[i]) <= 1e-12 * area[i]
    assert abs(area.sum() - 1.0) <= 1e-13


def clip_brute(x1, x2, pts, skip, rect):
    """Sutherland-Hodgman clip of the rectangle by the bisector of every other seed."""
    poly = [(rect[0], rect[1]), (rect[2], rect[1]), (rect[2], rect[3]), (rect[0], rect[3])]
    nbrs = set()
    for j, (y1, y2) in enumerate(pts):
        if j == skip:
            continue
        d1, d2 = y1 - x1, y2 - x2
        c = d1 * 0.5 * (y1 + x1) + d2 * 0.5 * (y2 + x2)
        f = [d1 * a + d2 * b - c for a, b in poly]
        out = []
        cut = False
        for k in range(len(poly)):
            a, b = poly[k], poly[(k + 1) % len(poly)]
            fa, fb = f[k], f[(k + 1) % len(poly)]
            if fa <= 0:
                out.append(a)
            if (fa < 0 < fb) or (fb < 0 < fa):
                t = fa / (fa - fb)
                out.append((a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])))
                cut = True
        if cut or any(v > 1e-14 for v in f):
            if len(out) < len(poly) or cut:
                nbrs.add(j)
        poly = out
    area = polygon_area(poly)
    return area, poly, nbrs


def test_wall_mesh_matches_brute_force(lv, oracle):
    g, p, x, dr = oracle_pair(lv, oracle, "walls", 14, 0x5EED0102)
    n = p.n
    pts = list(zip(x[:n], x[n:]))
    ptr, lab, ver = p.mesh()
    area = p.area()
    for i in range(n):
        a, poly, _ = clip_brute(x[i], x[n + i], pts, i, (0.0, 0.0, 1.0, 1.0))
        assert abs(a - area[i]) <= 1e-12 * area[i]
        # every neighbour label names a seed whose bisector carries that edge
        for k in range(ptr[i], ptr[i + 1]):
            j = lab[k]
            v1, v2 = ver[k, :2], ver[k, 2:]
            if j < 0:
                continue
            for v in (v1, v2):
                di = np.hypot(v[0] - x[i], v[1] - x[n + i])
                dj = np.hypot(v[0] - x[j], v[1] - x[n + j])
                assert abs(di - dj) <= 1e-12
    assert abs(area.sum() - 1.0) <= 1e-13


def test_edges_are_chained_and_symmetric(lv, oracle):
    g, p, x, dr = oracle_pair(lv, oracle, "periodic", 32, 0x5EED0103)
    ptr, lab, ver = p.mesh()
    n = p.n
    nbr = [set() for _ in range(n)]
    for i in range(n):
        e = ver[ptr[i]:ptr[i + 1]]
        # sort_edges! chains v2 of edge k to v1 of edge k+1 exactly, and the polygon closes
        assert np.array_equal(e[1:, :2], e[:-1, 2:])
        assert np.array_equal(e[0, :2], e[-1, 2:])
        nbr[i] = set(lab[ptr[i]:ptr[i + 1]])
    for i in range(n):
        for j in nbr[i]:
            assert i in nbr[j]
    deg = np.diff(ptr)
    assert abs(deg.mean() - 6.0) < 0.05


def test_hexagonal_lattice_closed_forms(lv, oracle):
    """A periodic regular hexagonal lattice: 6 neighbours, area a*b, lr_ratio 1/sqrt(3), centroid = seed."""
    dr = 1.0 / 16
    a = (4.0 / 3.0) ** 0.25 * dr
    b = (3.0 / 4.0) ** 0.25 * dr
    na, nb = 16, 18
    W, H = na * a, nb * b
    pts = [((i + (j % 2) / 2 + 0.25) * a, (j + 0.5) * b) for j in range(nb) for i in range(na)]
    x = np.concatenate([[q[0] for q in pts], [q[1] for q in pts]])
    g = oracle.Grid((0.0, 0.0), (W, H), dr, xperiodic=True, yperiodic=True)
    p = oracle.Particles(g, x)
    p.remesh()
    n = p.n
    assert np.all(p["ne"] == 6)
    assert np.allclose(p.area(), a * b, rtol=1e-12, atol=0)
    c = p.centroid()
    per = np.array([W] * n + [H] * n)
    dc = (c - x + per / 2) % per - per / 2
    assert np.max(np.abs(dc)) <= 1e-12 * dr
    p.set("rho", 1.0)
    p.set("mass", a * b)
    p.set("c2", np.inf)
    diag, coef = p.pressure_operator(1e-3)
    d, rp, cols, vals = p.csr(diag, coef)
    # c_ij = lr_ratio * (0.5/rho + 0.5/rho) = 1/sqrt(3)
    assert np.allclose(vals, 1.0 / math.sqrt(3.0), rtol=1e-12, atol=0)
    assert np.all(d == 0.0)


# ---- pressure system (src/pressure.jl:89-203) -----------------------------------------------------

def tg_state(lv, oracle, kind, K, seed):
    g, p, x, dr = oracle_pair(lv, oracle, kind, K, seed)
    ic = lv.inputs.taylor_green(x, p.area())
    for k in ("v", "rho", "mass", "P", "e", "mu"):
        p.set(k, ic[k])
    return g, p, x, dr, ic


def test_operator_symmetric_rowsum_spd(lv, oracle):
    g, p, x, dr, ic = tg_state(lv, oracle, "periodic", 12, 0x5EED0104)
    p.eos_stiffened(ic["gamma"], ic["P0"])
    dt = 1e-2
    diag, coef = p.pressure_operator(dt)
    d, rp, cols, vals = p.csr(diag, coef)
    n = p.n
    Adense = np.zeros((n, n))
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            j = cols[k]
            Adense[i, i] += vals[k]
            Adense[i, j] -= vals[k]
        Adense[i, i] += d[i]
    assert np.max(np.abs(Adense - Adense.T)) <= 1e-13 * np.max(np.abs(Adense))
    ones = np.ones(n)
    assert np.allclose(p.apply_A(diag, coef, ones), d, rtol=1e-12, atol=1e-14)
    ev = np.linalg.eigvalsh(0.5 * (Adense + Adense.T))
    assert ev.min() > 0
    # mul!(y, A, x) against the dense matrix
    rng = np.random.default_rng(7)
    v = rng.standard_normal(n)
    assert np.allclose(p.apply_A(diag, coef, v), Adense @ v, rtol=1e-12, atol=1e-12)
    # the oracle's PCG against a dense solve
    b, P0, GP = p.rhs(dt, False)
    xs = np.linalg.solve(Adense, b)
    xo, it, rr = p.pcg(diag, coef, b, P0, 1e-13)
    assert np.max(np.abs(xo - xs)) <= 1e-9 * np.max(np.abs(xs))


def test_minres_iterates_match_scipy(lv, oracle):
    """The oracle's MINRES (the reference's solver, src/pressure.jl:219) against scipy's independent
    implementation of the same Paige-Saunders recurrences: equal iterates after a fixed number of
    steps from the same warm start, and the dense solution once converged."""
    from scipy.sparse.linalg import minres as sp_minres

    g, p, x, dr, ic = tg_state(lv, oracle, "periodic", 12, 0x5EED0109)
    p.eos_stiffened(ic["gamma"], ic["P0"])
    dt = 1e-2
    diag, coef = p.pressure_operator(dt)
    d, rp, cols, vals = p.csr(diag, coef)
    n = p.n
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            A[i, i] += vals[k]
            A[i, cols[k]] -= vals[k]
        A[i, i] += d[i]
    b, P0, GP = p.rhs(dt, False)
    for k in (1, 3, 10, 25):
        xo, it, _ = p.minres(diag, coef, b, P0, 0.0, 0.0, itmax=k)
        xs, _ = sp_minres(A, b, x0=P0.copy(), rtol=0.0, maxiter=k)
        assert it == -1  # stopped by itmax
        assert np.max(np.abs(xo - xs)) <= 1e-10 * np.max(np.abs(xs)), k
    xo, it, rr = p.minres(diag, coef, b, P0, 0.0, 1e-12, itmax=2000)
    assert it > 0 and rr <= 1e-11
    assert np.max(np.abs(xo - np.linalg.solve(A, b))) <= 1e-9 * np.max(np.abs(xo))


def test_pressure_step_linear_pressure_closed_form(lv, oracle):
    """For P = g.x, every interior cell gets dv = -dt*area*g/m exactly (divergence theorem, midpoint
    rule exact on straight edges)."""
    g, p, x, dr = oracle_pair(lv, oracle, "walls", 20, 0x5EED0105)
    n = p.n
    gv = np.array([0.7, -1.3])
    p.set("P", gv[0] * x[:n] + gv[1] * x[n:])
    area = p.area()
    p.set("mass", 2.0 * area)
    p.set("v", 0.0)
    dt = 1e-3
    p.pressure_step(dt)
    v = p["v"]
    ptr, lab, _ = p.mesh()
    interior = np.array([np.all(lab[ptr[i]:ptr[i + 1]] >= 0) for i in range(n)])
    assert interior.sum() > 0.7 * n
    expect = -dt * area / (2.0 * area)
    assert np.allclose(v[:n][interior], expect[interior] * gv[0], rtol=1e-10, atol=1e-15)
    assert np.allclose(v[n:][interior], expect[interior] * gv[1], rtol=1e-10, atol=1e-15)


def test_conservation_of_fluxes(lv, oracle):
    """pressure_step! pass 2, viscous_step! pass 2 and relaxation_step! move energy/mass/momentum
    between neighbours antisymmetrically, so their totals change only by rounding."""
    g, p, x, dr, ic = tg_state(lv, oracle, "periodic", 40, 0x5EED0106)
    n = p.n
    dt = 1e-3
    p.eos_stiffened(ic["gamma"], ic["P0"])
    E0 = np.sum(p["mass"] * p["e"])
    p.pressure_step(dt)
    E1 = np.sum(p["mass"] * p["e"])
    assert abs(E1 - E0) <= 1e-13 * np.sum(np.abs(p["mass"] * p["e"]))
    p.find_D()
    p.viscous_step(dt, True)
    E2 = np.sum(p["mass"] * p["e"])
    assert abs(E2 - E1) <= 1e-13 * np.sum(np.abs(p["mass"] * p["e"]))
    p.find_dv(dt)
    M0 = p["mass"].sum()
    mom0 = np.array([np.sum(p["mass"] * p["v"][:n]), np.sum(p["mass"] * p["v"][n:])])
    E3 = np.sum(p["mass"] * p["e"])
    p.relaxation_step(dt, True, remesh=False)
    assert abs(p["mass"].sum() - M0) <= 1e-13 * M0
    mom1 = np.array([np.sum(p["mom"][:n]), np.sum(p["mom"][n:])])
    assert np.max(np.abs(mom1 - mom0)) <= 1e-13 * np.sum(np.abs(p["mass"] * p["v"][:n]))
    assert abs(np.sum(p["energy"]) - E3) <= 1e-13 * abs(E3)


def test_gp_step_vanishes_for_linear_pressure(lv, oracle):
    """GP is the exact gradient of a linear P on interior cells, so the gp_step correction
    sum lrr (GP_i - GP_j).(m - z) vanishes there."""
    g, p, x, dr = oracle_pair(lv, oracle, "walls", 20, 0x5EED0107)
    n = p.n
    gv = np.array([0.4, 0.9])
    p.set("P", gv[0] * x[:n] + gv[1] * x[n:])
    area = p.area()
    p.set("mass", area)
    p.set("rho", 1.0)
    p.set("c2", 1.0)
    b1, _, GP = p.rhs(1e-3, False)
    b2, _, _ = p.rhs(1e-3, True)
    ptr, lab, _ = p.mesh()
    interior = np.array([np.all(lab[ptr[i]:ptr[i + 1]] >= 0) for i in range(n)])
    assert np.allclose(GP[:n][interior], gv[0], rtol=1e-10)  # GP = grad P
    assert np.allclose(GP[n:][interior], gv[1], rtol=1e-10)
    # the correction only differs from zero on cells whose neighbours touch the walls
    inner2 = np.array([interior[i] and all(interior[j] for j in lab[ptr[i]:ptr[i + 1]]) for i in range(n)])
    assert np.max(np.abs((b2 - b1)[inner2])) <= 1e-9 * np.max(np.abs(b1))


# ---- move! (src/move.jl) --------------------------------------------------------------------------

def test_move_wraps_and_projects_at_walls(lv, oracle):
    g, p, x, dr = oracle_pair(lv, oracle, "xperiodic", 16, 0x5EED0108)
    n = p.n
    v = np.zeros(2 * n)
    v[:n] = 3.0     # crosses the periodic x boundary for seeds near x = 1
    v[n:] = -5.0    # drives the bottom row into the wall y = 0
    p.set("v", v)
    x0 = p["x"].copy()
    dt = 0.01
    p.move(dt, remesh=False)
    xn = p["x"]
    assert np.all((xn[:n] >= 0.0) & (xn[:n] <= 1.0))
    assert np.all((xn[n:] >= 0.0) & (xn[n:] <= 0.5))
    stuck = x0[n:] + dt * -5.0 < 0.0
    free = ~stuck
    assert stuck.any() and free.any()
    # periodic_wrap: x + dt*v taken modulo the period
    assert np.allclose(xn[:n][free], (x0[:n][free] + dt * 3.0) % 1.0, rtol=0, atol=1e-15)
    assert np.allclose(xn[n:][free], x0[n:][free] - dt * 5.0, rtol=0, atol=1e-15)
    # seeds that would leave through y = 0: the normal velocity is projected out (to rounding),
    # the tangential one kept, and the seed slides along the wall
    assert np.all(np.abs(p["v"][n:][stuck]) <= 1e-12)
    assert np.allclose(p["v"][:n][stuck], 3.0, rtol=1e-12)
    assert np.allclose(xn[n:][stuck], x0[n:][stuck], rtol=0, atol=1e-12)


def test_move_nan_velocity_raises(lv, oracle):
    g, p, x, dr = oracle_pair(lv, oracle, "periodic", 8, 1)
    v = np.zeros(2 * p.n)
    v[3] = np.nan
    p.set("v", v)
    with pytest.raises(oracle.OracleError, match="Velocity field invalidated"):
        p.move(1e-3)


# ---- the reference's own acceptance test (tests/taylorgreen.jl) -----------------------------------

def test_taylor_green_reference_thresholds(oracle):
    x = flows.populate_hex((0.0, 0.0), (1.0, 1.0), flows.TG_DR)
    g = oracle.Grid((0.0, 0.0), (1.0, 1.0), flows.TG_DR, xperiodic=True, yperiodic=True)
    p = oracle.Particles(g, x)
    p.remesh()
    for k, v in flows.tg_ic(x, p.area()).items():
        p.set(k, v)
    E0 = None
    last = None
    for k, t, frame in flows.tg_frame_schedule():
        p.step(flows.TG_DT, eos="stiffened", gamma=flows.TG_GAMMA, P0=flows.TG_P0, tol=1e-10,
               artificial_viscosity=False)
        if frame:
            E, ve, pe = flows.tg_errors(p["x"], p["v"], p["P"], p.area(), p["mass"], p["e"], t)
            if E0 is None:
                E0 = E
            last = (E - E0, ve, pe)
    E_err, v_err, P_err = last
    assert E_err < 1e-8 and v_err < 0.01 and P_err < 0.01, last


# ---- multiphase projection (src/relaxation.jl:75-206) ---------------------------------------------

def rt_particles(lv, oracle, K=24, seed=0x5EED0600):
    dr = 1.0 / K
    x = lv.inputs.jittered_lattice(K, 2 * K, dr, seed=seed)
    g = oracle.Grid((0.0, 0.0), (1.0, 2.0), dr)
    p = oracle.Particles(g, x)
    p.remesh()
    ic = lv.inputs.rayleigh_taylor(x, p.area())
    for k in ("v", "rho", "mass", "P", "e", "mu", "phase"):
        p.set(k, ic[k])
    p.find_D()
    p.find_dv(1e-3, 1.0)
    n = x.size // 2
    p.set("dv", 0.1 * np.random.default_rng(1).standard_normal(2 * n))
    return p, n


def test_multiphase_projector_is_spsd_on_the_interface(lv, oracle):
    """MultiphaseProjector: symmetric positive semi-definite, zero rows and columns off the phase
    interface, and refresh!'s right-hand side lies in its range."""
    p, n = rt_particles(lv, oracle)
    A = np.stack([p.multiphase_apply(e) for e in np.eye(n)], axis=1)
    b = p.multiphase_rhs()
    ph = p["phase"]
    iface = np.zeros(n, dtype=bool)
    ptr, lab, _ = p.mesh()
    for i in range(n):
        for j in lab[ptr[i]:ptr[i + 1]]:
            if j >= 0 and ph[j] != ph[i]:
                iface[i] = True
    assert iface.sum() > 10
    assert np.all(A[~iface] == 0.0) and np.all(A[:, ~iface] == 0.0) and np.all(b[~iface] == 0.0)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    assert ev.min() >= -1e-12 * ev.max()
    xs, *_ = np.linalg.lstsq(A, b, rcond=None)
    assert np.linalg.norm(A @ xs - b) <= 1e-10 * np.linalg.norm(b)


def test_multiphase_projection_removes_interphase_flux(lv, oracle):
    """After a tight projection without the quality cut-off, the right-hand side recomputed from
    the new dv (the flux of dv across phase boundaries) vanishes: b(dv - T(x)) = b - A x = 0."""
    p, n = rt_particles(lv, oracle)
    b0 = p.multiphase_rhs()
    it, res = p.multiphase_projection(0.0, 1e-13, 1e-13, 100000)
    assert it > 0
    b1 = p.multiphase_rhs()
    assert np.linalg.norm(b1) <= 1e-9 * np.linalg.norm(b0)


def test_multiphase_quality_threshold_skips_cells(lv, oracle):
    p, n = rt_particles(lv, oracle)
    dv0 = p["dv"].copy()
    q = p["quality"]
    thr = float(np.median(q))
    p.multiphase_projection(thr, 1e-10, 1e-10, 100000)
    changed = (p["dv"][:n] != dv0[:n]) | (p["dv"][n:] != dv0[n:])
    assert not np.any(changed & (q < thr)) and np.any(changed)


# ---- implicit viscosity (src/viscous_solver.jl:43-128) ---------------------------------------------

def test_viscous_operator_symmetric_with_mass_diagonal(lv, oracle):
    """ViscousSolver's operator: symmetric couplings 0.5 lrr (mu_i + mu_j), A 1 = m/dt, positive
    definite; the implicit step conserves momentum sum m v."""
    g, p, x, dr = oracle_pair(lv, oracle, "walls", 20, 0x5EED0702)
    n = x.size // 2
    rng = np.random.default_rng(4)
    p.set("mass", dr * dr * (1.0 + 0.1 * rng.random(n)))
    p.set("mu", 0.05 + 0.05 * rng.random(n))
    p.set("v", rng.standard_normal(2 * n))
    dt = 1e-2
    diag, coef = p.viscous_operator(dt)
    assert np.allclose(diag, p["mass"] / dt, rtol=0, atol=0)
    A = np.stack([p.apply_A(diag, coef, e) for e in np.eye(n)], axis=1)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    assert np.abs(A @ np.ones(n) - diag).max() <= 1e-12 * diag.max()
    assert np.linalg.eigvalsh(A).min() > 0.0
    mom0 = (p["mass"][None, :] * p["v"].reshape(2, n)).sum(axis=1)
    assert p.viscous_implicit(dt, 1e-13, 100000) >= 0
    mom1 = (p["mass"][None, :] * p["v"].reshape(2, n)).sum(axis=1)
    assert np.abs(mom1 - mom0).max() <= 1e-11 * np.abs(p["mass"] * p["v"][:n]).sum()
