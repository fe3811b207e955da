"""Independent pins of the oracle's stencil physics: find_D!, getS with the Stone-Norman term,
viscous_step!, bdary_friction!, find_dv!, ideal_eos! and stiffened_eos!.

None of these checks restates the oracle's loops.  Each expectation comes from mathematics that
holds for the reference's discretisation whatever its implementation:

* the divergence identity of Voronoi cells: for a linear field f and a cell with no boundary edge,
  sum_j lrr_ij (f_i - f_j)(m_ij - x_i) = -A_i grad f exactly (lrr (x_i - x_j) = -|e| n and the
  midpoint rule is exact on straight edges);
* the second-moment tensor of a cell, T_i = sum_j |e_ij| d_ij n n^T, taken from scipy's Qhull
  Voronoi diagram (T = 2 A I on a regular hexagon);
* closed forms of a regular hexagonal lattice (all cells congruent, centroid = seed; walls half a
  row spacing beyond the first and last rows, so that every wall edge runs between the vertical
  bisectors of a cell's row neighbours: length a at distance b/2, lrr = 2a/b = 4/sqrt(3));
* polygons from a brute-force clip against every other seed, for the wall edges of jittered cells.

References: src/diffusion.jl:8-78, src/relaxation.jl:10-25, src/pressure.jl:32-70.
"""
import math

import numpy as np
import pytest

from pairs import lattice
from test_oracle_pins import clip_brute

HEX_A = (4.0 / 3.0) ** 0.25   # neighbour distance of the hexagonal lattice, in units of dr
HEX_B = (3.0 / 4.0) ** 0.25   # its row spacing (b = a sqrt(3)/2, cell area a*b = dr^2)


def hex_lattice(na, nb, dr, x0=0.25):
    a, b = HEX_A * dr, HEX_B * dr
    pts = [((i + (j % 2) / 2 + x0) * a, (j + 0.5) * b) for j in range(nb) for i in range(na)]
    x = np.concatenate([[q[0] for q in pts], [q[1] for q in pts]])
    return x, a, b


def interior_mask(p):
    ptr, lab, _ = p.mesh()
    return np.array([np.all(lab[ptr[i]:ptr[i + 1]] >= 0) for i in range(p.n)]), ptr, lab


def ring2_mask(p):
    inner, ptr, lab = interior_mask(p)
    return np.array([inner[i] and all(inner[j] for j in lab[ptr[i]:ptr[i + 1]]) for i in range(p.n)])


def qhull_cells(x, cells):
    """Area A_i, centroid c_i and T_i = sum |e| d n n^T of the listed cells from scipy's Voronoi
    diagram of the free seeds (valid for cells whose Voronoi region the walls do not touch)."""
    from scipy.spatial import Voronoi

    n = x.size // 2
    pts = np.stack([x[:n], x[n:]], axis=1)
    vor = Voronoi(pts)
    T = {i: np.zeros((2, 2)) for i in cells}
    for (a, b), rv in zip(vor.ridge_points, vor.ridge_vertices):
        for i, j in ((a, b), (b, a)):
            if i in T:
                assert -1 not in rv
                e = np.linalg.norm(vor.vertices[rv[0]] - vor.vertices[rv[1]])
                dvec = pts[j] - pts[i]
                d = np.linalg.norm(dvec)
                nv = dvec / d
                T[i] += e * d * np.outer(nv, nv)
    A, C = {}, {}
    for i in cells:
        v = vor.vertices[vor.regions[vor.point_region[i]]]
        ang = np.arctan2(v[:, 1] - pts[i, 1], v[:, 0] - pts[i, 0])
        v = v[np.argsort(ang)]
        w = np.roll(v, -1, axis=0)
        cr = v[:, 0] * w[:, 1] - v[:, 1] * w[:, 0]
        A[i] = 0.5 * cr.sum()
        C[i] = ((v + w) * cr[:, None]).sum(axis=0) / (6.0 * A[i])
    return A, C, T


def set_D(p, D11, D21, D12, D22):
    n = p.n
    D = np.empty(4 * n)
    D[:n], D[n:2 * n], D[2 * n:3 * n], D[3 * n:] = D11, D21, D12, D22
    p.set("D", D)


def D_of(p, i):
    n = p.n
    D = p["D"]
    return np.array([[D[i], D[2 * n + i]], [D[n + i], D[3 * n + i]]])


# ---- find_D! (src/diffusion.jl:8-19) --------------------------------------------------------------

G = np.array([[0.37, -1.21], [0.83, -0.52]])  # a non-symmetric velocity gradient


def test_find_D_linear_velocity_on_hexagons(oracle):
    """v = G x on a regular hexagonal lattice: every cell without a boundary edge has D = sym(G)."""
    dr = 1.0 / 16
    x, a, b = hex_lattice(14, 16, dr)
    g = oracle.Grid((0.0, 0.0), (14.75 * a, 16 * b), dr)
    p = oracle.Particles(g, x)
    p.remesh()
    n = p.n
    X = np.stack([x[:n], x[n:]])
    v = G @ X
    p.set("v", np.concatenate([v[0], v[1]]))
    p.find_D()
    # cells with no boundary edge and the regular hexagon's area a*b (the last seed of a row has no
    # boundary edge either when its two diagonal neighbours close its cell, but that cell is a rhombus)
    inner, _, _ = interior_mask(p)
    inner &= np.abs(p.area() - a * b) <= 1e-12 * a * b
    assert inner.sum() > 0.6 * n
    S = 0.5 * (G + G.T)
    for i in np.nonzero(inner)[0]:
        assert np.allclose(D_of(p, i), S, rtol=0, atol=1e-12), i


def test_find_D_linear_velocity_on_jittered_cells(lv, oracle):
    """v = G x on a jittered lattice: sum lrr (v_i - v_j) (m - y)^T = -A G + G T, so
    D = sym(G T / A - G) with T and A from Qhull."""
    x, dr = lattice(lv, "walls", 18, 0x5EED0700)
    g = oracle.Grid((0.0, 0.0), (1.0, 1.0), dr)
    p = oracle.Particles(g, x)
    p.remesh()
    n = p.n
    X = np.stack([x[:n], x[n:]])
    v = G @ X
    p.set("v", np.concatenate([v[0], v[1]]))
    p.find_D()
    inner, _, _ = interior_mask(p)
    cells = list(np.nonzero(inner)[0])
    A, C, T = qhull_cells(x, cells)
    area = p.area()
    worst = 0.0
    for i in cells:
        assert abs(A[i] - area[i]) <= 1e-12 * area[i]
        M = G @ T[i] / A[i] - G
        expect = 0.5 * (M + M.T)
        worst = max(worst, np.max(np.abs(D_of(p, i) - expect)))
        # the jitter makes T differ from 2 A I, so sym(G) alone would not match
    assert worst <= 1e-11
    off = max(np.max(np.abs(D_of(p, i) - 0.5 * (G + G.T))) for i in cells)
    assert off > 1e-3


# ---- viscous_step! and getS (src/diffusion.jl:22-53) -------------------------------------------------

D0_COMPRESSIVE = np.array([[-0.5, 0.2], [0.1, -0.3]])   # tr = -0.8
D0_EXPANSIVE = np.array([[0.5, 0.2], [0.1, 0.3]])       # tr = +0.8
D1 = np.array([[0.6, -0.4], [0.9, -0.6]])               # traceless: tr D stays tr D0 everywhere
D2 = np.array([[-0.25, 0.7], [-0.35, 0.25]])


def _linear_D_state(p, x, D0, mu, rho):
    n = p.n
    X1, X2 = x[:n], x[n:]
    comp = [(0, 0), (1, 0), (0, 1), (1, 1)]  # D11, D21, D12, D22 (column-major planes)
    set_D(p, *[D0[r, c] + X1 * D1[r, c] + X2 * D2[r, c] for r, c in comp])
    p.set("mu", mu)
    p.set("rho", rho)


@pytest.mark.parametrize("D0,av,stone_norman", [(D0_COMPRESSIVE, True, True), (D0_EXPANSIVE, True, False),
                                                (D0_COMPRESSIVE, False, False)])
def test_viscous_pass1_linear_stress(lv, oracle, D0, av, stone_norman):
    """D linear in x with constant trace: S = 2 mu_eff (D - tr D I/3) is linear, and pass 1 gives
    v += (dt A / m) div S on every cell without a boundary edge.  mu_eff = mu - tr(D) rho dr^2 only
    for a compressive tr D < 0 with artificial viscosity (Stone-Norman), else mu."""
    x, dr = lattice(lv, "walls", 20, 0x5EED0701)
    g = oracle.Grid((0.0, 0.0), (1.0, 1.0), dr)
    p = oracle.Particles(g, x)
    p.remesh()
    n = p.n
    mu, rho = 1e-3, 2.0
    _linear_D_state(p, x, D0, mu, rho)
    area = p.area()
    p.set("mass", 1.7 * area)
    u = np.array([0.3, -0.7])
    p.set("v", np.concatenate([np.full(n, u[0]), np.full(n, u[1])]))
    dt = 1e-2
    e0 = p["e"].copy()
    p.viscous_step(dt, av)
    tr = np.trace(D0)
    mu_eff = mu - tr * rho * dr * dr if stone_norman else mu
    assert (mu_eff > 2 * mu) == stone_norman
    divS = 2.0 * mu_eff * np.array([D1[0, 0] + D2[0, 1], D1[1, 0] + D2[1, 1]])
    inner, _, _ = interior_mask(p)
    assert inner.sum() > 0.7 * n
    dv = (dt / 1.7) * divS  # dt A / m with m = 1.7 A
    vn = p["v"]
    assert np.allclose(vn[:n][inner], u[0] + dv[0], rtol=0, atol=1e-13)
    assert np.allclose(vn[n:][inner], u[1] + dv[1], rtol=0, atol=1e-13)
    assert np.max(np.abs(dv)) > 1e-5
    # pass 2 changed e somewhere (its closed form is pinned below)
    assert np.max(np.abs(p["e"] - e0)) > 0


def test_viscous_pass1_deviatoric_part(lv, oracle):
    """Without artificial viscosity mu_eff = mu, so a D whose trace varies linearly in x still gives a
    linear S = 2 mu (D - tr D I/3), and pass 1 gives v += (dt A/m) div S with the trace's
    gradient entering through the I/3 term."""
    x, dr = lattice(lv, "walls", 20, 0x5EED0705)
    g = oracle.Grid((0.0, 0.0), (1.0, 1.0), dr)
    p = oracle.Particles(g, x)
    p.remesh()
    n = p.n
    E1 = np.array([[0.9, -0.4], [0.2, 0.5]])   # tr 1.4
    E2 = np.array([[-0.3, 0.6], [-0.8, 1.1]])  # tr 0.8
    X1, X2 = x[:n], x[n:]
    comp = [(0, 0), (1, 0), (0, 1), (1, 1)]
    set_D(p, *[D0_EXPANSIVE[r, c] + X1 * E1[r, c] + X2 * E2[r, c] for r, c in comp])
    mu = 3e-3
    p.set("mu", mu)
    p.set("rho", 1.0)
    area = p.area()
    p.set("mass", area)
    p.set("v", 0.0)
    dt = 1e-2
    p.viscous_step(dt, False)
    dS1 = 2.0 * mu * (E1 - np.trace(E1) * np.eye(2) / 3.0)  # dS/dx1
    dS2 = 2.0 * mu * (E2 - np.trace(E2) * np.eye(2) / 3.0)  # dS/dx2
    divS = dS1[:, 0] + dS2[:, 1]
    inner, _, _ = interior_mask(p)
    assert inner.sum() > 0.7 * n
    assert np.allclose(p["v"][:n][inner], dt * divS[0], rtol=0, atol=1e-14)
    assert np.allclose(p["v"][n:][inner], dt * divS[1], rtol=0, atol=1e-14)


def test_viscous_pass2_energy_closed_form(lv, oracle):
    """With m = rho A, pass 1 leaves every inner cell at the same velocity u'.  For cells whose
    neighbours are inner too, pass 2 adds (dt/m) (tr(grad g T) - A div g) with g = S u' linear,
    T from Qhull (the identity sum lrr [(m-x_i).g_i - (m-y).g_j] = -A div g + tr(grad g T))."""
    x, dr = lattice(lv, "walls", 18, 0x5EED0702)
    g = oracle.Grid((0.0, 0.0), (1.0, 1.0), dr)
    p = oracle.Particles(g, x)
    p.remesh()
    n = p.n
    mu, rho = 2e-3, 1.5
    _linear_D_state(p, x, D0_COMPRESSIVE, mu, rho)
    area = p.area()
    p.set("mass", rho * area)
    u = np.array([0.4, 0.25])
    p.set("v", np.concatenate([np.full(n, u[0]), np.full(n, u[1])]))
    p.set("e", 3.0)
    dt = 1e-2
    p.viscous_step(dt, True)
    mu_eff = mu - np.trace(D0_COMPRESSIVE) * rho * dr * dr
    divS = 2.0 * mu_eff * np.array([D1[0, 0] + D2[0, 1], D1[1, 0] + D2[1, 1]])
    up = u + (dt / rho) * divS
    # g(x) = S(x) u' with S(x) = 2 mu_eff (D0 - tr D0 I/3 + x1 D1 + x2 D2): grad g = 2 mu_eff [D1 u', D2 u']
    grad_g = 2.0 * mu_eff * np.stack([D1 @ up, D2 @ up], axis=1)
    S0 = 2.0 * mu_eff * (D0_COMPRESSIVE - np.trace(D0_COMPRESSIVE) * np.eye(2) / 3.0)
    cells = list(np.nonzero(ring2_mask(p))[0])
    assert len(cells) > 0.5 * n
    inner, _, _ = interior_mask(p)
    assert np.allclose(p["v"][:n][inner], up[0], rtol=0, atol=1e-13)
    A, C, T = qhull_cells(x, cells)
    worst, scale = 0.0, 0.0
    for i in cells:
        de = (dt / (rho * area[i])) * (np.trace(grad_g @ T[i]) - A[i] * np.trace(grad_g))
        worst = max(worst, abs((p["e"][i] - 3.0) - de))
        scale = max(scale, abs(de))
    assert scale > 1e-6 and np.linalg.norm(S0) > 0
    assert worst <= 1e-10 * scale + 1e-15
    # the pair fluxes are antisymmetric, so the walled box keeps its total energy even though
    # g = S u' has a normal component on the walls (a flux written with m - y in place of m - x_i in
    # the first dot product would agree on every closed cell but add 2 dt (integral of div g) here)
    dE = p["mass"] * (p["e"] - 3.0)
    assert abs(dE.sum()) <= 1e-12 * np.abs(dE).sum()
    assert abs(2 * dt * np.trace(grad_g)) > 1e-3 * np.abs(dE).sum()


# ---- bdary_friction! (src/diffusion.jl:64-78) ---------------------------------------------------------

def test_friction_wall_rows_of_a_hexagonal_lattice(oracle):
    """x-periodic hexagonal lattice with walls half a row spacing beyond the first and last rows: a
    wall cell has no neighbour below (above), so its wall edge spans the two vertical bisectors of its
    row neighbours, length a at distance b/2: lrr = 2a/b, and
    v' = (v + dt mu lrr vD/m) / (1 + dt mu lrr/m), e' = e + dt (mu lrr vD/m).v; other cells keep v."""
    dr = 1.0 / 16
    na, nb = 16, 12
    x, a, b = hex_lattice(na, nb, dr)
    g = oracle.Grid((0.0, 0.0), (na * a, nb * b), dr, xperiodic=True)
    p = oracle.Particles(g, x)
    p.remesh()
    n = p.n
    rng = np.random.default_rng(5)
    v = rng.standard_normal(2 * n)
    p.set("v", v)
    p.set("e", 2.0 + rng.random(n))
    e0 = p["e"].copy()
    mu = 0.05 + 0.01 * rng.random(n)
    p.set("mu", mu)
    m = a * b * (1.0 + 0.5 * rng.random(n))
    p.set("mass", m)
    vbc = np.zeros(8)
    vbc[0:2] = (1.5, -0.25)   # BDARY_UP = -1
    vbc[4:6] = (-0.75, 0.5)   # BDARY_DOWN = -3
    dt = 1e-3
    p.bdary_friction(vbc, dt)
    lrr = 2.0 * a / b
    assert abs(lrr - 4.0 / math.sqrt(3.0)) <= 1e-14
    row = np.repeat(np.arange(nb), na)
    vn, en = p["v"], p["e"]
    for i in range(n):
        vi = np.array([v[i], v[n + i]])
        if row[i] in (0, nb - 1):
            vd = vbc[4:6] if row[i] == 0 else vbc[0:2]
            f = mu[i] * lrr * vd / m[i]
            expect_v = (vi + dt * f) / (1.0 + dt * mu[i] * lrr / m[i])
            expect_e = e0[i] + dt * np.dot(f, vi)
        else:
            expect_v, expect_e = vi, e0[i]
        assert np.allclose([vn[i], vn[n + i]], expect_v, rtol=1e-13, atol=1e-15), i
        assert abs(en[i] - expect_e) <= 1e-13 * abs(expect_e), i


@pytest.mark.parametrize("skip", [(), (-2, -4)])
def test_friction_on_jittered_walls_brute_force(lv, oracle, skip):
    """Every wall edge of a jittered walled lattice: lrr = (length of the polygon's side on that
    wall) / (distance of the seed to the wall), from a brute-force clip against every seed.  With
    `skip`, those walls' table entries are NaN: charfun is false there and their edges add nothing
    (the bubble example's top_and_bottom, examples/bubble.jl:81-93)."""
    x, dr = lattice(lv, "walls", 12, 0x5EED0703)
    g = oracle.Grid((0.0, 0.0), (1.0, 1.0), dr)
    p = oracle.Particles(g, x)
    p.remesh()
    n = p.n
    rng = np.random.default_rng(6)
    v = rng.standard_normal(2 * n)
    p.set("v", v)
    p.set("e", 1.0 + rng.random(n))
    e0 = p["e"].copy()
    mu = 0.02 + 0.01 * rng.random(n)
    p.set("mu", mu)
    area = p.area()
    p.set("mass", area)
    vbc = np.array([1.0, 0.0, 0.2, -0.3, -0.4, 0.1, 0.0, 0.6])  # per label -1..-4 (up, right, down, left)
    for lab in skip:
        vbc[2 * (-lab - 1):2 * (-lab - 1) + 2] = np.nan
    dt = 2e-3
    v_before = p["v"].copy()
    p.bdary_friction(vbc, dt)
    pts = list(zip(x[:n], x[n:]))
    walls = [(lab, ax, pos) for lab, ax, pos in [(-1, 1, 1.0), (-2, 0, 1.0), (-3, 1, 0.0), (-4, 0, 0.0)]
             if lab not in skip]  # (label, axis, position)
    corner_cells = 0
    for i in range(n):
        _, poly, _ = clip_brute(x[i], x[n + i], pts, i, (0.0, 0.0, 1.0, 1.0))
        poly = np.array(poly)
        xi = np.array([x[i], x[n + i]])
        edges = []
        for lab, ax, pos in walls:
            on = poly[np.abs(poly[:, ax] - pos) <= 1e-13]
            if len(on) >= 2:
                length = on[:, 1 - ax].max() - on[:, 1 - ax].min()
                if length > 0:
                    edges.append((lab, length / abs(pos - xi[ax])))
        vi = np.array([v[i], v[n + i]])
        vd = [vbc[2 * (-lab - 1):2 * (-lab - 1) + 2] for lab, _ in edges]
        f = [mu[i] * lrr * d / area[i] for (_, lrr), d in zip(edges, vd)]
        tmp = 1.0 + sum(dt * mu[i] * lrr / area[i] for _, lrr in edges)
        expect_v = (vi + dt * sum(f, np.zeros(2))) / tmp
        assert np.allclose([p["v"][i], p["v"][n + i]], expect_v, rtol=1e-12, atol=1e-14), i
        if len(edges) <= 1:  # with two wall edges e depends on their order by dt^2 f_a.f_b
            expect_e = e0[i] + sum(dt * np.dot(fk, vi) for fk in f)
            assert abs(p["e"][i] - expect_e) <= 1e-13 * abs(expect_e), i
        else:
            corner_cells += 1
        if not edges:  # no wall with friction: v and e exactly as they were
            assert p["v"][i] == v_before[i] and p["v"][n + i] == v_before[n + i] and p["e"][i] == e0[i], i
    assert corner_cells == (4 if not skip else 0)


# ---- find_dv! (src/relaxation.jl:10-25) -----------------------------------------------------------

def test_find_dv_zero_on_regular_hexagons(oracle):
    """Periodic regular hexagonal lattice: quality = rmin/rmax = 1 and centroid = seed, so dv = 0
    whatever D is."""
    dr = 1.0 / 16
    na, nb = 16, 18
    x, a, b = hex_lattice(na, nb, dr)
    g = oracle.Grid((0.0, 0.0), (na * a, nb * b), dr, xperiodic=True, yperiodic=True)
    p = oracle.Particles(g, x)
    p.remesh()
    n = p.n
    rng = np.random.default_rng(8)
    set_D(p, *[rng.standard_normal(n) for _ in range(4)])
    p.find_dv(1e-3, 1.0)
    assert np.allclose(p["quality"], 1.0, rtol=0, atol=1e-14)
    assert np.max(np.abs(p["dv"])) <= 1e-11 * dr


def test_find_dv_from_qhull_geometry(lv, oracle):
    """A jittered periodic lattice: quality from the Qhull neighbour distances, the centroid from
    the Qhull polygon, and dv = lambda/(1 + dt lambda) (c - x) with lambda = alpha |D|_F / quality^2.
    Checks the limits too: D = 0 gives dv = 0, and a huge dt gives dv -> (c - x)/dt."""
    from scipy.spatial import Voronoi

    x, dr = lattice(lv, "periodic", 16, 0x5EED0704)
    g = oracle.Grid((0.0, 0.0), (1.0, 1.0), dr, xperiodic=True, yperiodic=True)
    p = oracle.Particles(g, x)
    p.remesh()
    n = p.n
    pts = np.stack([x[:n], x[n:]], axis=1)
    allp = np.concatenate([pts + np.array([sx, sy]) for sx in (-1, 0, 1) for sy in (-1, 0, 1)])
    ctr = 4 * n
    vor = Voronoi(allp)
    rmin, rmax = np.full(n, np.inf), np.zeros(n)
    for a_, b_ in vor.ridge_points:
        for i, j in ((a_, b_), (b_, a_)):
            if ctr <= i < ctr + n:
                r = np.linalg.norm(allp[i] - allp[j])
                rmin[i - ctr] = min(rmin[i - ctr], r)
                rmax[i - ctr] = max(rmax[i - ctr], r)
    cen = np.empty((n, 2))
    for i in range(n):
        v = vor.vertices[vor.regions[vor.point_region[ctr + i]]]
        ang = np.arctan2(v[:, 1] - pts[i, 1], v[:, 0] - pts[i, 0])
        v = v[np.argsort(ang)]
        w = np.roll(v, -1, axis=0)
        cr = v[:, 0] * w[:, 1] - v[:, 1] * w[:, 0]
        cen[i] = ((v + w) * cr[:, None]).sum(axis=0) / (3.0 * cr.sum())
    rng = np.random.default_rng(9)
    Dp = [rng.standard_normal(n) for _ in range(4)]
    set_D(p, *Dp)
    alpha, dt = 0.7, 3e-3
    p.find_dv(dt, alpha)
    q = rmin / rmax
    assert np.allclose(p["quality"], q, rtol=1e-13, atol=0)
    nD = np.sqrt(sum(d * d for d in Dp))
    lam = alpha * nD / q ** 2
    f = lam / (1.0 + dt * lam)
    assert np.allclose(p["dv"][:n], f * (cen[:, 0] - pts[:, 0]), rtol=1e-9, atol=1e-13 * dr)
    assert np.allclose(p["dv"][n:], f * (cen[:, 1] - pts[:, 1]), rtol=1e-9, atol=1e-13 * dr)
    set_D(p, *[np.zeros(n)] * 4)
    p.find_dv(dt, alpha)
    assert np.all(p["dv"] == 0.0)
    set_D(p, *Dp)
    big = 1e9
    p.find_dv(big, alpha)
    assert np.allclose(p["dv"][:n] * big, cen[:, 0] - pts[:, 0], rtol=1e-6, atol=1e-12 * dr)


# ---- ideal_eos! and stiffened_eos! (src/pressure.jl:32-70) ----------------------------------------

def test_eos_closed_forms_with_pmin(oracle):
    """On a regular hexagonal lattice every cell has area a*b = dr^2, so rho = m/dr^2 and
    P = (gamma - 1) rho (e - |v|^2/2) are known numbers.  ideal_eos! clamps only the sound speed:
    c2 = gamma max(P, Pmin)/rho, and P itself may stay below Pmin (or be negative)."""
    dr = 1.0 / 16
    na, nb = 16, 18
    x, a, b = hex_lattice(na, nb, dr)
    g = oracle.Grid((0.0, 0.0), (na * a, nb * b), dr, xperiodic=True, yperiodic=True)
    p = oracle.Particles(g, x)
    p.remesh()
    n = p.n
    rng = np.random.default_rng(10)
    m = (0.5 + rng.random(n)) * dr * dr
    v = rng.standard_normal(2 * n)
    ekin = 0.5 * (v[:n] ** 2 + v[n:] ** 2)
    e = ekin + rng.uniform(-0.5, 1.5, n)  # some cells have e below their kinetic energy: P < 0
    gamma, Pmin = 1.4, 0.3
    for k, val in (("mass", m), ("v", v), ("e", e)):
        p.set(k, val)
    p.eos_ideal(gamma, Pmin)
    rho = m / (dr * dr)
    P = 0.4 * rho * (e - ekin)
    assert (P < 0).any() and ((P > 0) & (P < Pmin)).any() and (P > Pmin).any()
    assert np.allclose(p["rho"], rho, rtol=1e-13)
    assert np.allclose(p["P"], P, rtol=1e-12, atol=1e-14)
    assert np.allclose(p["c2"], gamma * np.maximum(P, Pmin) / rho, rtol=1e-12, atol=1e-14)
    P0 = 2.5
    p.eos_stiffened(gamma, P0)
    assert np.allclose(p["P"], P, rtol=1e-12, atol=1e-14)
    assert np.allclose(p["c2"], gamma * (P + P0) / rho, rtol=1e-12)
    # Pmin = 0 (the reference default) with a negative pressure: c2 = 0, not negative
    p.eos_ideal(gamma, 0.0)
    assert np.all(p["c2"][P < 0] == 0.0)
