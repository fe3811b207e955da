"""nearest_polygon, movingls, point_value and populate_* on the GPU against the oracle
(oracle/mls_ref.py, itself pinned by tests/test_mls_oracle.py).

Tolerances: the nearest polygon is an index (exact); the linear expansion and point_value differ
only by the order of FP64 roundings (1e-12 relative); the quadratic and cubic expansions solve
normal equations whose condition number grows like dr^-2 and dr^-4, so their coefficients agree to
1e-9 and 1e-6 of the largest.  Lloyd relaxation reuses the bit-exact remesh and centroid, so its
positions agree to 1e-12 after 20 iterations."""
import numpy as np
import pytest

import mls_ref
from pairs import make_pair, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["periodic", "walls"])
def test_nearest_polygon_and_point_value_match_oracle(lv, oracle, gpu, kind):
    print("nearest/point_value %s" % kind, flush=True)
    grid, op, x, dr = make_pair(lv, oracle, kind, K=40, seed=0x5EED0910)
    w = mls_ref.Walk(op.grid, x)
    rng = np.random.default_rng(4)
    q = rng.random((300, 2))
    idx = lv.nearest_polygon(grid, q)
    assert list(idx) == [mls_ref.nearest_polygon(w, a, b) for a, b in q]
    P = grid.get("P")
    v = grid.get_soa("v")
    for field, f in (("P", P), ("v[2]", v[grid.n:])):
        got = lv.point_value(grid, q, field)
        ref = np.array([mls_ref.point_value(w, f, a, b) for a, b in q])
        assert rel_err(got, ref) <= 1e-12, field


@pytest.mark.parametrize("order,tol,hf", [(1, 1e-12, None), (2, 1e-9, None), (3, 1e-6, 3.0)])
def test_movingls_matches_oracle(lv, oracle, gpu, order, tol, hf):
    print("movingls order %d" % order, flush=True)
    grid, op, x, dr = make_pair(lv, oracle, "periodic", K=32, seed=0x5EED0911)
    w = mls_ref.Walk(op.grid, x)
    h = grid.h if hf is None else hf * dr
    f = grid.get("P")
    got = lv.movingls(order, grid, "P", h=None if hf is None else h)
    ref = np.array([mls_ref.movingls(w, order, f, i, h) for i in range(grid.n)])
    scale = np.abs(ref).max(axis=1, keepdims=True)
    assert np.max(np.abs(got - ref) / scale) <= tol
    with pytest.raises(lv.LVError):
        lv.movingls(order, grid, "P", h=20 * dr)  # beyond r_max


def test_lloyd_matches_oracle(lv, oracle, gpu):
    """populate_lloyd! with 20 iterations: the device's positions equal the oracle's loop of remesh
    and centroid on the same random seeds"""
    print("lloyd", flush=True)
    dr = 1.0 / 24
    grid = lv.VoronoiGrid(lv.UnitRectangle(), dr)
    lv.populate_rand(grid, seed=77)
    x0 = grid.get_soa("x")
    lv.populate_lloyd(grid, niterations=20, seed=77)
    p = mls_ref.lloyd(oracle, oracle.Grid((0.0, 0.0), (1.0, 1.0), dr), x0, 20)
    assert rel_err(grid.get_soa("x"), p["x"]) <= 1e-12
    assert np.sum(grid.get("area")) == pytest.approx(1.0, rel=1e-12)


def test_populate_initialisers(lv, gpu):
    """seed counts of populate_rect! ((N-1) x (M-1): range(xmin, xmax, N) includes xmax, and the
    last row and column, shifted by dr/2, fall outside the rectangle), populate_hex! (the hexagonal
    lattice of the reference's Taylor-Green test), populate_circ! and populate_vogel! (filtered by
    charfun); each mesh tiles its domain"""
    print("populate", flush=True)
    import flows

    dr = 1.0 / 20
    g = lv.populate_rect(lv.VoronoiGrid(lv.UnitRectangle(), dr))
    assert g.n == 19 * 19
    g = lv.populate_hex(lv.VoronoiGrid(lv.UnitRectangle(), dr, xperiodic=True, yperiodic=True))
    assert g.n == flows.populate_hex((0.0, 0.0), (1.0, 1.0), dr).size // 2
    disk = lambda x: (x[0] - 0.5) ** 2 + (x[1] - 0.5) ** 2 < 0.16
    for pop in (lv.populate_circ, lv.populate_vogel):
        # the empty corners make the rim polygons reach beyond 10 dr: r_max = 20 dr keeps the mesh
        # (with the default the reference throws "The Voronoi Mesh has been destroyed.")
        g = lv.VoronoiGrid(lv.UnitRectangle(), dr, r_max=20 * dr)
        pop(g, charfun=disk, center=(0.5, 0.5))
        xs = g.get("x")
        assert g.n > 0 and np.all((xs[:, 0] - 0.5) ** 2 + (xs[:, 1] - 0.5) ** 2 < 0.16)
        assert np.sum(g.get("area")) == pytest.approx(1.0, rel=1e-12)
    seen = []
    g = lv.populate_rect(lv.VoronoiGrid(lv.UnitRectangle(), dr), ic=lambda grid: seen.append(grid.n))
    assert seen == [361]
