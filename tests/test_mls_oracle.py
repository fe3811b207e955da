"""Pins of the interpolation oracle (oracle/mls_ref.py) on the CPU.

  * nearest_polygon equals the brute-force nearest seed (raw distance, as the reference measures);
  * moving least squares reproduces a polynomial of its own order exactly: the linear expansion the
    gradient of a linear field, the quadratic one all five coefficients of a quadratic, the cubic
    one all nine of a cubic (up to the conditioning of the normal equations);
  * point_value returns f(seed) at a seed and the exact value of a linear field anywhere;
  * populate_lloyd!'s iterations never increase the centroidal Voronoi energy
    sum_i integral_{V_i} |y - x_i|^2 dy (Lloyd's algorithm is a descent method for it).
"""
import numpy as np
import pytest

import mls_ref
from pairs import domain, lattice


def _walk(lv, oracle, kind, K, seed=0x5EED0900):
    xmin, xmax, xp, yp = domain(kind)
    x, dr = lattice(lv, kind, K, seed)
    g = oracle.Grid(xmin, xmax, dr, xperiodic=xp, yperiodic=yp)
    return mls_ref.Walk(g, x), g, x, dr


@pytest.mark.parametrize("kind", ["periodic", "walls"])
def test_nearest_polygon_is_brute_force(lv, oracle, kind):
    w, g, x, dr = _walk(lv, oracle, kind, 24)
    rng = np.random.default_rng(1)
    n = w.n
    for q1, q2 in rng.random((200, 2)):
        d = (x[:n] - q1) ** 2 + (x[n:] - q2) ** 2
        assert mls_ref.nearest_polygon(w, q1, q2) == int(np.argmin(d))


def _poly(order, X, Y):
    """a polynomial of the given order and its Taylor coefficients at (X, Y) in power_vector order"""
    if order == 1:
        return 2.0 + 3.0 * X - 5.0 * Y, np.stack([3.0 + 0 * X, -5.0 + 0 * X], 1)
    if order == 2:
        f = 1.0 + 2.0 * X + 3.0 * Y + 4.0 * X * X - X * Y + 0.5 * Y * Y
        g = np.stack([2.0 + 8.0 * X - Y, 3.0 - X + Y, 4.0 + 0 * X, -1.0 + 0 * X, 0.5 + 0 * X], 1)
        return f, g
    f = 1.0 + X - Y + 2.0 * X * X + X * Y - Y * Y + 0.5 * X ** 3 - X * X * Y + 2.0 * X * Y * Y - 0.3 * Y ** 3
    fx = 1.0 + 4.0 * X + Y + 1.5 * X * X - 2.0 * X * Y + 2.0 * Y * Y
    fy = -1.0 + X - 2.0 * Y - X * X + 4.0 * X * Y - 0.9 * Y * Y
    fxx = 2.0 + 1.5 * X - Y          # f_xx / 2
    fxy = 1.0 - 2.0 * X + 4.0 * Y    # f_xy
    fyy = -1.0 + 2.0 * X - 0.9 * Y   # f_yy / 2
    g = np.stack([fx, fy, fxx, fxy, fyy, 0.5 + 0 * X, -1.0 + 0 * X, 2.0 + 0 * X, -0.3 + 0 * X], 1)
    return f, g


@pytest.mark.parametrize("order,hf,tol", [(1, 2.0, 1e-12), (2, 2.0, 1e-9), (3, 3.0, 1e-6)])
def test_movingls_reproduces_polynomials(lv, oracle, order, hf, tol):
    w, g, x, dr = _walk(lv, oracle, "walls", 24)
    n = w.n
    f, ref = _poly(order, x[:n], x[n:])
    h = hf * dr
    # interior seeds: near a wall fewer than the expansion's unknowns lie within h
    inner = np.nonzero((np.minimum(x[:n], 1 - x[:n]) > h + dr) & (np.minimum(x[n:], 1 - x[n:]) > h + dr))[0]
    for i in inner[::5]:
        a = mls_ref.movingls(w, order, f, i, h)
        assert np.allclose(a, ref[i], rtol=tol, atol=tol * np.abs(ref[i]).max()), (i, a, ref[i])


def test_point_value_at_seeds_and_of_linear_fields(lv, oracle):
    w, g, x, dr = _walk(lv, oracle, "walls", 24)
    n = w.n
    f, _ = _poly(1, x[:n], x[n:])
    for i in range(0, n, 11):
        assert mls_ref.point_value(w, f, x[i], x[n + i]) == pytest.approx(f[i], rel=1e-14)
    rng = np.random.default_rng(2)
    for q1, q2 in 0.05 + 0.9 * rng.random((50, 2)):
        assert mls_ref.point_value(w, f, q1, q2) == pytest.approx(2.0 + 3.0 * q1 - 5.0 * q2, rel=1e-12)


def _cvt_energy(p):
    ptr, lab, ev = p.mesh()
    x = p["x"]
    n = p.n
    E = 0.0
    for i in range(n):
        for e in ev[ptr[i]:ptr[i + 1]]:
            a = e[:2] - (x[i], x[n + i])
            b = e[2:] - (x[i], x[n + i])
            A = 0.5 * abs(a[0] * b[1] - a[1] * b[0])
            E += A / 6.0 * (a @ a + a @ b + b @ b)
    return E


def test_lloyd_iterations_decrease_the_centroidal_energy(lv, oracle):
    xmin, xmax, xp, yp = domain("walls")
    dr = 1.0 / 20
    rng = np.random.default_rng(3)
    s = rng.random((400, 2))
    x = np.concatenate([s[:, 0], s[:, 1]])
    g = oracle.Grid(xmin, xmax, dr)
    energies = [_cvt_energy(mls_ref.lloyd(oracle, g, x, k)) for k in (0, 1, 2, 5, 10)]
    assert all(b <= a * (1 + 1e-12) for a, b in zip(energies, energies[1:])), energies
    assert energies[-1] < 0.8 * energies[0]
