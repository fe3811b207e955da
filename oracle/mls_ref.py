"""FP64 oracle of nearest_polygon (src/voronoigrid.jl:154-178), movingls with the Gaussian kernel
(src/movingls.jl:22-96), point_value (:140-145) and the iterations of populate_lloyd!
(src/populate.jl:132-145).

TEST INFRASTRUCTURE ONLY: tests/ compare lagrangianvoronoi.jl_amd's HIP kernels (csrc/lv_mls.hip)
and lv_lloyd with these loops on the same seeds.  The walks follow the reference over the oracle's
cell list (lvoracle.Grid.cell_list: every in-grid periodic image of a label in its cell, labels
ascending) and magic path:
  * nearest_polygon: |q.x - x|^2 at the seed's own position, stop at the first node with rr >
    the best distance;
  * movingls: y = x + get_arrow(q.x, x), skip y == x, keep |x - y| < h, stop at the first node with
    rr > h; a += w (f(q) - f(p)) dx, R += (w dx) dx^T in visiting order; R \\ a by Cramer's rule
    (2x2, as StaticArrays solves 2x2 systems) or numpy.linalg.solve (LU with partial pivoting).
Pinned in tests/test_mls_oracle.py by brute force (nearest seed) and by exact reproduction of
linear, quadratic and cubic fields.
"""
import math

import numpy as np


class Walk:
    """the cell list and magic path of an lvoracle.Grid for seeds x (SoA planes)"""

    def __init__(self, grid, x):
        self.g = grid
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.n = self.x.size // 2
        self.cs, self.cl = grid.cell_list(self.x)
        self.i1, self.i2, self.rr = grid.magic_path()
        d = grid.desc
        self.h = d.h
        self.origin = (grid.cropping_min[0] - d.h, grid.cropping_min[1] - d.h)
        self.per = (d.xperiodic, d.yperiodic)
        self.L = (d.bxmax[0] - d.bxmin[0], d.bxmax[1] - d.bxmin[1])

    def key(self, q1, q2):
        return int(math.floor((q1 - self.origin[0]) / self.h)) + 1, int(math.floor((q2 - self.origin[1]) / self.h)) + 1

    def cells(self, k1, k2, bound):
        """the labels of the path's cells in order, while node rr <= bound (bound a callable)"""
        for k in range(self.g.npath):
            if self.rr[k] > bound():
                return
            c1, c2 = k1 + self.i1[k], k2 + self.i2[k]
            if c1 < 1 or c1 > self.g.n1 or c2 < 1 or c2 > self.g.n2:
                continue
            c = (c1 - 1) + (c2 - 1) * self.g.n1
            for s in range(self.cs[c], self.cs[c + 1]):
                yield int(self.cl[s])

    def arrow_target(self, q1, q2, x1, x2):
        """x + get_arrow(q, x), src/voronoigrid.jl:116-125"""
        v1, v2 = q1 - x1, q2 - x2
        if self.per[0] and abs(v1) > 0.5 * self.L[0]:
            sp = (1.0 if v1 > 0.0 else -1.0) * self.L[0]
            v1, v2 = v1 - sp * 1.0, v2 - sp * 0.0
        if self.per[1] and abs(v2) > 0.5 * self.L[1]:
            sp = (1.0 if v2 > 0.0 else -1.0) * self.L[1]
            v1, v2 = v1 - sp * 0.0, v2 - sp * 1.0
        return x1 + v1, x2 + v2


def nearest_polygon(w, q1, q2):
    k1, k2 = w.key(q1, q2)
    best = [math.inf]
    ib = -1
    for j in w.cells(k1, k2, lambda: best[0]):
        a, b = w.x[j] - q1, w.x[w.n + j] - q2
        rr = a * a + b * b
        if rr < best[0]:
            best[0], ib = rr, j
    return ib


def power_vector(order, d1, d2):
    p = [d1, d2]
    if order >= 2:
        p += [d1 * d1, d1 * d2, d2 * d2]
    if order >= 3:
        p += [(d1 * d1) * d1, (d1 * d1) * d2, (d1 * d2) * d2, (d2 * d2) * d2]
    return np.array(p)


def gaussian(h, r):
    q = r / h
    return math.exp(-0.5 * (q * q)) / math.sqrt((2.0 * math.pi) * (h * h))


def movingls(w, order, f, i, h):
    x1, x2 = w.x[i], w.x[w.n + i]
    fp = f[i]
    N = {1: 2, 2: 5, 3: 9}[order]
    a = np.zeros(N)
    R = np.zeros((N, N))
    k1, k2 = w.key(x1, x2)
    for j in w.cells(k1, k2, lambda: h):
        y1, y2 = w.arrow_target(w.x[j], w.x[w.n + j], x1, x2)
        if x1 == y1 and x2 == y2:
            continue
        r = math.sqrt((x1 - y1) ** 2 + (x2 - y2) ** 2)
        if not r < h:
            continue
        k = gaussian(h, r)
        dx = power_vector(order, y1 - x1, y2 - x2)
        a += (k * (f[j] - fp)) * dx
        R += np.outer(k * dx, dx)
    if N == 2:
        det = R[0, 0] * R[1, 1] - R[0, 1] * R[1, 0]
        return np.array([(R[1, 1] * a[0] - R[0, 1] * a[1]) / det, (R[0, 0] * a[1] - R[1, 0] * a[0]) / det])
    return np.linalg.solve(R, a)


def point_value(w, f, q1, q2):
    p = nearest_polygon(w, q1, q2)
    L = movingls(w, 1, f, p, w.h)
    return f[p] + (L[0] * (q1 - w.x[p]) + L[1] * (q2 - w.x[w.n + p]))


def lloyd(oracle, grid, x, niterations):
    """populate_lloyd!'s loop on the oracle: niterations times remesh! and x = centroid(p)"""
    p = oracle.Particles(grid, x)
    for _ in range(niterations):
        p.remesh()
        c = p.centroid()
        p.set("x", c)
    p.remesh()
    return p
