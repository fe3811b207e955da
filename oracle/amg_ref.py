"""FP64 CPU oracle of the multilevel preconditioner (hierarchy and V(1,1) cycle).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the HIP cycle of
lagrangianvoronoi.jl_amd/csrc/lv_amg.hip vector for vector.  It restates the same construction in
numpy/scipy FP64:

  * level 1: cells of a regular grid of width >= CELL_DR dr (2 dr) over the boundary rectangle, each particle in
    the cell holding it; Galerkin sums of d and of the couplings to the 8 neighbouring cells
    (couplings reaching further are dropped from both sides);
  * level l+1: BLOCK x BLOCK (2 x 2) blocks of level-l cells; the coarsest level (<= 64 cells) inverted densely,
    with sigma/nc added to every entry of a singular system (sigma = the sum of its diagonal / nc);
  * symmetric V(1,1): fine smoother s*M (the learned M, s from 10 power iterations towards
    lambda_max(s M A) = 1.2) or (2/3) D^-1; coarse smoothers (2/3) D^-1;
  * the prolongated coarse correction scaled by 1.6 into the fine level and by 1.4 into a coarse
    level, except for corrections from levels of at most TAIL_MAX cells (scale 1);
  * for a singular system, the levels of at most TAIL_MAX cells (other than the dense one) are
    visited twice (W-cycle): the second visit solves for the residual the first correction leaves;
  * fine_bf16 (the device's aggregate-order layout): the fine residuals of the cycle use the
    couplings rounded to FP32 and then to BF16 (nearest even); D^-1 and the power iteration keep
    the operator's own.

No counterpart exists in the reference (src/pressure.jl:219 runs MINRES unpreconditioned).  The
oracle is itself pinned by tests/test_amg_oracle.py (Galerkin identity, symmetry and definiteness
of the cycle, iteration counts against Jacobi).
"""
import numpy as np
import scipy.sparse as sp

DENSE_MAX = 64
OMEGA = 2.0 / 3.0
CELL_DR = 2.0  # level-1 cell width in units of dr
BLOCK = 2  # coarse levels: BLOCK x BLOCK blocks of the finer level's cells
TAIL_MAX = 1024
ALPHA_FINE = 1.6
ALPHA_COARSE = 1.4
FINE_TARGET = 1.2
OFFS = [(dx, dy) for dx in (-1, 0, 1) for dy in (-1, 0, 1) if (dx, dy) != (0, 0)]


def _fold(d, K, per):
    if not per:
        return d
    d = np.fmod(d, K)
    d = np.where(d > K // 2, d - K, d)
    d = np.where(d < -((K - 1) // 2), d + K, d)
    return d


def _wrap(i, K, per):
    if per:
        return np.mod(i, K), np.ones_like(i, dtype=bool)
    return i, (i >= 0) & (i < K)


class Level:
    """Kx x Ky cells; difference-form stencil d[n], c[8, n]"""

    def __init__(self, Kx, Ky, d, c, xper, yper):
        self.Kx, self.Ky, self.n = Kx, Ky, Kx * Ky
        self.d, self.c = d, c
        self.xper, self.yper = xper, yper
        self.dinv = np.where(d + c.sum(0) > 0, 1.0 / np.maximum(d + c.sum(0), 1e-300), 1.0)
        I = np.arange(self.n)
        Ix, Iy = I // Ky, I % Ky
        rows, cols, vals = [], [], []
        diag = d.copy()
        for k, (ox, oy) in enumerate(OFFS):
            jx, okx = _wrap(Ix + ox, Kx, xper)
            jy, oky = _wrap(Iy + oy, Ky, yper)
            ok = okx & oky & (c[k] != 0)
            rows.append(I[ok])
            cols.append((jx * Ky + jy)[ok])
            vals.append(-c[k][ok])
            diag = diag + np.where(ok, c[k], 0.0)
        rows.append(I)
        cols.append(I)
        vals.append(diag)
        self.A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                               shape=(self.n, self.n))


class Hierarchy:
    def __init__(self, x, d, rp, cols, coef, bxmin, bxmax, dr, xper, yper, singular, M=None, fine_bf16=True):
        """x: SoA positions (2n); (d, rp, cols, coef): the operator as lv_get_pressure_operator
        returns it; M: the learned preconditioner as a sparse matrix, or None for Jacobi."""
        n = d.size
        self.n = n
        src = np.repeat(np.arange(n), np.diff(rp))
        C = sp.csr_matrix((coef, (src, cols)), shape=(n, n))
        self.A = (sp.diags(d + np.asarray(C.sum(axis=1)).ravel()) - C).tocsr()
        if fine_bf16:
            Ct = sp.csr_matrix((bf16(coef), (src, cols)), shape=(n, n))
            self.Af = (sp.diags(d + np.asarray(Ct.sum(axis=1)).ravel()) - Ct).tocsr()
        else:
            self.Af = self.A
        self.Dinv = 1.0 / self.A.diagonal()
        Lx, Ly = bxmax[0] - bxmin[0], bxmax[1] - bxmin[1]
        Kx, Ky = max(1, int(np.floor(Lx / (CELL_DR * dr)))), max(1, int(np.floor(Ly / (CELL_DR * dr))))
        ix = np.clip(np.floor((x[:n] - bxmin[0]) / (Lx / Kx)).astype(int), 0, Kx - 1)
        iy = np.clip(np.floor((x[n:] - bxmin[1]) / (Ly / Ky)).astype(int), 0, Ky - 1)
        agg = ix * Ky + iy
        self.agg = agg
        na = Kx * Ky
        cnt = np.bincount(agg, minlength=na)
        d1 = np.bincount(agg, weights=d, minlength=na)
        d1[cnt == 0] = 1.0
        c1 = np.zeros((8, na))
        I, J = agg[src], agg[cols]
        dx = _fold(J // Ky - I // Ky, Kx, xper)
        dy = _fold(J % Ky - I % Ky, Ky, yper)
        keep = (I != J) & (np.abs(dx) <= 1) & (np.abs(dy) <= 1) & ~((dx == 0) & (dy == 0))
        slot = np.array([OFFS.index((a, b)) if k else -1 for a, b, k in zip(dx, dy, keep)], dtype=int)
        for k in range(8):
            m = slot == k
            c1[k] = np.bincount(I[m], weights=coef[m], minlength=na)
        self.levels = [Level(Kx, Ky, d1, c1, xper, yper)]
        while self.levels[-1].n > DENSE_MAX:
            F = self.levels[-1]
            Kxc, Kyc = (F.Kx + BLOCK - 1) // BLOCK, (F.Ky + BLOCK - 1) // BLOCK
            f = np.arange(F.n)
            fx, fy = f // F.Ky, f % F.Ky
            P = (fx // BLOCK) * Kyc + fy // BLOCK
            dc = np.bincount(P, weights=F.d, minlength=Kxc * Kyc)
            cc = np.zeros((8, Kxc * Kyc))
            for k, (ox, oy) in enumerate(OFFS):
                nx, okx = _wrap(fx + ox, F.Kx, xper)
                ny, oky = _wrap(fy + oy, F.Ky, yper)
                ddx = _fold(nx // BLOCK - fx // BLOCK, Kxc, xper)
                ddy = _fold(ny // BLOCK - fy // BLOCK, Kyc, yper)
                ok = okx & oky & (F.c[k] != 0) & ~((ddx == 0) & (ddy == 0))
                for q in np.where(ok)[0]:
                    cc[OFFS.index((int(ddx[q]), int(ddy[q]))), P[q]] += F.c[k][q]
            self.levels.append(Level(Kxc, Kyc, dc, cc, xper, yper))
        Dn = self.levels[-1].A.toarray()
        if singular:
            Dn = Dn + np.trace(Dn) / Dn.shape[0] / Dn.shape[0]
        self.inv = np.linalg.inv(Dn)
        self.W = bool(singular)
        nl = len(self.levels)
        # alpha[k]: factor on the correction prolongated into levels[k] from levels[k+1]
        self.alpha = [ALPHA_COARSE if (k + 1 < nl and self.levels[k + 1].n > TAIL_MAX) else 1.0 for k in range(nl)]
        self.alpha_fine = ALPHA_FINE if self.levels[0].n > TAIL_MAX else 1.0
        self.M = M
        self.scale = OMEGA
        if M is not None:
            v = _probe(n)
            for _ in range(10):
                u = M @ (self.A @ v)
                lam = np.sqrt((u @ u) / (v @ v))
                v = u / np.linalg.norm(u)
            self.scale = FINE_TARGET / lam

    def _fine(self, r):
        return self.scale * (self.M @ r if self.M is not None else self.Dinv * r)

    def _visit(self, l, r):
        """the correction for levels[l] from its right-hand side: one V-cycle visit, or two (the
        second on the remaining residual) for a W-visited level"""
        e = self._cycle(l, r)
        if self.W and l < len(self.levels) - 1 and self.levels[l].n <= TAIL_MAX:
            e = e + self._cycle(l, r - self.levels[l].A @ e)
        return e

    def _cycle(self, l, r):
        if l == len(self.levels) - 1:
            return self.inv @ r
        L = self.levels[l]
        z = OMEGA * L.dinv * r
        t = r - L.A @ z
        C = self.levels[l + 1]
        f = np.arange(L.n)
        P = (f // L.Ky // BLOCK) * C.Ky + (f % L.Ky) // BLOCK
        ec = self._visit(l + 1, np.bincount(P, weights=t, minlength=C.n))
        z = z + self.alpha[l] * ec[P]
        t = r - L.A @ z
        return z + OMEGA * L.dinv * t

    def apply(self, r):
        """z = B r, the symmetric V(1,1) cycle"""
        z = self._fine(r)
        t = r - self.Af @ z
        ec = self._visit(0, np.bincount(self.agg, weights=t, minlength=self.levels[0].n))
        z = z + self.alpha_fine * ec[self.agg]
        t = r - self.Af @ z
        return z + self._fine(t)


def bf16(v):
    """v rounded to FP32, then to BF16 (round to nearest even), as FP64"""
    u = np.asarray(v, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _probe(n):
    """+-1 from the SplitMix64 finaliser of 0x5EED5EED + (i+1)*golden, as the device's power iteration"""
    with np.errstate(over="ignore"):
        z = np.uint64(0x5EED5EED) + (np.arange(n, dtype=np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return np.where((z >> np.uint64(63)) != 0, 1.0, -1.0)


def pcg(A, b, apply, tol, x0=None, itmax=5000):
    """FP64 PCG iterations to |b - A x| <= tol |b|"""
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - A @ x
    nb = np.linalg.norm(b)
    z = apply(r)
    p = z.copy()
    rz = r @ z
    for it in range(1, itmax + 1):
        q = A @ p
        al = rz / (p @ q)
        x += al * p
        r -= al * q
        if np.linalg.norm(r) <= tol * nb:
            return x, it
        z = apply(r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return x, -1
