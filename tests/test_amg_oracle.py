"""Pins of the multilevel-preconditioner oracle (oracle/amg_ref.py) on the CPU.

The cycle has no reference counterpart, so the oracle is pinned by what the construction fixes:
the coarse operators are Galerkin products P^T A P of the piecewise-constant aggregation (checked
against an explicit sparse product), the V(1,1) cycle is a symmetric positive-definite operator
(checked densely), and as a CG preconditioner it cuts the iterations of Jacobi-PCG several times
on the benchmark's regime, with a periodic and with a singular walled system.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import amg_ref
from pairs import domain, lattice, initial_fields


def system(lv, oracle, kind, K, ic, dt, seed=0x5EED0400):
    xmin, xmax, xp, yp = domain(kind)
    x, dr = lattice(lv, kind, K, seed)
    g = oracle.Grid(xmin, xmax, dr, xperiodic=xp, yperiodic=yp)
    p = oracle.Particles(g, x)
    p.remesh()
    for k, v in initial_fields(lv, ic, x, p.area()).items():
        p.set(k, v)
    if ic == "tg":
        p.eos_stiffened(1.4, 2500.0 / 1.4)
    else:
        p.set("c2", np.full(x.size // 2, np.inf))
    diag, coef = p.pressure_operator(dt)
    d, rp, cols, c = p.csr(diag, coef)
    singular = not np.any(d > 0)
    H = amg_ref.Hierarchy(x, d, rp, cols, c, xmin, xmax, dr, xp, yp, singular)
    return H, p, d, singular


def test_coarse_operators_are_galerkin_products(lv, oracle):
    H, p, d, singular = system(lv, oracle, "periodic", 128, "tg", 8e-3)
    n = H.n
    P1 = sp.csr_matrix((np.ones(n), (np.arange(n), H.agg)), shape=(n, H.levels[0].n))
    G1 = (P1.T @ H.A @ P1).toarray()
    assert np.abs(G1 - H.levels[0].A.toarray()).max() <= 1e-12 * np.abs(G1).max()
    L, C = H.levels[0], H.levels[1]
    f = np.arange(L.n)
    par = (f // L.Ky // 2) * C.Ky + (f % L.Ky) // 2  # 2 x 2 blocks
    P2 = sp.csr_matrix((np.ones(L.n), (f, par)), shape=(L.n, C.n))
    G2 = (P2.T @ L.A @ P2).toarray()
    assert np.abs(G2 - C.A.toarray()).max() <= 1e-12 * np.abs(G2).max()


@pytest.mark.parametrize("kind,ic", [("periodic", "tg"), ("walls", "cavity")])
def test_cycle_is_symmetric_positive_definite(lv, oracle, kind, ic):
    H, p, d, singular = system(lv, oracle, kind, 24, ic, 8e-2)
    n = H.n
    Bm = np.stack([H.apply(e) for e in np.eye(n)], axis=1)
    assert np.abs(Bm - Bm.T).max() <= 1e-10 * np.abs(Bm).max()
    ev = np.linalg.eigvalsh(0.5 * (Bm + Bm.T))
    assert ev.min() > 0.0


@pytest.mark.parametrize("kind,ic", [("periodic", "tg"), ("walls", "cavity")])
def test_deep_cycle_with_overcorrection_is_symmetric_positive_definite(lv, oracle, kind, ic, monkeypatch):
    """Four levels below the particles (144, 36, 9 and a dense 4 cells, with the dense and
    small-level thresholds lowered to 4 and 20 cells so that a 24^2 system has them): the fine and
    first coarse corrections are scaled by 1.6 and 1.4, and for the singular walled system the
    9-cell level is W-visited.  The cycle stays symmetric and positive definite (dense check)."""
    monkeypatch.setattr(amg_ref, "DENSE_MAX", 4)
    monkeypatch.setattr(amg_ref, "TAIL_MAX", 20)
    H, p, d, singular = system(lv, oracle, kind, 24, ic, 2e-2)
    assert [L.n for L in H.levels] == [144, 36, 9, 4]
    assert H.alpha_fine == amg_ref.ALPHA_FINE and H.alpha[:3] == [amg_ref.ALPHA_COARSE, 1.0, 1.0]
    assert H.W == singular
    n = H.n
    Bm = np.stack([H.apply(e) for e in np.eye(n)], axis=1)
    assert np.abs(Bm - Bm.T).max() <= 1e-10 * np.abs(Bm).max()
    ev = np.linalg.eigvalsh(0.5 * (Bm + Bm.T))
    assert ev.min() > 1e-8 * ev.max()


@pytest.mark.parametrize("kind,ic,K,dt", [("periodic", "tg", 128, 4e-3), ("walls", "cavity", 96, 1e-3)])
def test_cycle_cuts_cg_iterations(lv, oracle, kind, ic, K, dt):
    H, p, d, singular = system(lv, oracle, kind, K, ic, dt)
    rng = np.random.default_rng(7)
    b = rng.standard_normal(H.n)
    if singular:
        b -= b.mean()
    _, it_j = amg_ref.pcg(H.A, b, lambda r: H.Dinv * r, 1e-8)
    x, it_m = amg_ref.pcg(H.A, b, H.apply, 1e-8)
    assert it_j > 0 and it_m > 0
    assert it_m <= 0.2 * it_j, (it_m, it_j)
    assert np.linalg.norm(b - H.A @ x) <= 1e-8 * np.linalg.norm(b)


def test_bf16_rounding_matches_torch():
    """amg_ref.bf16 (the device's FP32 -> BF16 round to nearest even) against torch's conversion,
    on random magnitudes and on exact ties."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-6, 6, 4000),
                        1.0 + np.array([2.0 ** -8, 3 * 2.0 ** -8, 2.0 ** -9, 3 * 2.0 ** -9])])
    ref = torch.tensor(v.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(amg_ref.bf16(v), ref)
    assert amg_ref.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0  # a tie goes to the even neighbour
