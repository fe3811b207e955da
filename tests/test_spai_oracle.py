"""CPU pins of the learned-preconditioner oracle (oracle/spai_ref.py).

The oracle's gradients (autograd in float64) are checked against central finite differences; the
untrained network (zero decoders) must emit exactly the Neumann-type inverse w = a; and the
preconditioner must be symmetric positive definite for arbitrary parameters (diagonal dominance).
"""
import numpy as np
import pytest

spai_ref = pytest.importorskip("spai_ref")


def small_system(lv, oracle, K=6, singular=False):
    dr = 1.0 / K
    x = lv.inputs.jittered_lattice(K, K, dr, seed=0x5EED0200)
    g = oracle.Grid((0.0, 0.0), (1.0, 1.0), dr, xperiodic=True, yperiodic=True)
    p = oracle.Particles(g, x)
    p.remesh()
    ic = lv.inputs.taylor_green(x, p.area())
    for k in ("v", "rho", "mass", "P", "e", "mu"):
        p.set(k, ic[k])
    p.eos_stiffened(ic["gamma"], ic["P0"])
    if singular:
        p.set("c2", np.inf)
    diag, coef = p.pressure_operator(2e-2)
    d, rp, cols, vals = p.csr(diag, coef)
    return spai_ref.Graph(d, rp, cols, vals), p


def random_theta(seed, scale_dec=0.3):
    rng = np.random.default_rng(seed)
    th = np.zeros(spai_ref.NPARAM)
    th[: spai_ref.OFF_VE] = rng.standard_normal(spai_ref.OFF_VE) * 0.2
    th[spai_ref.OFF_VE:] = rng.standard_normal(spai_ref.NPARAM - spai_ref.OFF_VE) * scale_dec
    return th


def test_oracle_gradient_matches_finite_differences(lv, oracle):
    g, p = small_system(lv, oracle)
    th = random_theta(3)
    v = spai_ref.probes(g, 3, seed=11, x=p["P"])
    L, grad = spai_ref.loss_and_grad(th, g, v)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(spai_ref.OFF_VE, 24, replace=False),
                          np.arange(spai_ref.OFF_VE, spai_ref.NPARAM, 9)])
    import torch
    for k in idx:
        h = 1e-6
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        fd = (float(spai_ref.loss(torch.tensor(tp), g, v)) - float(spai_ref.loss(torch.tensor(tm), g, v))) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, grad[k])


def test_zero_decoder_gives_scaled_couplings(lv, oracle):
    import torch
    g, p = small_system(lv, oracle)
    th = random_theta(4, scale_dec=0.0)
    oe, on, w, mdiag, delta = spai_ref.forward(torch.tensor(th), g)
    assert np.allclose(w.numpy(), g.a, rtol=1e-15, atol=0)
    expect = np.bincount(g.src, weights=np.abs(g.a), minlength=g.n) + g.rfl
    assert np.allclose(mdiag.numpy(), expect, rtol=1e-14)


@pytest.mark.parametrize("singular", [False, True])
def test_preconditioner_spd_for_any_parameters(lv, oracle, singular):
    g, p = small_system(lv, oracle, singular=singular)
    for seed in range(3):
        M = spai_ref.dense_preconditioner(random_theta(seed, scale_dec=2.0), g)
        assert np.allclose(M, M.T, rtol=1e-12, atol=1e-14)
        assert np.linalg.eigvalsh(M).min() > 0
