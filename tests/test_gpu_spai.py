"""The learned SPAI preconditioner on the GPU: BF16-MFMA forward and hand-written backward against
the FP64 oracle (oracle/spai_ref.py), SPD/fallback behaviour, and what training gains.

Tolerances: the network runs with BF16 operands and FP32 accumulation, so outputs and gradients
are compared in relative L2 norm (3e-2 and 1e-1) and by cosine similarity, not elementwise.
"""
import numpy as np
import pytest

from pairs import make_pair

pytestmark = pytest.mark.gpu

spai_ref = pytest.importorskip("spai_ref")


def _system(lv, oracle, K=24, kind="periodic", ic="tg", dt=2e-2):
    grid, op, x, dr = make_pair(lv, oracle, kind, K=K, seed=0x5EED0300, ic=ic)
    if ic == "tg":
        lv.stiffened_eos(grid, 1.4, 2500.0 / 1.4)
    solver = lv.PressureSolver(grid, method="pcg64", tol=1e-10)
    solver.assemble(dt)
    d, rp, cols, coef = solver.operator()
    g = spai_ref.Graph(d, rp, cols, coef)
    return grid, solver, g


def _theta(seed, dec=0.3):
    rng = np.random.default_rng(seed)
    th = np.zeros(spai_ref.NPARAM, dtype=np.float32)
    th[: spai_ref.OFF_VE] = (rng.standard_normal(spai_ref.OFF_VE) * 0.2).astype(np.float32)
    th[spai_ref.OFF_VE:] = (rng.standard_normal(spai_ref.NPARAM - spai_ref.OFF_VE) * dec).astype(np.float32)
    return th


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("kind,ic,K", [("periodic", "tg", 24), ("walls", "cavity", 24), ("periodic", "tg", 301)])
def test_spai_forward_matches_oracle(lv, oracle, gpu, kind, ic, K):
    """The solver's inference pass (persistent MFMA waves; at K = 301 every wave loops over several
    32-edge tiles and the last tile is partial) against the FP64 oracle."""
    import torch

    grid, solver, g = _system(lv, oracle, K=K, kind=kind, ic=ic)
    assert K < 301 or g.cols.size > 4096 * 4 * 32 and g.cols.size % 32 != 0
    th = _theta(1)
    lv.spai.set_params(grid, th)
    oe, on, w, md = lv.spai.outputs(grid, g.cols.size)
    with torch.no_grad():
        roe, ron, rw, rmd, rdl = spai_ref.forward(torch.tensor(th.astype(np.float64)), g)
    assert _rel(oe, roe.numpy()) <= 3e-2
    assert _rel(on, ron.numpy()) <= 3e-2
    assert _rel(w, rw.numpy()) <= 3e-2
    assert _rel(md, rmd.numpy()) <= 3e-2


def test_spai_zero_decoder_is_exact(lv, oracle, gpu):
    grid, solver, g = _system(lv, oracle)
    th = _theta(2, dec=0.0)
    lv.spai.set_params(grid, th)
    oe, on, w, md = lv.spai.outputs(grid, g.cols.size)
    assert np.all(oe == 0) and np.all(on == 0)
    assert _rel(w, g.a) <= 1e-6
    assert _rel(md, np.bincount(g.src, weights=np.abs(g.a), minlength=g.n) + g.rfl) <= 1e-6


def test_spai_gradient_matches_oracle(lv, oracle, gpu):
    grid, solver, g = _system(lv, oracle)
    th = _theta(3)
    lv.spai.set_params(grid, th)
    P = 4
    L_gpu, G_gpu = lv.spai.loss_grad(grid, nprobe=P, seed=99, use_data=True)
    v = spai_ref.probes(g, P, 99, x=grid.get("P"))
    L_ref, G_ref = spai_ref.loss_and_grad(th.astype(np.float64), g, v)
    assert abs(L_gpu - L_ref) <= 2e-2 * abs(L_ref)
    blocks = [(spai_ref.OFF_WN, spai_ref.OFF_WE), (spai_ref.OFF_WE, spai_ref.OFF_W(0))]
    blocks += [(spai_ref.OFF_W(l), spai_ref.OFF_B(l) + 64) for l in range(3)]
    blocks += [(spai_ref.OFF_VE, spai_ref.OFF_UN), (spai_ref.OFF_UN, spai_ref.NPARAM)]
    for a, b in blocks:
        ga, gr = G_gpu[a:b].astype(np.float64), G_ref[a:b]
        if np.linalg.norm(gr) == 0.0:  # outputs inside the clamp: no gradient on either side
            assert np.linalg.norm(ga) == 0.0
            continue
        cos = float(ga @ gr / (np.linalg.norm(ga) * np.linalg.norm(gr)))
        assert cos >= 0.99 and _rel(ga, gr) <= 0.1, (a, b, cos, _rel(ga, gr))


def _cg_iters(A, M, b, tol, itmax):
    """PCG iterations (FP64, dense, CPU) to |r| <= tol |b| from x = 0"""
    x = np.zeros_like(b)
    r = b.copy()
    z = M @ r
    p = z.copy()
    rz = r @ z
    for it in range(1, itmax + 1):
        q = A @ p
        al = rz / (p @ q)
        x += al * p
        r -= al * q
        if np.linalg.norm(r) <= tol * np.linalg.norm(b):
            return it
        z = M @ r
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return itmax + 1


def test_spai_solver_meets_tolerance_and_falls_back(lv, oracle, gpu):
    """The learned preconditioner inside FP64 refinement meets the FP64 tolerance without a fallback;
    a poor network (edge outputs of both signs, mostly beyond the +-8 clamp, so couplings are scaled by
    up to e^8 at random) stalls within itmax, which the CPU confirms on the same M, and each sweep
    is redone with Jacobi, judged by the same FP64 residual."""
    grid, solver, g = _system(lv, oracle, K=32, dt=2e-3)
    b, _ = solver.rhs(2e-3, False)
    x0 = grid.get("P")
    lv.spai.set_params(grid, _theta(4, dec=0.0))
    x = solver.solve(b, x0, method="spai", tol=1e-9)
    assert max(solver.last_stats["relres"]) <= 1e-9 and solver.last_stats["fallbacks"] == 0
    bad = _theta(4, dec=0.0)
    bad[spai_ref.OFF_VE:spai_ref.OFF_CE] = np.random.default_rng(5).standard_normal(64).astype(np.float32) * 40.0
    import torch

    with torch.no_grad():
        oe0 = spai_ref.forward(torch.tensor(bad.astype(np.float64)), g)[0].numpy()
    bad[spai_ref.OFF_CE] = -np.median(oe0)  # outputs of both signs, most of them beyond the clamp
    # the oracle's M for this network needs more than itmax iterations for the first sweep's 1e-4
    itmax = 25
    d, rp, cols, coef = solver.operator()
    A = np.diag(g.D)
    A[g.src, g.cols] -= g.coef
    Mt = spai_ref.dense_preconditioner(bad.astype(np.float64), g)
    s = 1.0 / np.sqrt(g.D)
    M = s[:, None] * Mt * s[None, :]
    rb = np.random.default_rng(6).standard_normal(g.n)
    assert _cg_iters(A, M, rb, 1e-4, itmax) > itmax
    assert _cg_iters(A, np.diag(1.0 / g.D), rb, 1e-4, 400) <= 400
    lv.spai.set_params(grid, bad)
    x = solver.solve(b, x0, method="spai", tol=1e-9, itmax=itmax, outer_max=400)
    st = solver.last_stats
    assert st["fallbacks"] >= 1 and max(st["relres"]) <= 1e-9, st


def test_spai_without_parameters_uses_jacobi(lv, oracle, gpu):
    grid, solver, g = _system(lv, oracle, K=16)
    b, _ = solver.rhs(2e-2, False)
    x = solver.solve(b, grid.get("P"), method="spai", tol=1e-9)
    assert solver.last_stats["fallbacks"] == 1 and max(solver.last_stats["relres"]) <= 1e-9


def test_training_on_the_gpu_lowers_the_loss(lv, oracle, gpu):
    """Adam through the hand-written forward and backward kernels, on systems harvested from
    plain-solver runs of three flows, lowers the probe loss."""
    th0 = lv.spai.init_params(0)
    th, losses, systems = lv.spai.train(th0, [("tg", 48, 1.0, 6), ("cavity", 40, 1.0, 4), ("tg", 48, 0.5, 6)],
                                        steps_per_system=8, lr=3e-4, nprobe=8)
    assert systems == 16
    first, last = np.mean(losses[:8]), np.mean(losses[-8:])
    assert last < 0.97 * first, (first, last)


def test_trained_network_gains_over_jacobi(lv, oracle, gpu):
    """The committed network (trained on 20,000 systems harvested from plain-solver runs of every
    configuration family, lagrangianvoronoi.jl_amd/train_spai.py) on a held-out system of the benchmark
    regime (Taylor-Green 128^2, c0 dt/dr of the 2048^2 benchmark, a seed not used in training): it meets
    the FP64 tolerance in at most 0.6x the inner iterations of Jacobi-PCG (the method without it), and
    in no more than the untrained network's.  A preconditioner on A's own pattern can shrink the
    condition number by at most 4x, i.e. halve the iterations; the untrained network is the Neumann
    form I + N, which already reaches about 1.9x, so training lowers the probe loss (by 41% over the
    20,000 systems) but the iterations by at most about 1% (169 against 169 here, 250 against 252 at
    256^2)."""
    th = lv.spai.load_params()
    assert th is not None
    g, desc, dt = lv.spai.flow("tg", 128, 1.0, seed=0x5EED0999)
    lv.stiffened_eos(g, 1.4, 2500.0 / 1.4)
    solver = lv.PressureSolver(g, method="mixed", tol=1e-8, itmax=4000)
    solver.assemble(dt)
    b, _ = solver.rhs(dt, False)
    x0 = g.get("P")
    solver.solve(b, x0, method="mixed")
    it_jac = solver.last_stats["iters"][0]
    its = {}
    for name, theta in (("untrained", lv.spai.init_params(0)), ("trained", th)):
        lv.spai.set_params(g, theta)
        solver.solve(b, x0, method="spai")
        assert max(solver.last_stats["relres"]) <= 1e-8 and solver.last_stats["fallbacks"] == 0
        its[name] = solver.last_stats["iters"][0]
    print("inner iterations: jacobi %d, untrained %d, trained %d" % (it_jac, its["untrained"], its["trained"]))
    assert its["trained"] <= 0.6 * it_jac, (its, it_jac)
    assert its["trained"] <= its["untrained"], its
