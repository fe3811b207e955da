"""The implicit viscous step (ViscousSolver, src/viscous_solver.jl:43-128) on the GPU against the oracle.

The operator is assembled in FP64 with the reference's operand order (compared at 1e-14); the solve
meets its FP64 tolerance on both sides, so velocities agree to the tolerance.  Pins: the viscous
operator has zero column sums off the diagonal mass term, so the implicit step conserves momentum
sum m v exactly (to the solve tolerance); a uniform velocity is a fixed point.
"""
import numpy as np
import pytest

from pairs import make_pair, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,method", [("periodic", "pcg64"), ("walls", "mixed"), ("walls", "amg")])
def test_implicit_viscous_step_matches_oracle(lv, oracle, gpu, kind, method):
    print("implicit viscous %s %s" % (kind, method), flush=True)
    grid, op, x, dr = make_pair(lv, oracle, kind, K=40, seed=0x5EED0700, ic="cavity")
    n = grid.n
    rng = np.random.default_rng(3)
    mu = 0.05 + 0.05 * rng.random(n)
    grid.set("mu", mu)
    op.set("mu", mu)
    dt = 2e-2
    mom0 = (grid.get("mass")[:, None] * grid.get("v")).sum(axis=0)
    vs = lv.ViscousSolver(grid, method=method, tol=1e-12, itmax=5000)
    st = lv.viscous_step(vs, dt)
    assert max(st["relres"]) <= 1e-12
    it = op.viscous_implicit(dt, 1e-13, 100000)
    assert it >= 0
    assert rel_err(grid.get_soa("v"), op["v"]) <= 1e-9
    mom1 = (grid.get("mass")[:, None] * grid.get("v")).sum(axis=0)
    assert np.abs(mom1 - mom0).max() <= 1e-10 * np.abs(grid.get("mass")[:, None] * grid.get("v")).sum()
    # the solver's system now holds the viscous operator: the pressure accessors refuse it
    with pytest.raises(lv.LVError):
        lv.PressureSolver(grid).operator()


def test_uniform_velocity_is_a_fixed_point(lv, oracle, gpu):
    """A 1 = (m/dt) 1 (the coupling part has zero row sums), so a uniform velocity solves the implicit
    step exactly; checked on the GPU, with the oracle's operator as the closed form's witness"""
    grid, op, x, dr = make_pair(lv, oracle, "walls", K=32, seed=0x5EED0701, ic="cavity")
    n = grid.n
    mu = np.full(n, 0.1)
    grid.set("mu", mu)
    op.set("mu", mu)
    dt = 1e-2
    diag, coef = op.viscous_operator(dt)
    # uniform velocity is a fixed point of the implicit step: A 1 = (m/dt) 1
    grid.set_soa("v", np.concatenate([np.full(n, 0.3), np.full(n, -0.2)]))
    st = lv.viscous_step(lv.ViscousSolver(grid, method="pcg64", tol=1e-12), dt)
    v = grid.get("v")
    assert np.abs(v[:, 0] - 0.3).max() <= 1e-11 and np.abs(v[:, 1] + 0.2).max() <= 1e-11
    ones = np.ones(n)
    assert rel_err(op.apply_A(diag, coef, ones), diag) <= 1e-14


def test_rayleigh_taylor_steps_with_implicit_viscosity(lv, oracle, gpu):
    """four Rayleigh-Taylor steps with the implicit viscous step, GPU against the oracle"""
    print("rayleigh-taylor implicit viscosity", flush=True)
    K = 20
    dr = 1.0 / K
    x = lv.inputs.jittered_lattice(K, 2 * K, dr, seed=0x5EED0004)
    grid = lv.VoronoiGrid(lv.Rectangle((0.0, 0.0), (1.0, 2.0)), dr, positions=x)
    lv.remesh(grid)
    op = oracle.Particles(oracle.Grid((0.0, 0.0), (1.0, 2.0), dr), x)
    op.remesh()
    ic = lv.inputs.rayleigh_taylor(x, op.area())
    for k in ("v", "rho", "mass", "P", "e", "mu", "phase"):
        grid.set_soa(k, ic[k])
        op.set(k, ic[k])
    dt = 2e-3
    solver = lv.PressureSolver(grid, tol=1e-12)
    vs = lv.ViscousSolver(grid, tol=1e-12)
    desc = lv.make_step_desc(dt, eos="ideal", gamma=1.4, gravity=(0.0, -1.0), solver=solver, viscous=vs)
    for _ in range(4):
        lv.step_fused(grid, desc)
        op.step(dt, eos="ideal", gamma=1.4, gravity=(0.0, -1.0), tol=1e-13, viscous="implicit", viscous_tol=1e-13)
    for k in ("x", "v", "P", "e", "mass"):
        assert rel_err(grid.get_soa(k), op[k]) <= 1e-8, k
