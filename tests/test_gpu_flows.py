"""Multi-step flows of every configuration family on the GPU, against the oracle step for step.

Both sides run the same seeded input through the same step order; the only non-bit-exact stage is
the iterative pressure (solved to 1e-12 on both sides), so trajectories agree to ~1e-9 over a few
steps.  Incompressible runs compare pressure up to its mean (tests/taylorgreen.jl:79-86 does the
same).  The reference's own acceptance test runs on the GPU with its thresholds.
"""
import math

import numpy as np
import pytest

import flows
from pairs import rel_err

pytestmark = pytest.mark.gpu


def _pair(lv, oracle, xmin, xmax, nx, ny, xp, yp, seed):
    dr = (xmax[0] - xmin[0]) / nx
    x = lv.inputs.jittered_lattice(nx, ny, dr, xmin=xmin, seed=seed)
    grid = lv.VoronoiGrid(lv.Rectangle(xmin, xmax), dr, xperiodic=xp, yperiodic=yp, positions=x)
    lv.remesh(grid)
    op = oracle.Particles(oracle.Grid(xmin, xmax, dr, xperiodic=xp, yperiodic=yp), x)
    op.remesh()
    return grid, op, x, dr


METHODS = ["pcg64", "amg", "amg-learned"]


def _solver(lv, grid, method, tol, **kw):
    """The pressure solver of a flow test; "amg-learned" loads the trained network as the multilevel
    cycle's fine smoother."""
    if method == "amg-learned":
        lv.spai.set_params(grid, lv.spai.load_params())
        method = "amg"
    return lv.PressureSolver(grid, method=method, tol=tol, **kw)


def _set(grid, op, fields):
    for k, v in fields.items():
        grid.set_soa(k, v)
        op.set(k, v)


def _compare(grid, op, names, tol, mean_free=()):
    for k in names:
        g = grid.get_soa(k)
        o = op[k]
        if k in mean_free:
            g = g - g.mean()
            o = o - o.mean()
        assert rel_err(g, o) <= tol, (k, rel_err(g, o))


@pytest.mark.parametrize("method", METHODS)
def test_gresho_incompressible_periodic(lv, oracle, gpu, method):
    """Config 1 family: Gresho vortex, c2 = Inf (singular pressure system), periodic."""
    print("gresho %s" % method, flush=True)
    grid, op, x, dr = _pair(lv, oracle, (0.0, 0.0), (1.0, 1.0), 40, 40, True, True, 0x5EED0001)
    ic = lv.inputs.gresho(x, op.area())
    _set(grid, op, {k: ic[k] for k in ("v", "rho", "mass", "P", "e", "c2")})
    dt = 1e-3
    solver = _solver(lv, grid, method, 1e-12)
    desc = lv.make_step_desc(dt, eos=None, solver=solver, viscous=False)
    for _ in range(4):
        lv.step_fused(grid, desc)
        op.step(dt, eos=None, tol=1e-13, viscous=False)
    _compare(grid, op, ["x", "v", "mass"], 1e-8)
    _compare(grid, op, ["P"], 1e-7, mean_free=("P",))


@pytest.mark.parametrize("method", METHODS)
def test_rayleigh_taylor_walls_gravity(lv, oracle, gpu, method):
    """Config 4 family: two phases (density 3:1), gravity, ideal gas, no-slip walls via
    bdary_friction!, artificial viscosity on (examples/rayleightaylor.jl:90-101 without the
    multiphase projection)."""
    print("rayleigh-taylor %s" % method, flush=True)
    grid, op, x, dr = _pair(lv, oracle, (0.0, 0.0), (1.0, 2.0), 20, 40, False, False, 0x5EED0004)
    ic = lv.inputs.rayleigh_taylor(x, op.area())
    _set(grid, op, {k: ic[k] for k in ("v", "rho", "mass", "P", "e", "mu", "phase")})
    dt = 2e-3
    zero = np.zeros(8)
    solver = _solver(lv, grid, method, 1e-12)
    desc = lv.make_step_desc(dt, eos="ideal", gamma=1.4, gravity=(0.0, -1.0), solver=solver, viscous=True,
                             artificial_viscosity=True, friction=zero)
    for _ in range(4):
        lv.step_fused(grid, desc)
        op.step(dt, eos="ideal", gamma=1.4, gravity=(0.0, -1.0), tol=1e-13, viscous=True,
                artificial_viscosity=True, friction=zero)
    _compare(grid, op, ["x", "v", "P", "e", "mass"], 1e-8)
    assert np.array_equal(grid.get("phase"), op["phase"])


@pytest.mark.parametrize("method", METHODS)
def test_lid_driven_cavity(lv, oracle, gpu, method):
    """Config 5 family: incompressible cavity, lid u = 1 on BDARY_UP through bdary_friction!
    (examples/cavity.jl:70-80)."""
    print("cavity %s" % method, flush=True)
    grid, op, x, dr = _pair(lv, oracle, (0.0, 0.0), (1.0, 1.0), 32, 32, False, False, 0x5EED0005)
    ic = lv.inputs.cavity(x, op.area())
    _set(grid, op, {k: ic[k] for k in ("v", "rho", "mass", "mu", "c2")})
    lid = {lv.BDARY_UP: (1.0, 0.0)}
    vbc = lv._vbc_table(lid)
    dt = 1e-3
    solver = _solver(lv, grid, method, 1e-12)
    desc = lv.make_step_desc(dt, eos=None, solver=solver, viscous=True, artificial_viscosity=False, friction=lid)
    for _ in range(4):
        lv.step_fused(grid, desc)
        op.step(dt, eos=None, tol=1e-13, viscous=True, artificial_viscosity=False, friction=vbc)
    _compare(grid, op, ["x", "v", "mass"], 1e-8)
    _compare(grid, op, ["P"], 1e-7, mean_free=("P",))
    v = grid.get("v")
    assert v[:, 0].max() > 0.01  # the lid drags the top row


@pytest.mark.parametrize("method", ["pcg64", "mixed", "amg", "amg-learned"])
def test_taylor_green_reference_thresholds_on_gpu(lv, gpu, method):
    """tests/taylorgreen.jl end to end on the GPU: N = 80 hexagonal lattice, 80 steps, and the
    reference's asserts E_err < 1e-8, v_err < 0.01, P_err < 0.01."""
    print("taylor-green reference %s" % method, flush=True)
    x = flows.populate_hex((0.0, 0.0), (1.0, 1.0), flows.TG_DR)
    grid = lv.VoronoiGrid(lv.UnitRectangle(), flows.TG_DR, xperiodic=True, yperiodic=True, positions=x)
    lv.remesh(grid)
    for k, v in flows.tg_ic(x, grid.get("area")).items():
        grid.set_soa(k, v)
    solver = _solver(lv, grid, method, 1e-10, picard_window=8)
    desc = lv.make_step_desc(flows.TG_DT, eos="stiffened", gamma=flows.TG_GAMMA, P0=flows.TG_P0, solver=solver,
                             viscous=True, artificial_viscosity=False)
    E0, last = None, None
    for k, t, frame in flows.tg_frame_schedule():
        lv.step_fused(grid, desc)
        if frame:
            E, ve, pe = flows.tg_errors(grid.get_soa("x"), grid.get_soa("v"), grid.get("P"), grid.get("area"),
                                        grid.get("mass"), grid.get("e"), t)
            E0 = E if E0 is None else E0
            last = (E - E0, ve, pe)
    E_err, v_err, P_err = last
    assert E_err < 1e-8 and v_err < 0.01 and P_err < 0.01, last
    assert abs(E_err) < 1e-10
