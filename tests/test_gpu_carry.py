"""Carried Picard directions (lv_solve_opts.picard_carry) across consecutive steps.

The previous find_pressure!'s first two corrections and its total change of P enter the next call's
window (multiplied by the new operator).  They only move where each solve starts: every solve still
meets the tolerance, the steps agree with the same steps without the carry to that tolerance, and
the early Picard solves of the later steps start from smaller residuals and take fewer iterations.
"""
import numpy as np
import pytest

from pairs import make_pair, rel_err

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _steps(lv, oracle, carry, nsteps, method="amg", K=128):
    grid, op, x, dr = make_pair(lv, oracle, "periodic", K=K, seed=0x5EED0061, ic="tg")
    solver = lv.PressureSolver(grid, method=method, tol=1e-10, itmax=4000, picard_window=8, picard_carry=carry)
    desc = lv.make_step_desc(1e-3, eos="stiffened", gamma=1.4, P0=2500.0 / 1.4, solver=solver, viscous=True,
                             artificial_viscosity=False, alpha=1.0, rusanov=True)
    stats = []
    for _ in range(nsteps):
        st = lv.step_fused(grid, desc)
        assert st["fallbacks"] == 0 and max(st["relres"]) <= 1e-10, st
        stats.append(st)
    return grid, stats


@pytest.mark.parametrize("method", ["amg", "pcg64"])
def test_carry_same_steps_fewer_iterations(lv, oracle, gpu, method):
    _flush("picard carry %s" % method)
    g0, s0 = _steps(lv, oracle, 0, 4, method)
    g1, s1 = _steps(lv, oracle, 1, 4, method)
    for f in ("P", "v", "x", "e"):
        assert rel_err(g1.get_soa(f), g0.get_soa(f)) <= 1e-7, f
    # the first step has nothing carried: identical solves
    assert s1[0]["iters"] == s0[0]["iters"]
    # later steps: Picard solves 2 and 3 start lower, the solves after the fourth are left to the
    # window as they were, and the steps take fewer iterations in all
    for k in (1, 2, 3):
        print("step %d relres0 without %s with %s" % (k + 1, ["%.1e" % v for v in s0[k]["relres0"][:5]],
                                                      ["%.1e" % v for v in s1[k]["relres0"][:5]]))
        print("step %d iterations without %s with %s" % (k + 1, s0[k]["iters"], s1[k]["iters"]))
    for k in (2, 3):
        assert s1[k]["relres0"][1] < 0.7 * s0[k]["relres0"][1], (k, s0[k]["relres0"], s1[k]["relres0"])
        assert s1[k]["relres0"][2] < 0.7 * s0[k]["relres0"][2], (k, s0[k]["relres0"], s1[k]["relres0"])
    it0 = sum(sum(st["iters"]) for st in s0[1:])
    it1 = sum(sum(st["iters"]) for st in s1[1:])
    print("iterations of steps 2-4: without carry %d, with %d" % (it0, it1))
    assert it1 < it0


def test_carry_dropped_for_a_new_particle_set(lv, oracle, gpu):
    """After set_positions (a new particle set, even of the same size) nothing is carried: the first
    find_pressure! takes exactly the iterations of a fresh grid."""
    _flush("picard carry reset")
    grid, op, x, dr = make_pair(lv, oracle, "periodic", K=64, seed=0x5EED0062, ic="tg")
    solver = lv.PressureSolver(grid, method="amg", tol=1e-10, picard_window=8, picard_carry=1)
    lv.stiffened_eos(grid, 1.4, 2500.0 / 1.4)
    st_a = lv.find_pressure(solver, 1e-3, 10)
    fresh, _, _, _ = make_pair(lv, oracle, "periodic", K=64, seed=0x5EED0062, ic="tg")
    grid.set_positions(fresh.get_soa("x"))
    lv.remesh(grid)
    for f in ("v", "rho", "mass", "P", "e", "mu"):
        grid.set_soa(f, fresh.get_soa(f))
    lv.stiffened_eos(grid, 1.4, 2500.0 / 1.4)
    st_b = lv.find_pressure(solver, 1e-3, 10)
    assert st_b["iters"] == st_a["iters"] and max(st_b["relres"]) <= 1e-10
