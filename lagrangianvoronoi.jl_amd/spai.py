"""Learned SPAI preconditioner: parameters, training on harvested pressure systems, evaluation.

The network and the preconditioner it emits are described in csrc/lv_spai.hip.  Parameters are a
flat float32 vector (layout in csrc/lv_spai.h).  Training data are the (operator, solution) pairs
that plain-solver runs of the method produce: every harvested system is the pressure operator of a
step of a seeded flow together with the pressure it was solved to.
"""
import ctypes
import os

import numpy as np

from . import _check, _ptr, lib, make_step_desc, PressureSolver, Rectangle, VoronoiGrid, remesh, step_fused
from . import inputs

H = 64
OFF_WN, OFF_BN, OFF_WE, OFF_BE = 0, 256, 320, 576


def OFF_W(l):
    return 640 + l * (192 * 64 + 64)


def OFF_B(l):
    return OFF_W(l) + 192 * 64


OFF_VE = 640 + 3 * (192 * 64 + 64)
OFF_CE = OFF_VE + 64
OFF_UN = OFF_CE + 1
OFF_CN = OFF_UN + 64
NPARAM = OFF_CN + 1

PARAMS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spai_params.npy")


def num_params():
    c = ctypes.c_int64()
    _check(None, lib().lv_spai_num_params(ctypes.byref(c)))
    assert c.value == NPARAM
    return c.value


def init_params(seed=0):
    """He-initialised encoders and layers; zero decoders, so the untrained network emits exactly the
    first-order Neumann approximate inverse of the Jacobi-scaled operator."""
    rng = np.random.default_rng(seed)
    th = np.zeros(NPARAM, dtype=np.float32)
    th[OFF_WN:OFF_BN] = rng.standard_normal(256) * np.sqrt(2.0 / 4)
    th[OFF_WE:OFF_BE] = rng.standard_normal(256) * np.sqrt(2.0 / 4)
    for l in range(3):
        th[OFF_W(l):OFF_B(l)] = rng.standard_normal(192 * 64) * np.sqrt(2.0 / 192) * 0.5
    return th


def load_params(path=PARAMS_PATH):
    """The committed trained parameters, or None."""
    if not os.path.exists(path):
        return None
    th = np.load(path, allow_pickle=False)
    assert th.dtype == np.float32 and th.size == NPARAM
    return th


def set_params(grid, theta):
    """Loads the network's parameters into the grid's context; None unloads them."""
    if theta is None:
        _check(grid.ctx, lib().lv_spai_set_params(grid.ctx, None))
        return
    th = np.ascontiguousarray(np.asarray(theta, dtype=np.float32))
    assert th.size == NPARAM
    _check(grid.ctx, lib().lv_spai_set_params(grid.ctx, _ptr(th)))


def get_params(grid):
    th = np.empty(NPARAM, dtype=np.float32)
    _check(grid.ctx, lib().lv_spai_get_params(grid.ctx, _ptr(th)))
    return th


def loss_grad(grid, nprobe=8, seed=1, use_data=True, want_grad=True):
    L = ctypes.c_double()
    g = np.empty(NPARAM, dtype=np.float32) if want_grad else None
    _check(grid.ctx, lib().lv_spai_loss_grad(grid.ctx, int(nprobe), ctypes.c_uint64(seed), int(bool(use_data)),
                                             ctypes.byref(L), _ptr(g) if g is not None else None))
    return L.value, g


def train_step(grid, lr=1e-3, nprobe=8, seed=1):
    L = ctypes.c_double()
    _check(grid.ctx, lib().lv_spai_train_step(grid.ctx, float(lr), int(nprobe), ctypes.c_uint64(seed),
                                              ctypes.byref(L)))
    return L.value


def outputs(grid, nnz):
    oe = np.empty(nnz, dtype=np.float32)
    on = np.empty(grid.n, dtype=np.float32)
    w = np.empty(nnz, dtype=np.float32)
    md = np.empty(grid.n, dtype=np.float32)
    _check(grid.ctx, lib().lv_spai_outputs(grid.ctx, _ptr(oe), _ptr(on), _ptr(w), _ptr(md)))
    return oe, on, w, md


# ---- harvesting and training ----------------------------------------------------------------------

def flow(kind, K, dt_scale=1.0, seed=0x5EED0100):
    """A small seeded flow of one configuration family: (grid, step descriptor, dt)."""
    if kind == "tg":
        dr = 1.0 / K
        x = inputs.jittered_lattice(K, K, dr, seed=seed)
        g = VoronoiGrid(Rectangle((0.0, 0.0), (1.0, 1.0)), dr, xperiodic=True, yperiodic=True, positions=x)
        remesh(g)
        ic = inputs.taylor_green(x, g.get("area"))
        for k in ("v", "rho", "mass", "P", "e", "mu"):
            g.set_soa(k, ic[k])
        dt = 2.5e-4 * (2048.0 / K) * dt_scale  # c0 dt / dr of the 2048^2 benchmark times dt_scale
        solver = PressureSolver(g, method="pcg64", tol=1e-10, picard_window=8)
        desc = make_step_desc(dt, eos="stiffened", gamma=ic["gamma"], P0=ic["P0"], solver=solver, niter=4)
        return g, desc, dt
    if kind == "cavity":
        dr = 1.0 / K
        x = inputs.jittered_lattice(K, K, dr, seed=seed)
        g = VoronoiGrid(Rectangle((0.0, 0.0), (1.0, 1.0)), dr, positions=x)
        remesh(g)
        ic = inputs.cavity(x, g.get("area"))
        for k in ("v", "rho", "mass", "mu", "c2"):
            g.set_soa(k, ic[k])
        dt = 0.1 * dr * dt_scale
        solver = PressureSolver(g, method="pcg64", tol=1e-10, picard_window=8)
        desc = make_step_desc(dt, eos=None, solver=solver, niter=4, friction={-1: (1.0, 0.0)})
        return g, desc, dt
    if kind == "gresho":
        dr = 1.0 / K
        x = inputs.jittered_lattice(K, K, dr, seed=seed)
        g = VoronoiGrid(Rectangle((0.0, 0.0), (1.0, 1.0)), dr, xperiodic=True, yperiodic=True, positions=x)
        remesh(g)
        ic = inputs.gresho(x, g.get("area"))
        for k in ("v", "rho", "mass", "P", "e", "c2"):
            g.set_soa(k, ic[k])
        dt = 0.1 * dr * dt_scale
        solver = PressureSolver(g, method="pcg64", tol=1e-10, picard_window=8)
        desc = make_step_desc(dt, eos=None, solver=solver, niter=4, viscous=False)
        return g, desc, dt
    if kind == "rt":
        dr = 1.0 / K
        x = inputs.jittered_lattice(K, 2 * K, dr, seed=seed)
        g = VoronoiGrid(Rectangle((0.0, 0.0), (1.0, 2.0)), dr, positions=x)
        remesh(g)
        ic = inputs.rayleigh_taylor(x, g.get("area"))
        for k in ("v", "rho", "mass", "P", "e", "mu", "phase"):
            g.set_soa(k, ic[k])
        dt = 0.05 * dr * dt_scale
        solver = PressureSolver(g, method="pcg64", tol=1e-10, picard_window=8)
        desc = make_step_desc(dt, eos="ideal", gamma=1.4, gravity=(0.0, -1.0), solver=solver, niter=4,
                              artificial_viscosity=True, friction=np.zeros(8))
        return g, desc, dt
    raise ValueError(kind)


def train(theta, schedule, steps_per_system=4, lr=1e-3, nprobe=8, log=None, seed=7, log_every=1):
    """Trains theta on systems harvested from plain-solver runs: for each (kind, K, dt_scale, nsteps)
    of the schedule, run the flow with Jacobi-PCG, and after every step assemble its pressure
    operator and take steps_per_system Adam steps on it (probes: the step's pressure and Rademacher
    vectors).  A flow that ends in one of the method's errors (a destroyed mesh, say) contributes the
    steps it made.  Returns (theta, losses, systems): systems counts the harvested operators."""
    from . import LVError

    losses = []
    th = np.asarray(theta, dtype=np.float32)
    systems = 0
    for it, (kind, K, dt_scale, nsteps) in enumerate(schedule):
        g, desc, dt = flow(kind, K, dt_scale, seed=seed + it)
        set_params(g, th)
        for s in range(nsteps):
            try:
                step_fused(g, desc)
            except LVError as e:
                if log:
                    log("%s K=%d dt*%g: flow ended at step %d (%s)" % (kind, K, dt_scale, s, e))
                break
            nnz = ctypes.c_int64()
            _check(g.ctx, lib().lv_assemble_pressure(g.ctx, float(dt), ctypes.byref(nnz)))
            for k in range(steps_per_system):
                L = train_step(g, lr=lr, nprobe=nprobe, seed=1000003 * it + 17 * s + k)
                losses.append(L)
            systems += 1
            if log and (s % log_every == 0 or s == nsteps - 1):
                log("%s K=%d dt*%g step %d loss %.4e" % (kind, K, dt_scale, s, losses[-1]))
        th = get_params(g)
    return th, losses, systems
