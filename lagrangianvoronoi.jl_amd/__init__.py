"""Python mirror of LagrangianVoronoi.jl's hot-path API over liblvhip.so (MI355X / gfx950).

Names follow the reference with the Julia `!` dropped: `remesh(grid)` is `remesh!(grid)`,
`find_pressure(solver, dt)` is `find_pressure!(solver, dt)`, and so on.  Every operation runs
through the HIP library; there is no CPU fallback: importing this package on a machine without
the built library raises, and the library reports a HIP error when no GPU is present.

Fields use the reference's PolygonNS names (src/celldefs.jl:7-39).  Arrays are numpy, shaped
(n,) for scalars, (n, 2) for vectors and (n, 2, 2) for the tensor D; labels are 0-based and
boundary labels are BDARY_UP = -1, BDARY_RIGHT = -2, BDARY_DOWN = -3, BDARY_LEFT = -4.
"""
import ctypes
import os

import numpy as np

from . import inputs  # noqa: F401

__all__ = [
    "LVError", "Rectangle", "UnitRectangle", "VoronoiGrid", "GridNS", "PressureSolver", "remesh", "move",
    "stiffened_eos", "ideal_eos", "find_rho", "gravity_step", "find_pressure", "pressure_step", "find_D",
    "viscous_step", "bdary_friction", "find_dv", "relaxation_step", "step_fused", "StepDesc",
    "BDARY_UP", "BDARY_RIGHT", "BDARY_DOWN", "BDARY_LEFT", "lib", "library_path", "export_grid", "export_points",
    "SimulationWorkspace", "run", "movingavg", "nearest_polygon", "movingls", "point_value", "poly_eval",
    "power_vector", "integral", "LinearExpansion", "QuadraticExpansion", "CubicExpansion", "populate_circ",
    "populate_hex", "populate_lloyd", "populate_rand", "populate_rect", "populate_vogel",
]

BDARY_UP, BDARY_RIGHT, BDARY_DOWN, BDARY_LEFT = -1, -2, -3, -4

LV_OK = 0
LV_E_MESH_DESTROYED = 1
LV_E_NAN_VELOCITY = 2
LV_E_POLY_OVERFLOW = 3
LV_E_NOT_CONVERGED = 4
LV_E_HIP = 5
LV_E_INVALID = 6
LV_E_NOMEM = 7
LV_E_NO_MESH = 8
LV_MAX_EDGES = 16

SOLVERS = {"pcg64": 0, "mixed": 1, "spai": 2, "amg": 3}

FIELDS = {"x": (0, 2), "v": (1, 2), "dv": (2, 2), "momentum": (3, 2), "rho": (4, 1), "e": (5, 1), "P": (6, 1),
          "c2": (7, 1), "mass": (8, 1), "energy": (9, 1), "quality": (10, 1), "mu": (11, 1), "D": (12, 4),
          "phase": (13, 1), "area": (14, 1), "centroid": (15, 2)}

library_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblvhip.so")


class LVError(RuntimeError):
    """An error status of liblvhip; the message of the reference's throw where one exists."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


# ---- C structures (include/lvhip.h) ---------------------------------------------------------------

class GridDesc(ctypes.Structure):
    _fields_ = [("xmin", ctypes.c_double * 2), ("xmax", ctypes.c_double * 2), ("dr", ctypes.c_double),
                ("h", ctypes.c_double), ("r_max", ctypes.c_double), ("xperiodic", ctypes.c_int32),
                ("yperiodic", ctypes.c_int32), ("capacity", ctypes.c_int32)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("tol", ctypes.c_double), ("itmax", ctypes.c_int32),
                ("inner_tol", ctypes.c_double), ("outer_max", ctypes.c_int32), ("vbc", ctypes.c_double * 8),
                ("picard_window", ctypes.c_int32), ("amg_layout", ctypes.c_int32), ("picard_carry", ctypes.c_int32),
                ("amg_level_kernels", ctypes.c_int32)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int32 * 32), ("outer", ctypes.c_int32 * 32), ("relres", ctypes.c_double * 32),
                ("relres0", ctypes.c_double * 32), ("solve_ms", ctypes.c_double * 32), ("total_iters", ctypes.c_int32),
                ("fallbacks", ctypes.c_int32), ("nsolves", ctypes.c_int32), ("multiphase_iters", ctypes.c_int32),
                ("multiphase_relres", ctypes.c_double), ("agg_order", ctypes.c_int32)]

    def as_dict(self):
        k = min(self.nsolves, 32)
        return dict(iters=list(self.iters[:k]), outer=list(self.outer[:k]), relres=list(self.relres[:k]),
                    relres0=list(self.relres0[:k]),
                    solve_ms=list(self.solve_ms[:k]), total_iters=self.total_iters, fallbacks=self.fallbacks,
                    nsolves=self.nsolves, multiphase_iters=self.multiphase_iters,
                    multiphase_relres=self.multiphase_relres, agg_order=self.agg_order)


class StepDesc(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("eos", ctypes.c_int32), ("gamma", ctypes.c_double),
                ("P0", ctypes.c_double), ("Pmin", ctypes.c_double), ("gravity", ctypes.c_int32),
                ("g", ctypes.c_double * 2), ("niter", ctypes.c_int32), ("solve", SolveOpts),
                ("viscous", ctypes.c_int32), ("artificial_viscosity", ctypes.c_int32), ("friction", ctypes.c_int32),
                ("friction_vbc", ctypes.c_double * 8), ("alpha", ctypes.c_double), ("rusanov", ctypes.c_int32),
                ("multiphase", ctypes.c_int32), ("multiphase_quality", ctypes.c_double),
                ("multiphase_tol", ctypes.c_double), ("multiphase_itmax", ctypes.c_int32),
                ("viscous_tol", ctypes.c_double)]


_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int32

_SIGNATURES = {
    "lv_create": [ctypes.POINTER(GridDesc), ctypes.POINTER(ctypes.c_void_p)],
    "lv_destroy": [_P],
    "lv_sync": [_P],
    "lv_grid_info": [_P, _P, _P, _P, _P],
    "lv_grid_path": [_P, _P, _P, _P],
    "lv_set_particles": [_P, _I, _P],
    "lv_num_particles": [_P, _P],
    "lv_upload": [_P, _I, _P],
    "lv_download": [_P, _I, _P],
    "lv_field_device_ptr": [_P, _I, ctypes.POINTER(ctypes.c_void_p)],
    "lv_remesh": [_P],
    "lv_mesh_size": [_P, _P],
    "lv_get_mesh": [_P, _P, _P, _P],
    "lv_cell_list_size": [_P, _P],
    "lv_get_cell_list": [_P, _P, _P],
    "lv_move": [_P, _D],
    "lv_gravity_step": [_P, _P, _D],
    "lv_find_rho": [_P],
    "lv_eos_stiffened": [_P, _D, _D],
    "lv_eos_ideal": [_P, _D, _D],
    "lv_find_pressure": [_P, _D, _I, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats)],
    "lv_pressure_step": [_P, _D],
    "lv_find_D": [_P],
    "lv_viscous_step": [_P, _D, _I],
    "lv_bdary_friction": [_P, _P, _D],
    "lv_find_dv": [_P, _D, _D],
    "lv_relaxation_step": [_P, _D, _I],
    "lv_step_fused": [_P, ctypes.POINTER(StepDesc), ctypes.POINTER(SolveStats)],
    "lv_assemble_pressure": [_P, _D, _P],
    "lv_get_pressure_operator": [_P, _P, _P, _P, _P],
    "lv_pressure_rhs": [_P, _D, _I, _P, _P, _P],
    "lv_solve_pressure": [_P, _P, _P, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats)],
    "lv_apply_preconditioner": [_P, _I, _P, _P],
    "lv_multiphase_projection": [_P, _D, _D, _D, _I, _P, _P],
    "lv_viscous_implicit": [_P, _D, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveStats)],
    "lv_multiphase_rhs": [_P, _P],
    "lv_multiphase_apply": [_P, _P, _P],
    "lv_spai_num_params": [_P],
    "lv_spai_set_params": [_P, _P],
    "lv_spai_get_params": [_P, _P],
    "lv_spai_train_step": [_P, _D, _I, ctypes.c_uint64, _P],
    "lv_spai_loss_grad": [_P, _I, ctypes.c_uint64, _I, _P, _P],
    "lv_spai_outputs": [_P, _P, _P, _P, _P],
    "lv_lloyd": [_P, _I],
    "lv_nearest_polygon": [_P, _I, _P, _P],
    "lv_movingls": [_P, _I, _I, _I, _D, _P],
    "lv_point_values": [_P, _I, _P, _I, _I, _P],
    "lv_point_values_noalloc": [_P, _I, _P, _I, _I, _P],
    "lv_export_grid": [_P, ctypes.c_char_p, _P, _I, _I],
    "lv_export_points": [_P, ctypes.c_char_p, _P, _I, _I],
    "lv_timing_enable": [_P, _I, _I],
    "lv_timing_read": [_P, _I, _P, _P, _P],
    "lv_timing_seen": [_P, _I, _P],
}

_lib = None


def lib():
    """The loaded liblvhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise ImportError("liblvhip.so is not built: run `python __graft_entry__.py` or "
                              "`python lagrangianvoronoi.jl_amd/build.py` first")
        L = ctypes.CDLL(library_path)
        for name, args in _SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.lv_last_error.argtypes = [_P]
        L.lv_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(ctx, rc):
    if rc != LV_OK:
        msg = lib().lv_last_error(ctx).decode() if ctx else "invalid argument"
        raise LVError(rc, msg or "liblvhip error %d" % rc)
    return rc


# ---- geometry -------------------------------------------------------------------------------------

class Rectangle:
    """Rectangle(xmin, xmax), src/geometry.jl:94-111"""

    def __init__(self, xmin=None, xmax=None, xlims=None, ylims=None):
        if xlims is not None:
            xmin, xmax = (xlims[0], ylims[0]), (xlims[1], ylims[1])
        self.xmin = (float(xmin[0]), float(xmin[1]))
        self.xmax = (float(xmax[0]), float(xmax[1]))


def UnitRectangle():
    return Rectangle((0.0, 0.0), (1.0, 1.0))


def _vbc_table(spec):
    """Per-label boundary velocity table vbc[2*(-label-1) + comp] from a dict {label: (vx, vy)},
    an (8,) array, or a callable f(x, label) that is constant on each boundary side."""
    t = np.zeros(8)
    if spec is None:
        return t
    if isinstance(spec, dict):
        for lab, val in spec.items():
            t[2 * (-lab - 1):2 * (-lab - 1) + 2] = val
        return t
    if callable(spec):
        for lab in (BDARY_UP, BDARY_RIGHT, BDARY_DOWN, BDARY_LEFT):
            t[2 * (-lab - 1):2 * (-lab - 1) + 2] = spec(None, lab)
        return t
    return np.asarray(spec, dtype=np.float64).reshape(8).copy()


class VoronoiGrid:
    """VoronoiGrid{T}(boundary_rect, dr; h = 2dr, r_max = 10dr, xperiodic, yperiodic),
    src/voronoigrid.jl:14-50, with its polygons held on the GPU."""

    def __init__(self, boundary_rect, dr, h=None, r_max=None, xperiodic=False, yperiodic=False, capacity=None,
                 positions=None):
        self.boundary_rect = boundary_rect
        self.dr = float(dr)
        self.h = 2.0 * self.dr if h is None else float(h)
        self.r_max = 10.0 * self.dr if r_max is None else float(r_max)
        self.xperiodic, self.yperiodic = bool(xperiodic), bool(yperiodic)
        if capacity is None:
            capacity = 0 if positions is None else np.asarray(positions).size // 2
        self._create(capacity)
        if positions is not None:
            self.set_positions(positions)

    def reserve(self, n):
        """room for n seeds; an empty grid (the populate_* pattern) is re-created with it"""
        if n <= self.capacity:
            return
        if self.n:
            raise ValueError("the grid holds %d seeds and room for %d" % (self.n, self.capacity))
        if self._ctx and _lib is not None:
            _lib.lv_destroy(self._ctx)
        self._ctx = None
        self._create(n)

    def _create(self, capacity):
        boundary_rect = self.boundary_rect
        d = GridDesc()
        d.xmin[:] = boundary_rect.xmin
        d.xmax[:] = boundary_rect.xmax
        d.dr, d.h, d.r_max = self.dr, self.h, self.r_max
        d.xperiodic, d.yperiodic = int(self.xperiodic), int(self.yperiodic)
        d.capacity = max(int(capacity), 1)
        self.capacity = d.capacity
        h_ = ctypes.c_void_p()
        rc = lib().lv_create(ctypes.byref(d), ctypes.byref(h_))
        if rc != LV_OK:
            raise LVError(rc, "lv_create failed with status %d" % rc)
        self._ctx = h_
        self.n = 0

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx and _lib is not None:
            _lib.lv_destroy(ctx)
            self._ctx = None

    @property
    def ctx(self):
        return self._ctx

    # push! of many seeds at once, then the caller remeshes (src/populate.jl pattern)
    def set_positions(self, x):
        x = np.asarray(x, dtype=np.float64)
        if x.ndim == 2:
            x = np.ascontiguousarray(x.T).reshape(-1)
        x = np.ascontiguousarray(x)
        n = x.size // 2
        _check(self._ctx, lib().lv_set_particles(self._ctx, n, _ptr(x)))
        self.n = n

    def get(self, name):
        fid, comps = FIELDS[name]
        dtype = np.int32 if name == "phase" else np.float64
        out = np.empty(comps * self.n, dtype=dtype)
        _check(self._ctx, lib().lv_download(self._ctx, fid, _ptr(out)))
        if comps == 1:
            return out
        if comps == 2:
            return out.reshape(2, self.n).T.copy()
        return out.reshape(2, 2, self.n).transpose(2, 1, 0).copy()  # (n, row, col), column-major planes

    def get_soa(self, name):
        """The raw SoA planes of a field, as the C ABI holds them."""
        fid, comps = FIELDS[name]
        dtype = np.int32 if name == "phase" else np.float64
        out = np.empty(comps * self.n, dtype=dtype)
        _check(self._ctx, lib().lv_download(self._ctx, fid, _ptr(out)))
        return out

    def set(self, name, value):
        fid, comps = FIELDS[name]
        dtype = np.int32 if name == "phase" else np.float64
        a = np.asarray(value, dtype=dtype)
        if a.ndim == 0:
            a = np.full(comps * self.n, a, dtype=dtype)
        elif comps == 2 and a.ndim == 2:
            a = a.T.reshape(-1)
        elif comps == 4 and a.ndim == 3:
            a = a.transpose(2, 1, 0).reshape(-1)
        a = np.ascontiguousarray(a)
        if a.size != comps * self.n:
            raise ValueError("field %s expects %d values" % (name, comps * self.n))
        _check(self._ctx, lib().lv_upload(self._ctx, fid, _ptr(a)))

    def set_soa(self, name, planes):
        fid, comps = FIELDS[name]
        dtype = np.int32 if name == "phase" else np.float64
        a = np.ascontiguousarray(np.asarray(planes, dtype=dtype).reshape(-1))
        _check(self._ctx, lib().lv_upload(self._ctx, fid, _ptr(a)))

    def mesh(self):
        """(edge_ptr[n+1], labels[E], verts[E,4]) with every polygon's edges after sort_edges!"""
        ne = ctypes.c_int64()
        _check(self._ctx, lib().lv_mesh_size(self._ctx, ctypes.byref(ne)))
        ptr = np.empty(self.n + 1, dtype=np.int32)
        lab = np.empty(ne.value, dtype=np.int32)
        ver = np.empty((ne.value, 4), dtype=np.float64)
        _check(self._ctx, lib().lv_get_mesh(self._ctx, _ptr(ptr), _ptr(lab), _ptr(ver)))
        return ptr, lab, ver

    def cell_list(self):
        """(cell_start[n1*n2+1], cell_labels) of the last remesh"""
        n1, n2, npath = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(self._ctx, lib().lv_grid_info(self._ctx, ctypes.byref(n1), ctypes.byref(n2), ctypes.byref(npath), None))
        m = ctypes.c_int64()
        _check(self._ctx, lib().lv_cell_list_size(self._ctx, ctypes.byref(m)))
        cs = np.empty(n1.value * n2.value + 1, dtype=np.int32)
        cl = np.empty(m.value, dtype=np.int32)
        _check(self._ctx, lib().lv_get_cell_list(self._ctx, _ptr(cs), _ptr(cl)))
        return cs, cl

    def info(self):
        n1, n2, npath = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        out6 = np.empty(6)
        _check(self._ctx, lib().lv_grid_info(self._ctx, ctypes.byref(n1), ctypes.byref(n2), ctypes.byref(npath),
                                             _ptr(out6)))
        return dict(n1=n1.value, n2=n2.value, npath=npath.value, cropping_min=tuple(out6[:2]),
                    cropping_max=tuple(out6[2:4]), next_rr=out6[4], rr_max=out6[5])

    def magic_path(self):
        npath = self.info()["npath"]
        i1 = np.empty(npath, dtype=np.int32)
        i2 = np.empty(npath, dtype=np.int32)
        rr = np.empty(npath)
        _check(self._ctx, lib().lv_grid_path(self._ctx, _ptr(i1), _ptr(i2), _ptr(rr)))
        return i1, i2, rr

    def sync(self):
        _check(self._ctx, lib().lv_sync(self._ctx))


GridNS = VoronoiGrid


# ---- operators (src/voronoigrid.jl, src/move.jl, src/pressure.jl, src/diffusion.jl, src/relaxation.jl) --

def remesh(grid):
    _check(grid.ctx, lib().lv_remesh(grid.ctx))


def move(grid, dt):
    _check(grid.ctx, lib().lv_move(grid.ctx, float(dt)))


def gravity_step(grid, g, dt):
    g = np.ascontiguousarray(np.asarray(g, dtype=np.float64).reshape(2))
    _check(grid.ctx, lib().lv_gravity_step(grid.ctx, _ptr(g), float(dt)))


def find_rho(grid):
    _check(grid.ctx, lib().lv_find_rho(grid.ctx))


def stiffened_eos(grid, gamma=1.4, P0=0.0):
    _check(grid.ctx, lib().lv_eos_stiffened(grid.ctx, float(gamma), float(P0)))


def ideal_eos(grid, gamma=1.4, Pmin=0.0):
    _check(grid.ctx, lib().lv_eos_ideal(grid.ctx, float(gamma), float(Pmin)))


def pressure_step(grid, dt):
    _check(grid.ctx, lib().lv_pressure_step(grid.ctx, float(dt)))


def find_D(grid):
    _check(grid.ctx, lib().lv_find_D(grid.ctx))


class ViscousSolver:
    """ViscousSolver(grid; verbose), src/viscous_solver.jl:43-67: the implicit viscous step.  Its
    system is solved with this build's inner solver (`method`), to the reference's minres!
    tolerance 1e-6 with itmax 2000 by default."""

    def __init__(self, grid, verbose=0, method="mixed", tol=1e-6, itmax=2000, inner_tol=1e-4, outer_max=20):
        self.grid = grid
        self.verbose = verbose
        self.method, self.tol, self.itmax = method, tol, itmax
        self.inner_tol, self.outer_max = inner_tol, outer_max
        self.last_stats = None

    def opts(self):
        return solve_opts(self.method, self.tol, self.itmax, self.inner_tol, self.outer_max)


def viscous_step(obj, dt, artificial_viscosity=True):
    """viscous_step!(grid, dt; artificial_viscosity) (explicit, src/diffusion.jl:39-53), or
    viscous_step!(solver::ViscousSolver, dt) (implicit, src/viscous_solver.jl:120-128)."""
    if isinstance(obj, ViscousSolver):
        o = obj.opts()
        st = SolveStats()
        _check(obj.grid.ctx, lib().lv_viscous_implicit(obj.grid.ctx, float(dt), ctypes.byref(o), ctypes.byref(st)))
        obj.last_stats = st.as_dict()
        return obj.last_stats
    _check(obj.ctx, lib().lv_viscous_step(obj.ctx, float(dt), int(bool(artificial_viscosity))))


def _wall_midpoints(grid):
    """the midpoint of each wall of the boundary rectangle, by label"""
    (x0, y0), (x1, y1) = grid.boundary_rect.xmin, grid.boundary_rect.xmax
    xm, ym = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
    return {BDARY_UP: np.array([xm, y1]), BDARY_RIGHT: np.array([x1, ym]), BDARY_DOWN: np.array([xm, y0]),
            BDARY_LEFT: np.array([x0, ym])}


def bdary_friction(grid, vDirichlet, dt, charfun=None):
    """bdary_friction!(grid, vDirichlet, dt; charfun), src/diffusion.jl:64-78.  vDirichlet as in
    _vbc_table; charfun a callable of a point, evaluated at each wall's midpoint (walls where it is
    false get no friction), or a collection of the labels that get it; None: every wall."""
    t = _vbc_table(vDirichlet)
    if charfun is not None:
        mids = _wall_midpoints(grid)
        for lab in (BDARY_UP, BDARY_RIGHT, BDARY_DOWN, BDARY_LEFT):
            on = charfun(mids[lab]) if callable(charfun) else lab in charfun
            if not on:
                t[2 * (-lab - 1):2 * (-lab - 1) + 2] = np.nan
    _check(grid.ctx, lib().lv_bdary_friction(grid.ctx, _ptr(t), float(dt)))


def find_dv(grid, dt, alpha=1.0):
    _check(grid.ctx, lib().lv_find_dv(grid.ctx, float(dt), float(alpha)))


def relaxation_step(grid, dt, rusanov=True):
    _check(grid.ctx, lib().lv_relaxation_step(grid.ctx, float(dt), int(bool(rusanov))))


def solve_opts(method="pcg64", tol=1e-10, itmax=1000, inner_tol=1e-4, outer_max=20, boundary_velocity=None,
               picard_window=0, amg_layout=0, picard_carry=0, amg_level_kernels=0):
    o = SolveOpts()
    o.method = SOLVERS[method] if isinstance(method, str) else int(method)
    o.tol, o.itmax, o.inner_tol, o.outer_max = float(tol), int(itmax), float(inner_tol), int(outer_max)
    o.vbc[:] = _vbc_table(boundary_velocity)
    o.picard_window = int(picard_window)
    o.amg_layout = int(amg_layout)
    o.picard_carry = int(bool(picard_carry))
    o.amg_level_kernels = int(bool(amg_level_kernels))
    return o


class PressureSolver:
    """PressureSolver(grid; verbose), src/pressure.jl:142-159.  The Krylov method and tolerance are
    this build's (the reference hard-codes MINRES with atol = rtol = 1e-6, :219)."""

    def __init__(self, grid, verbose=False, method="pcg64", tol=1e-10, itmax=1000, inner_tol=1e-4, outer_max=20,
                 picard_window=0, amg_layout=0, picard_carry=0, amg_level_kernels=0):
        self.grid = grid
        self.verbose = verbose
        self.method, self.tol, self.itmax = method, tol, itmax
        self.inner_tol, self.outer_max = inner_tol, outer_max
        self.picard_window = picard_window
        self.amg_layout = amg_layout
        self.picard_carry = picard_carry
        self.amg_level_kernels = amg_level_kernels
        self.last_stats = None

    def opts(self, boundary_velocity=None):
        return solve_opts(self.method, self.tol, self.itmax, self.inner_tol, self.outer_max, boundary_velocity,
                          self.picard_window, self.amg_layout, self.picard_carry, self.amg_level_kernels)

    # PressureOperator access (src/pressure.jl:89-135)
    def assemble(self, dt):
        nnz = ctypes.c_int64()
        _check(self.grid.ctx, lib().lv_assemble_pressure(self.grid.ctx, float(dt), ctypes.byref(nnz)))
        return nnz.value

    def operator(self):
        """(diag[n], row_ptr[n+1], cols[nnz], coef[nnz]) of the assembled operator"""
        n = self.grid.n
        diag = np.empty(n)
        rp = np.empty(n + 1, dtype=np.int32)
        _check(self.grid.ctx, lib().lv_get_pressure_operator(self.grid.ctx, _ptr(diag), _ptr(rp), None, None))
        nnz = int(rp[-1])
        cols = np.empty(nnz, dtype=np.int32)
        coef = np.empty(nnz)
        _check(self.grid.ctx, lib().lv_get_pressure_operator(self.grid.ctx, _ptr(diag), _ptr(rp), _ptr(cols),
                                                             _ptr(coef)))
        return diag, rp, cols, coef

    def rhs(self, dt, gp_step, boundary_velocity=None):
        n = self.grid.n
        b = np.empty(n)
        GP = np.empty(2 * n)
        t = _vbc_table(boundary_velocity)
        _check(self.grid.ctx, lib().lv_pressure_rhs(self.grid.ctx, float(dt), int(bool(gp_step)), _ptr(t), _ptr(b),
                                                    _ptr(GP)))
        return b, GP

    def apply_preconditioner(self, r, method=None):
        """z = B r with the FP32 preconditioner of the inner solve of `method` ("mixed": Jacobi,
        "spai": the learned M, "amg": the multilevel cycle) on the assembled operator."""
        m = self.method if method is None else method
        r = np.ascontiguousarray(np.asarray(r, dtype=np.float64))
        z = np.empty_like(r)
        _check(self.grid.ctx, lib().lv_apply_preconditioner(self.grid.ctx, SOLVERS[m] if isinstance(m, str) else int(m),
                                                             _ptr(r), _ptr(z)))
        return z

    def solve(self, b, x0, raise_on_fail=True, **kw):
        o = solve_opts(**{**dict(method=self.method, tol=self.tol, itmax=self.itmax, inner_tol=self.inner_tol,
                                  outer_max=self.outer_max), **kw})
        x = np.ascontiguousarray(np.array(x0, dtype=np.float64))
        b = np.ascontiguousarray(np.asarray(b, dtype=np.float64))
        st = SolveStats()
        rc = lib().lv_solve_pressure(self.grid.ctx, _ptr(b), _ptr(x), ctypes.byref(o), ctypes.byref(st))
        if rc != LV_E_NOT_CONVERGED or raise_on_fail:
            _check(self.grid.ctx, rc)
        self.last_stats = st.as_dict()
        return x


def find_pressure(solver, dt, niter=10, boundary_velocity=None):
    """find_pressure!(solver, dt, niter; boundary_velocity), src/pressure.jl:215-224.  Returns the
    solve statistics.  Unlike the reference (which ignores MINRES convergence), a missed tolerance
    raises LVError(LV_E_NOT_CONVERGED) after P has been written."""
    st = SolveStats()
    o = solver.opts(boundary_velocity)
    rc = lib().lv_find_pressure(solver.grid.ctx, float(dt), int(niter), ctypes.byref(o), ctypes.byref(st))
    solver.last_stats = st.as_dict()
    _check(solver.grid.ctx, rc)
    return solver.last_stats


def step_fused(grid, desc, stats=True):
    """One whole step! on the device (move! ... relaxation_step!), see StepDesc."""
    st = SolveStats()
    rc = lib().lv_step_fused(grid.ctx, ctypes.byref(desc), ctypes.byref(st) if stats else None)
    _check(grid.ctx, rc)
    return st.as_dict()


def make_step_desc(dt, eos=None, gamma=1.4, P0=0.0, Pmin=0.0, gravity=None, niter=10, solver=None,
                   viscous=True, artificial_viscosity=False, friction=None, alpha=1.0, rusanov=True,
                   boundary_velocity=None, multiphase=None):
    """The step! of tests/taylorgreen.jl:61-71 and its relatives as one device call; multiphase = a
    MultiphaseSolver (or its quality_threshold) runs multiphase_projection! after find_dv!
    (examples/rayleightaylor.jl:99)."""
    d = StepDesc()
    d.dt = float(dt)
    d.eos = {None: 0, "stiffened": 1, "ideal": 2, "rho": 3}[eos]
    d.gamma, d.P0, d.Pmin = float(gamma), float(P0), float(Pmin)
    d.gravity = int(gravity is not None)
    d.g[:] = (0.0, 0.0) if gravity is None else tuple(gravity)
    d.niter = int(niter)
    d.solve = solver.opts(boundary_velocity) if solver is not None else solve_opts(boundary_velocity=boundary_velocity)
    d.viscous = 2 if isinstance(viscous, ViscousSolver) else int(bool(viscous))
    d.viscous_tol = float(viscous.tol) if isinstance(viscous, ViscousSolver) else 1e-6
    d.artificial_viscosity = int(bool(artificial_viscosity))
    d.friction = int(friction is not None)
    d.friction_vbc[:] = _vbc_table(friction)
    d.alpha = float(alpha)
    d.rusanov = int(bool(rusanov))
    if multiphase is None:
        d.multiphase, d.multiphase_quality, d.multiphase_tol, d.multiphase_itmax = 0, 0.25, 1e-4, 200
    else:
        d.multiphase = 1
        d.multiphase_quality = float(getattr(multiphase, "quality_threshold", multiphase))
        d.multiphase_tol = float(getattr(multiphase, "atol", 1e-4))
        d.multiphase_itmax = int(getattr(multiphase, "itmax", 200))
    return d


class MultiphaseSolver:
    """MultiphaseSolver(grid; quality_threshold = 0.25), src/relaxation.jl:134-163.  The linear solve
    of multiphase_projection! uses the reference's tolerances (atol = rtol = 1e-4, itmax = 200)."""

    def __init__(self, grid, quality_threshold=0.25, atol=1e-4, rtol=1e-4, itmax=200):
        self.grid = grid
        self.quality_threshold = float(quality_threshold)
        self.atol, self.rtol, self.itmax = float(atol), float(rtol), int(itmax)
        self.last_stats = None

    def rhs(self):
        """refresh!(solver), src/relaxation.jl:165-180 (zero off the interface)"""
        b = np.empty(self.grid.n)
        _check(self.grid.ctx, lib().lv_multiphase_rhs(self.grid.ctx, _ptr(b)))
        return b

    def apply(self, x):
        """mul!(res, MultiphaseProjector(grid), x), src/relaxation.jl:92-125"""
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        y = np.empty_like(x)
        _check(self.grid.ctx, lib().lv_multiphase_apply(self.grid.ctx, _ptr(x), _ptr(y)))
        return y


def multiphase_projection(solver):
    """multiphase_projection!(solver), src/relaxation.jl:190-206.  Returns (iterations, |r|/|b|);
    iterations = -1 when itmax stopped the solve (the reference warns and applies it anyway)."""
    it, rel = ctypes.c_int32(), ctypes.c_double()
    _check(solver.grid.ctx, lib().lv_multiphase_projection(solver.grid.ctx, solver.quality_threshold, solver.atol,
                                                           solver.rtol, solver.itmax, ctypes.byref(it),
                                                           ctypes.byref(rel)))
    solver.last_stats = dict(iters=it.value, relres=rel.value)
    if it.value < 0:
        import warnings

        warnings.warn("multiphase projector did not converge within tolerance")
    return it.value, rel.value


# ---- output (src/IO.jl:9-71) ---------------------------------------------------------------------

# the PolygonNS fields export_grid can write (src/celldefs.jl:7-39); area and centroid are functions
# of the mesh in the reference, not fields
EXPORTABLE = ("x", "rho", "v", "e", "P", "c2", "dv", "mass", "momentum", "energy", "phase", "quality", "D", "mu")


def _export_ids(vars):
    ids = []
    for v in vars:
        name = v if isinstance(v, str) else str(v)
        name = name.lstrip(":")
        if name not in EXPORTABLE:
            raise ValueError("Cannot export variable %s because it does not exist." % name)
        ids.append(FIELDS[name][0])
    return np.asarray(ids, dtype=np.int32)


def vtp_filename(filename):
    """the file vtk_grid(filename, ...) writes: .vtp appended unless already there"""
    return filename if filename.endswith(".vtp") else filename + ".vtp"


def export_grid(grid, filename, *vars, compress=True):
    """export_grid(grid, filename, vars...), src/IO.jl:9-34: the Voronoi polygons as a VTK XML
    PolyData file with the named fields as cell data.  Returns the file name."""
    ids = _export_ids(vars)
    _check(grid.ctx, lib().lv_export_grid(grid.ctx, str(filename).encode(), _ptr(ids) if ids.size else None,
                                          int(ids.size), int(bool(compress))))
    return vtp_filename(str(filename))


def export_points(grid, filename, *vars, compress=True):
    """export_points(grid, filename, vars...), src/IO.jl:50-59: the seeds as a point cloud"""
    ids = _export_ids(vars)
    _check(grid.ctx, lib().lv_export_points(grid.ctx, str(filename).encode(), _ptr(ids) if ids.size else None,
                                            int(ids.size), int(bool(compress))))
    return vtp_filename(str(filename))


# ---- interpolation (src/movingls.jl, src/voronoigrid.jl:154-178) --------------------------------

LinearExpansion, QuadraticExpansion, CubicExpansion = 1, 2, 3
_NCOEF = {1: 2, 2: 5, 3: 9}


def _field_component(name):
    """("v", 0) for "v[1]" / ("v", 0) / "P" -> ("P", 0): the scalar a fun(p) of the reference reads"""
    if isinstance(name, tuple):
        return name[0], int(name[1])
    if name.endswith("]"):
        base, k = name[:-1].split("[")
        return base, int(k) - 1  # Julia's 1-based component
    return name, 0


def _queries(x):
    q = np.asarray(x, dtype=np.float64).reshape(-1, 2)
    return q.shape[0], np.ascontiguousarray(np.concatenate([q[:, 0], q[:, 1]]))


def nearest_polygon(grid, x):
    """nearest_polygon(grid, x), src/voronoigrid.jl:154-178: 0-based indices for points x ((m, 2) or
    one point)"""
    m, q = _queries(x)
    idx = np.empty(m, dtype=np.int32)
    _check(grid.ctx, lib().lv_nearest_polygon(grid.ctx, m, _ptr(q), _ptr(idx)))
    return idx


def movingls(order, grid, field, h=None):
    """movingls(T, grid, p, p -> field; h) for every polygon, src/movingls.jl:41-96: an (n, 2|5|9)
    array of Taylor coefficients (T = LinearExpansion, QuadraticExpansion, CubicExpansion; field
    "P", "v[1]", ("v", 0), ...), with the Gaussian kernel"""
    name, comp = _field_component(field)
    out = np.empty((grid.n, _NCOEF[order]))
    _check(grid.ctx, lib().lv_movingls(grid.ctx, int(order), FIELDS[name][0], comp, 0.0 if h is None else float(h),
                                       _ptr(out)))
    return out


def point_value(grid, x, field):
    """point_value(grid, x, p -> field), src/movingls.jl:140-145, at one point or (m, 2) points"""
    name, comp = _field_component(field)
    m, q = _queries(x)
    out = np.empty(m)
    _check(grid.ctx, lib().lv_point_values(grid.ctx, m, _ptr(q), FIELDS[name][0], comp, _ptr(out)))
    return out[0] if np.ndim(x) == 1 else out


NOALLOC_MAX_QUERIES = 65536


def point_value_noalloc(grid, x, field):
    """point_value at up to NOALLOC_MAX_QUERIES points without any device allocation (the queries
    and values are staged in the library's static per-device array; same values as point_value)"""
    name, comp = _field_component(field)
    m, q = _queries(x)
    out = np.empty(m)
    _check(grid.ctx, lib().lv_point_values_noalloc(grid.ctx, m, _ptr(q), FIELDS[name][0], comp, _ptr(out)))
    return out[0] if np.ndim(x) == 1 else out


def power_vector(order, dx):
    """power_vector(T, x), src/movingls.jl:26-36"""
    a, b = dx
    p = [a, b]
    if order >= 2:
        p += [a * a, a * b, b * b]
    if order >= 3:
        p += [a * a * a, a * a * b, a * b * b, b * b * b]
    return np.array(p)


def poly_eval(val, taylor, dx):
    """poly_eval(val, taylor, dx), src/movingls.jl:98-105"""
    order = {2: 1, 5: 2, 9: 3}[len(taylor)]
    return val + float(np.dot(taylor, power_vector(order, dx)))


def integral(x, verts, val, taylor):
    """integral(p, val, taylor), src/movingls.jl:107-138: the integral over the polygon with seed x and
    chained edge vertices verts ((k, 4): v1x, v1y, v2x, v2y, as grid.mesh() gives them) of the
    polynomial val + taylor . power_vector(dx)"""
    x = np.asarray(x, dtype=np.float64)
    order = {2: 1, 5: 2, 9: 3}[len(taylor)]
    out = 0.0
    for e in np.asarray(verts, dtype=np.float64):
        v1, v2 = e[:2], e[2:]
        d1, d2 = v1 - x, v2 - x
        A = 0.5 * abs(d1[0] * d2[1] - d1[1] * d2[0])
        if order == 1:
            out += A * poly_eval(val, taylor, -2 / 3 * x + 1 / 3 * v1 + 1 / 3 * v2)
        elif order == 2:
            out += (A / 3) * poly_eval(val, taylor, -1 / 3 * x + 1 / 6 * v1 + 1 / 6 * v2)
            out += (A / 3) * poly_eval(val, taylor, -5 / 6 * x + 2 / 3 * v1 + 1 / 6 * v2)
            out += (A / 3) * poly_eval(val, taylor, -5 / 6 * x + 1 / 6 * v1 + 2 / 3 * v2)
        else:
            out -= (9 * A / 16) * poly_eval(val, taylor, -2 / 3 * x + 1 / 3 * v1 + 1 / 3 * v2)
            out += (25 * A / 48) * poly_eval(val, taylor, -2 / 5 * x + 1 / 5 * v1 + 1 / 5 * v2)
            out += (25 * A / 48) * poly_eval(val, taylor, -4 / 5 * x + 3 / 5 * v1 + 1 / 5 * v2)
            out += (25 * A / 48) * poly_eval(val, taylor, -4 / 5 * x + 1 / 5 * v1 + 3 / 5 * v2)
    return out


from . import spai  # noqa: E402,F401  (learned preconditioner helpers)
from .populate import (populate_circ, populate_hex, populate_lloyd, populate_rand, populate_rect,  # noqa: E402,F401
                       populate_vogel)
from .simulation import SimulationWorkspace, movingavg, run  # noqa: E402,F401
