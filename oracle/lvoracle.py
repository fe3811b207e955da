"""ctypes binding of the CPU FP64 oracle (oracle/lvoracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's
cpu_baseline leg, always as the checker or the timed CPU baseline, never as the product.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "lvoracle.c")
LIB = os.path.join(HERE, "liblvoracle.so")
MAXE = 32

LV_OK, LV_E_MESH_DESTROYED, LV_E_NAN_VELOCITY, LV_E_POLY_OVERFLOW, LV_E_NOT_CONVERGED = 0, 1, 2, 3, 4


def build():
    """gcc -O2 -ffp-contract=off -fopenmp: the oracle must round exactly as the reference's
    uncontracted IEEE FP64 arithmetic does."""
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared", "-o", LIB + ".tmp",
           SRC, "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


class Desc(ctypes.Structure):
    _fields_ = [("bxmin", ctypes.c_double * 2), ("bxmax", ctypes.c_double * 2), ("dr", ctypes.c_double),
                ("h", ctypes.c_double), ("r_max", ctypes.c_double), ("xperiodic", ctypes.c_int),
                ("yperiodic", ctypes.c_int)]


class State(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("maxe", ctypes.c_int)] + \
               [(k, ctypes.c_void_p) for k in ("x", "v", "dv", "mom", "rho", "e", "P", "c2", "mass", "energy",
                                               "quality", "mu", "D", "phase", "ne", "ev", "el")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        L = ctypes.CDLL(LIB)
        P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        sig = {
            "lvo_grid_new": ([P], P), "lvo_grid_free": ([P], None), "lvo_grid_info": ([P, P, P, P, P], I),
            "lvo_grid_path": ([P, P, P, P], I), "lvo_cell_list": ([P, I, P, P, P, ctypes.c_longlong], ctypes.c_longlong),
            "lvo_remesh": ([P, P], I), "lvo_area": ([P, P], I), "lvo_centroid": ([P, P], I),
            "lvo_find_rho": ([P], I), "lvo_eos_stiffened": ([P, D, D], I), "lvo_eos_ideal": ([P, D, D], I),
            "lvo_gravity": ([P, D, D, D], I), "lvo_move_positions": ([P, P, D], I),
            "lvo_pressure_operator": ([P, P, D, P, P], I), "lvo_apply_A": ([P, P, P, P, P], I),
            "lvo_rhs": ([P, P, D, I, P, P, P, P], I), "lvo_pcg": ([P, P, P, P, P, D, I, P], I),
            "lvo_find_pressure": ([P, P, D, I, P, D, I, P], I), "lvo_pressure_step": ([P, P, D], I),
            "lvo_find_pressure_m": ([P, P, D, I, P, D, I, P, I], I), "lvo_minres": ([P, P, P, P, P, D, D, I, P], I),
            "lvo_find_D": ([P, P], I), "lvo_viscous_step": ([P, P, D, I], I), "lvo_bdary_friction": ([P, P, D], I),
            "lvo_find_dv": ([P, P, D, D], I), "lvo_relaxation_step": ([P, P, D, I], I),
            "lvo_multiphase_rhs": ([P, P, P], I), "lvo_multiphase_apply": ([P, P, P, P, P], I),
            "lvo_multiphase_update": ([P, P, P, D], I), "lvo_multiphase_projection": ([P, P, D, D, D, I, P], I),
            "lvo_viscous_operator": ([P, P, D, P, P], I), "lvo_viscous_implicit": ([P, P, D, D, I], I),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__({1: "The Voronoi Mesh has been destroyed.", 2: "Velocity field invalidated.",
                          3: "polygon capacity exceeded", 4: "not converged"}.get(code, "error %d" % code))
        self.code = code


def _chk(rc):
    if rc != 0:
        raise OracleError(rc)


class Grid:
    """VoronoiGrid geometry + CellList + magic path, restated (src/voronoigrid.jl:26-49)."""

    def __init__(self, xmin, xmax, dr, h=None, r_max=None, xperiodic=False, yperiodic=False):
        d = Desc()
        d.bxmin[:] = xmin
        d.bxmax[:] = xmax
        d.dr = dr
        d.h = 2.0 * dr if h is None else h
        d.r_max = 10.0 * dr if r_max is None else r_max
        d.xperiodic, d.yperiodic = int(xperiodic), int(yperiodic)
        self.desc = d
        self.ptr = lib().lvo_grid_new(ctypes.byref(d))
        n1, n2, npath = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        out6 = np.empty(6)
        lib().lvo_grid_info(self.ptr, ctypes.byref(n1), ctypes.byref(n2), ctypes.byref(npath), _p(out6))
        self.n1, self.n2, self.npath = n1.value, n2.value, npath.value
        self.cropping_min, self.cropping_max = tuple(out6[:2]), tuple(out6[2:4])
        self.next_rr, self.rr_max = out6[4], out6[5]

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.lvo_grid_free(self.ptr)
            self.ptr = None

    def magic_path(self):
        i1 = np.empty(self.npath, dtype=np.int32)
        i2 = np.empty(self.npath, dtype=np.int32)
        rr = np.empty(self.npath)
        lib().lvo_grid_path(self.ptr, _p(i1), _p(i2), _p(rr))
        return i1, i2, rr

    def cell_list(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = x.size // 2
        cs = np.empty(self.n1 * self.n2 + 1, dtype=np.int32)
        cap = 9 * n + 16
        cl = np.empty(cap, dtype=np.int32)
        m = lib().lvo_cell_list(self.ptr, n, _p(x), _p(cs), _p(cl), cap)
        return cs, cl[:m].copy()


FIELDS1 = ("rho", "e", "P", "c2", "mass", "energy", "quality", "mu")
FIELDS2 = ("x", "v", "dv", "mom")


class Particles:
    """SoA state of n polygons (PolygonNS fields) plus their edges, owned by numpy."""

    def __init__(self, grid, x):
        self.grid = grid
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = x.size // 2
        self.n = n
        self.a = {k: np.zeros(2 * n) for k in FIELDS2}
        self.a["x"][:] = x
        for k in FIELDS1:
            self.a[k] = np.zeros(n)
        self.a["D"] = np.zeros(4 * n)
        self.a["phase"] = np.zeros(n, dtype=np.int32)
        self.a["ne"] = np.zeros(n, dtype=np.int32)
        self.a["ev"] = np.zeros(n * MAXE * 4)
        self.a["el"] = np.zeros(n * MAXE, dtype=np.int32)
        s = State()
        s.n, s.maxe = n, MAXE
        for k in FIELDS2 + FIELDS1 + ("D", "phase", "ne", "ev", "el"):
            setattr(s, k, _p(self.a[k]))
        self.s = s

    def __getitem__(self, k):
        return self.a[k]

    def set(self, k, val):
        self.a[k][:] = val

    # ---- mesh ----
    def remesh(self):
        _chk(lib().lvo_remesh(self.grid.ptr, ctypes.byref(self.s)))

    def mesh(self):
        """(edge_ptr, labels, verts[E,4]) in the same CSR form as the product's lv_get_mesh"""
        ne = self.a["ne"]
        ptr = np.zeros(self.n + 1, dtype=np.int32)
        ptr[1:] = np.cumsum(ne)
        el = self.a["el"].reshape(self.n, MAXE)
        ev = self.a["ev"].reshape(self.n, MAXE, 4)
        mask = np.arange(MAXE)[None, :] < ne[:, None]
        return ptr, el[mask].copy(), ev[mask].copy()

    def area(self):
        out = np.empty(self.n)
        lib().lvo_area(ctypes.byref(self.s), _p(out))
        return out

    def centroid(self):
        out = np.empty(2 * self.n)
        lib().lvo_centroid(ctypes.byref(self.s), _p(out))
        return out

    # ---- operators ----
    def find_rho(self):
        lib().lvo_find_rho(ctypes.byref(self.s))

    def eos_stiffened(self, gamma, P0):
        lib().lvo_eos_stiffened(ctypes.byref(self.s), gamma, P0)

    def eos_ideal(self, gamma, Pmin):
        lib().lvo_eos_ideal(ctypes.byref(self.s), gamma, Pmin)

    def gravity(self, g, dt):
        lib().lvo_gravity(ctypes.byref(self.s), g[0], g[1], dt)

    def move(self, dt, remesh=True):
        _chk(lib().lvo_move_positions(self.grid.ptr, ctypes.byref(self.s), dt))
        if remesh:
            self.remesh()

    def pressure_operator(self, dt):
        diag = np.empty(self.n)
        coef = np.empty(self.n * MAXE)
        lib().lvo_pressure_operator(self.grid.ptr, ctypes.byref(self.s), dt, _p(diag), _p(coef))
        return diag, coef

    def apply_A(self, diag, coef, x):
        y = np.empty(self.n)
        lib().lvo_apply_A(ctypes.byref(self.s), _p(diag), _p(coef), _p(np.ascontiguousarray(x)), _p(y))
        return y

    def csr(self, diag, coef):
        """The operator as (diag, row_ptr, cols, coef) in the product's CSR edge order"""
        ne = self.a["ne"]
        el = self.a["el"].reshape(self.n, MAXE)
        cf = coef.reshape(self.n, MAXE)
        mask = (np.arange(MAXE)[None, :] < ne[:, None]) & (el >= 0)
        rp = np.zeros(self.n + 1, dtype=np.int32)
        rp[1:] = np.cumsum(mask.sum(axis=1))
        return diag, rp, el[mask].copy(), cf[mask].copy()

    def rhs(self, dt, gp_step, vbc=None):
        b = np.empty(self.n)
        P0 = np.empty(self.n)
        GP = np.empty(2 * self.n)
        vb = None if vbc is None else np.ascontiguousarray(vbc, dtype=np.float64)
        lib().lvo_rhs(self.grid.ptr, ctypes.byref(self.s), dt, int(gp_step), _p(vb), _p(b), _p(P0), _p(GP))
        return b, P0, GP

    def pcg(self, diag, coef, b, x0, tol, itmax=100000):
        x = np.ascontiguousarray(np.array(x0, dtype=np.float64))
        rr = ctypes.c_double()
        it = lib().lvo_pcg(ctypes.byref(self.s), _p(diag), _p(coef), _p(np.ascontiguousarray(b)), _p(x), tol, itmax,
                           ctypes.byref(rr))
        return x, it, rr.value

    def find_pressure(self, dt, niter=10, vbc=None, tol=1e-10, itmax=100000, method="pcg"):
        """method "pcg": Jacobi-PCG to tol (relative to |b|); "minres": MINRES with atol = rtol = tol, as
        the reference calls Krylov.jl's minres! (src/pressure.jl:219); a missed tolerance is returned
        as -1 iterations, not raised, since the reference ignores it too"""
        iters = np.zeros(niter, dtype=np.int32)
        vb = None if vbc is None else np.ascontiguousarray(vbc, dtype=np.float64)
        m = {"pcg": 0, "minres": 1}[method]
        rc = lib().lvo_find_pressure_m(self.grid.ptr, ctypes.byref(self.s), dt, niter, _p(vb), tol, itmax, _p(iters),
                                       m)
        if m == 0:
            _chk(rc)
        return iters

    def minres(self, diag, coef, b, x0, atol, rtol, itmax=1000):
        x = np.ascontiguousarray(np.array(x0, dtype=np.float64))
        rr = ctypes.c_double()
        it = lib().lvo_minres(ctypes.byref(self.s), _p(diag), _p(coef), _p(np.ascontiguousarray(b)), _p(x), atol,
                              rtol, itmax, ctypes.byref(rr))
        return x, it, rr.value

    def pressure_step(self, dt):
        lib().lvo_pressure_step(self.grid.ptr, ctypes.byref(self.s), dt)

    def find_D(self):
        lib().lvo_find_D(self.grid.ptr, ctypes.byref(self.s))

    def viscous_step(self, dt, artificial_viscosity=True):
        lib().lvo_viscous_step(self.grid.ptr, ctypes.byref(self.s), dt, int(artificial_viscosity))

    def viscous_operator(self, dt):
        """ViscousSolver's operator (diag, coef) in the pressure operator's layout"""
        diag = np.empty(self.n)
        coef = np.empty(self.n * MAXE)
        lib().lvo_viscous_operator(self.grid.ptr, ctypes.byref(self.s), dt, _p(diag), _p(coef))
        return diag, coef

    def viscous_implicit(self, dt, tol=1e-6, itmax=2000):
        return lib().lvo_viscous_implicit(self.grid.ptr, ctypes.byref(self.s), dt, tol, itmax)

    def bdary_friction(self, vbc, dt):
        lib().lvo_bdary_friction(ctypes.byref(self.s), _p(np.ascontiguousarray(vbc, dtype=np.float64)), dt)

    def find_dv(self, dt, alpha=1.0):
        lib().lvo_find_dv(self.grid.ptr, ctypes.byref(self.s), dt, alpha)

    def multiphase_rhs(self):
        b = np.empty(self.n)
        lib().lvo_multiphase_rhs(self.grid.ptr, ctypes.byref(self.s), _p(b))
        return b

    def multiphase_apply(self, x):
        T = np.empty(2 * self.n)
        y = np.empty(self.n)
        lib().lvo_multiphase_apply(self.grid.ptr, ctypes.byref(self.s), _p(np.ascontiguousarray(x, dtype=np.float64)),
                                   _p(T), _p(y))
        return y

    def multiphase_update(self, res, quality_threshold=0.25):
        lib().lvo_multiphase_update(self.grid.ptr, ctypes.byref(self.s),
                                    _p(np.ascontiguousarray(res, dtype=np.float64)), quality_threshold)

    def multiphase_projection(self, quality_threshold=0.25, atol=1e-4, rtol=1e-4, itmax=200):
        res = np.empty(self.n)
        it = lib().lvo_multiphase_projection(self.grid.ptr, ctypes.byref(self.s), quality_threshold, atol, rtol,
                                             itmax, _p(res))
        return it, res

    def relaxation_step(self, dt, rusanov=True, remesh=True):
        lib().lvo_relaxation_step(self.grid.ptr, ctypes.byref(self.s), dt, int(rusanov))
        if remesh:
            self.remesh()

    def step(self, dt, eos=None, gamma=1.4, P0=0.0, Pmin=0.0, gravity=None, niter=10, tol=1e-10, pvbc=None,
             viscous=True, artificial_viscosity=False, friction=None, alpha=1.0, rusanov=True, multiphase=None,
             multiphase_tol=1e-4, multiphase_itmax=200, viscous_tol=1e-6, itmax=100000, solver="pcg"):
        """step!(sim, t) in the order of tests/taylorgreen.jl:61-71 (optional parts as in
        examples/cavity.jl:70-80 and examples/rayleightaylor.jl:90-101; multiphase = the
        MultiphaseSolver's quality_threshold, or None for no multiphase_projection!)"""
        self.move(dt)
        if gravity is not None:
            self.gravity(gravity, dt)
        if eos == "stiffened":
            self.eos_stiffened(gamma, P0)
        elif eos == "ideal":
            self.eos_ideal(gamma, Pmin)
        elif eos == "rho":
            self.find_rho()
        iters = self.find_pressure(dt, niter, pvbc, tol, itmax, solver)
        self.pressure_step(dt)
        self.find_D()
        if viscous == "implicit":
            self.viscous_implicit(dt, viscous_tol, 100000)
        elif viscous:
            self.viscous_step(dt, artificial_viscosity)
        if friction is not None:
            self.bdary_friction(friction, dt)
        self.find_dv(dt, alpha)
        if multiphase is not None:
            self.multiphase_projection(multiphase, multiphase_tol, multiphase_tol, multiphase_itmax)
        self.relaxation_step(dt, rusanov)
        return iters
