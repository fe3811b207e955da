"""The closed-form pins of tests/test_oracle_pins_physics.py run on the GPU library.

Each pin there takes an `oracle` module and drives its Grid/Particles; here the same test functions
drive the HIP library through the C ABI instead (an adapter with the oracle's method names), so the
GPU kernels k_find_D, k_visc1/2, k_friction, k_find_dv and k_eos meet the same closed forms and
Qhull identities directly, and every field they write is also compared with the oracle at 1e-12.
"""
import types

import numpy as np
import pytest

import test_oracle_pins_physics as pins
from pairs import rel_err

pytestmark = pytest.mark.gpu

_NAMES = {"mom": "momentum"}


def _adapter(lv, oracle):
    """A module with Grid and Particles classes whose operators run on the GPU; every operator is
    mirrored on an oracle twin and the written fields compared at 1e-12."""

    class Grid:
        def __init__(self, xmin, xmax, dr, h=None, r_max=None, xperiodic=False, yperiodic=False):
            self.args = (xmin, xmax, dr, h, r_max, xperiodic, yperiodic)
            self.twin = oracle.Grid(xmin, xmax, dr, h=h, r_max=r_max, xperiodic=xperiodic, yperiodic=yperiodic)

    class Particles:
        def __init__(self, grid, x):
            xmin, xmax, dr, h, r_max, xp, yp = grid.args
            self.g = lv.VoronoiGrid(lv.Rectangle(xmin, xmax), dr, h=h, r_max=r_max, xperiodic=xp, yperiodic=yp,
                                    positions=x)
            self.twin = oracle.Particles(grid.twin, x)
            self.n = x.size // 2

        def set(self, k, val):
            fid, comps = lv.FIELDS[_NAMES.get(k, k)]
            a = np.asarray(val, dtype=np.int32 if k == "phase" else np.float64)
            if a.ndim == 0:
                a = np.full(comps * self.n, a)
            self.g.set_soa(_NAMES.get(k, k), a)
            self.twin.set(k, a)

        def __getitem__(self, k):
            return self.g.get_soa(_NAMES.get(k, k))

        def _same(self, *names):
            for k in names:
                assert rel_err(self[k], self.twin[k]) <= 1e-12, k

        def remesh(self):
            lv.remesh(self.g)
            self.twin.remesh()
            pg, lg, vg = self.g.mesh()
            po, lo, vo = self.twin.mesh()
            assert np.array_equal(pg, po) and np.array_equal(lg, lo) and np.array_equal(vg, vo)

        def mesh(self):
            return self.g.mesh()

        def area(self):
            return self.g.get("area")

        def find_D(self):
            lv.find_D(self.g)
            self.twin.find_D()
            self._same("D")

        def viscous_step(self, dt, av=True):
            lv.viscous_step(self.g, dt, artificial_viscosity=av)
            self.twin.viscous_step(dt, av)
            self._same("v", "e")

        def bdary_friction(self, vbc, dt):
            lv.bdary_friction(self.g, np.asarray(vbc, dtype=np.float64), dt)
            self.twin.bdary_friction(vbc, dt)
            self._same("v", "e")

        def find_dv(self, dt, alpha=1.0):
            lv.find_dv(self.g, dt, alpha)
            self.twin.find_dv(dt, alpha)
            self._same("dv", "quality")

        def eos_ideal(self, gamma, Pmin):
            lv.ideal_eos(self.g, gamma, Pmin)
            self.twin.eos_ideal(gamma, Pmin)
            self._same("rho", "P", "c2")

        def eos_stiffened(self, gamma, P0):
            lv.stiffened_eos(self.g, gamma, P0)
            self.twin.eos_stiffened(gamma, P0)
            self._same("rho", "P", "c2")

    return types.SimpleNamespace(Grid=Grid, Particles=Particles)


PINS = [
    ("test_find_D_linear_velocity_on_hexagons", {}),
    ("test_find_D_linear_velocity_on_jittered_cells", {}),
    ("test_viscous_pass1_linear_stress", dict(D0=pins.D0_COMPRESSIVE, av=True, stone_norman=True)),
    ("test_viscous_pass1_linear_stress", dict(D0=pins.D0_EXPANSIVE, av=True, stone_norman=False)),
    ("test_viscous_pass1_linear_stress", dict(D0=pins.D0_COMPRESSIVE, av=False, stone_norman=False)),
    ("test_viscous_pass1_deviatoric_part", {}),
    ("test_viscous_pass2_energy_closed_form", {}),
    ("test_friction_wall_rows_of_a_hexagonal_lattice", {}),
    ("test_friction_on_jittered_walls_brute_force", dict(skip=())),
    ("test_friction_on_jittered_walls_brute_force", dict(skip=(-2, -4))),
    ("test_find_dv_zero_on_regular_hexagons", {}),
    ("test_find_dv_from_qhull_geometry", {}),
    ("test_eos_closed_forms_with_pmin", {}),
]


@pytest.mark.parametrize("name,kw", PINS, ids=["%s-%d" % (p[0][5:], k) for k, p in enumerate(PINS)])
def test_gpu_meets_closed_form(lv, oracle, gpu, name, kw):
    print("closed-form pin on the GPU: %s %s" % (name, sorted(kw)), flush=True)
    fn = getattr(pins, name)
    gpu_oracle = _adapter(lv, oracle)
    params = fn.__code__.co_varnames[:fn.__code__.co_argcount]
    args = dict(kw)
    if "lv" in params:
        args["lv"] = lv
    args["oracle"] = gpu_oracle
    fn(**args)


def test_friction_charfun_is_the_per_wall_table(lv, oracle, gpu):
    """bdary_friction(grid, vD, dt; charfun) with the bubble example's top_and_bottom (examples/bubble.jl:
    81-93) equals the table with NaN on the side walls, bit for bit, and leaves cells that touch only
    the side walls unchanged."""
    from pairs import lattice

    x, dr = lattice(lv, "walls", 24, 0x5EED0704)
    out = []
    for mode in ("charfun", "table"):
        g = lv.VoronoiGrid(lv.UnitRectangle(), dr, positions=x)
        lv.remesh(g)
        n = g.n
        rng = np.random.default_rng(9)
        g.set_soa("v", rng.standard_normal(2 * n))
        g.set_soa("e", 1.0 + rng.random(n))
        g.set_soa("mu", np.full(n, 0.03))
        g.set_soa("mass", g.get("area"))
        v0 = g.get_soa("v")
        vD = {lv.BDARY_UP: (1.0, 0.0), lv.BDARY_DOWN: (0.0, 0.0), lv.BDARY_LEFT: (0.0, 0.5), lv.BDARY_RIGHT: (0.0, -0.5)}
        if mode == "charfun":
            lv.bdary_friction(g, vD, 1e-3, charfun=lambda m: abs(m[1] - 0.0) < 1e-12 or abs(m[1] - 1.0) < 1e-12)
        else:
            t = lv._vbc_table(vD)
            t[2:4] = np.nan  # right
            t[6:8] = np.nan  # left
            lv.bdary_friction(g, t, 1e-3)
        out.append((g.get_soa("v"), g.get_soa("e")))
        if mode == "table":
            ptr, lab, _ = g.mesh()
            side_only = [i for i in range(n) if any(lab[ptr[i]:ptr[i + 1]] < 0)
                         and all(l in (-2, -4) for l in lab[ptr[i]:ptr[i + 1]] if l < 0)]
            assert side_only
            for i in side_only:
                assert out[-1][0][i] == v0[i] and out[-1][0][n + i] == v0[n + i]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
