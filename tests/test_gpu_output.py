"""export_grid / export_points / run! on the GPU (src/IO.jl:9-71, src/simulation.jl:35-94).

The polygons written to the .vtp file are the mesh itself: every polygon's vertices (the v1 of its
chained edges) equal the oracle's mesh bit for bit, the offsets are the running edge counts, and
the cell data are the device fields bit for bit (vectors as (x, y, 0)).  run() of the reference's
Taylor-Green test writes the files that test checks (tests/taylorgreen.jl:115-119)."""
import os

import numpy as np
import pytest

import flows
from pairs import make_pair
from vtp import read_vtp

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,compress", [("periodic", True), ("walls", False)])
def test_export_grid_is_the_mesh(lv, oracle, gpu, tmp_path, kind, compress):
    print("export_grid %s" % kind, flush=True)
    grid, op, x, dr = make_pair(lv, oracle, kind, K=40, seed=0x5EED0700)
    f = lv.export_grid(grid, str(tmp_path / "cframe0"), "P", "v", "D", "phase", compress=compress)
    assert f.endswith("cframe0.vtp") and os.path.exists(f)
    d = read_vtp(f)
    assert d["compressed"] == compress
    ptr, lab, ev = op.mesh()
    n = grid.n
    assert int(d["piece"]["NumberOfPolys"]) == n and int(d["piece"]["NumberOfPoints"]) == ptr[-1]
    assert np.array_equal(d["cells"]["Polys"]["offsets"], ptr[1:].astype(np.int64))
    assert np.array_equal(d["cells"]["Polys"]["connectivity"], np.arange(ptr[-1], dtype=np.int64))
    assert np.array_equal(d["points"][:, 0], ev[:, 0]) and np.array_equal(d["points"][:, 1], ev[:, 1])
    assert not np.any(d["points"][:, 2])
    assert np.array_equal(d["celldata"]["P"], grid.get("P"))
    v = grid.get_soa("v")
    assert np.array_equal(d["celldata"]["v"], np.stack([v[:n], v[n:], np.zeros(n)], axis=1))
    D = grid.get_soa("D")
    assert np.array_equal(d["celldata"]["D"], np.stack([D[:n], D[n:2 * n], np.zeros(n)], axis=1))
    assert np.array_equal(d["celldata"]["phase"], grid.get("phase").astype(np.float64))


def test_export_points_and_unknown_field(lv, oracle, gpu, tmp_path):
    print("export_points", flush=True)
    grid, op, x, dr = make_pair(lv, oracle, "periodic", K=24, seed=0x5EED0701)
    f = lv.export_points(grid, str(tmp_path / "pframe0.vtp"), "rho", "momentum")
    d = read_vtp(f)
    n = grid.n
    assert np.array_equal(d["points"], np.stack([x[:n], x[n:], np.zeros(n)], axis=1))
    assert np.array_equal(d["cells"]["Verts"]["offsets"], np.arange(1, n + 1))
    assert np.array_equal(d["celldata"]["rho"], grid.get("rho"))
    with pytest.raises(ValueError, match="Cannot export variable area"):
        lv.export_grid(grid, str(tmp_path / "bad"), "area")


def test_run_writes_the_reference_test_files(lv, gpu, tmp_path):
    """run! of tests/taylorgreen.jl:101-119 (nframes = 20, save_csv/points/grid, vtp_vars P and v)
    on the N = 80 Taylor-Green case: the files the test asserts exist, 20 frames in each
    collection, and the csv's error columns within the test's thresholds."""
    print("run taylor-green", flush=True)

    class TG(lv.SimulationWorkspace):
        def __init__(self):
            x = flows.populate_hex((0.0, 0.0), (1.0, 1.0), flows.TG_DR)
            self.grid = lv.VoronoiGrid(lv.UnitRectangle(), flows.TG_DR, xperiodic=True, yperiodic=True, positions=x)
            lv.remesh(self.grid)
            for k, v in flows.tg_ic(x, self.grid.get("area")).items():
                self.grid.set_soa(k, v)
            solver = lv.PressureSolver(self.grid, method="amg", tol=1e-10, picard_window=8)
            self.desc = lv.make_step_desc(flows.TG_DT, eos="stiffened", gamma=flows.TG_GAMMA, P0=flows.TG_P0,
                                          solver=solver, viscous=True, artificial_viscosity=False)
            self.E0 = None
            self.E_err = self.v_err = self.P_err = 0.0

    def step(sim, t):
        lv.step_fused(sim.grid, sim.desc)

    def postproc(sim, t):
        g = sim.grid
        E, sim.v_err, sim.P_err = flows.tg_errors(g.get_soa("x"), g.get_soa("v"), g.get("P"), g.get("area"),
                                                  g.get("mass"), g.get("e"), t + flows.TG_DT)
        sim.E0 = E if sim.E0 is None else sim.E0
        sim.E_err = E - sim.E0

    sim = TG()
    path = str(tmp_path / "results")
    lv.run(sim, flows.TG_DT, flows.TG_TEND, step, postproc=postproc, nframes=20, path=path, save_csv=True,
           save_points=True, save_grid=True, vtp_vars=("P", "v"), csv_vars=("E_err", "v_err", "P_err"))
    for f in ("simdata.csv", "cells.pvd", "cframe19.vtp", "points.pvd", "pframe19.vtp"):
        assert os.path.exists(os.path.join(path, f)), f
    import xml.etree.ElementTree as ET

    for c in ("cells.pvd", "points.pvd"):
        assert len(ET.parse(os.path.join(path, c)).getroot().findall("Collection/DataSet")) == 20
    rows = open(os.path.join(path, "simdata.csv")).read().splitlines()
    assert rows[0] == "E_err,P_err,time,v_err" and len(rows) == 21
    E_err, P_err, t, v_err = map(float, rows[-1].split(","))
    assert E_err < 1e-8 and v_err < 0.01 and P_err < 0.01
    d = read_vtp(os.path.join(path, "cframe19.vtp"))
    assert np.array_equal(d["celldata"]["P"], sim.grid.get("P"))
