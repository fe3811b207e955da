"""The multilevel preconditioner on the GPU (csrc/lv_amg.hip) against its FP64 oracle
(oracle/amg_ref.py), and the solver built on it against the pressure oracle.

Tolerances: the cycle runs in FP32, so z = B r is compared in relative L2 norm at 1e-4 (Jacobi
fine smoother, everything else fixed by the construction) and 2e-3 with the learned smoother (its
M is BF16-network output in FP32 and its scale a 10-step power-iteration estimate).  By default
the device renumbers the rows cell by cell and smooths with BF16 couplings; the oracle models that
rounding (fine_bf16), and the operator's own order (amg_layout = 1) is checked against the oracle
without it.  Solves meet
the FP64 residual tolerance, recomputed on the host.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from pairs import make_pair, rel_err

pytestmark = pytest.mark.gpu

amg_ref = pytest.importorskip("amg_ref")
spai_ref = pytest.importorskip("spai_ref")


def _flush(msg):
    print(msg, flush=True)


def _system(lv, oracle, kind, K, ic, dt, seed=0x5EED0500):
    grid, op, x, dr = make_pair(lv, oracle, kind, K=K, seed=seed, ic=ic)
    if ic == "tg":
        lv.stiffened_eos(grid, 1.4, 2500.0 / 1.4)
        op.eos_stiffened(1.4, 2500.0 / 1.4)
    else:
        inf = np.full(grid.n, np.inf)
        grid.set("c2", inf)
        op.set("c2", inf)
    solver = lv.PressureSolver(grid, method="amg", tol=1e-8, itmax=4000)
    solver.assemble(dt)
    d, rp, cols, coef = solver.operator()
    return grid, op, solver, x, dr, d, rp, cols, coef


def _hier(kind, x, dr, d, rp, cols, coef, M=None, fine_bf16=True):
    from pairs import domain

    xmin, xmax, xp, yp = domain(kind)
    return amg_ref.Hierarchy(x, d, rp, cols, coef, xmin, xmax, dr, xp, yp, not np.any(d > 0), M=M,
                             fine_bf16=fine_bf16)


@pytest.mark.parametrize("kind,ic,K,dt", [("periodic", "tg", 64, 8e-3), ("periodic", "tg", 301, 2e-3),
                                          ("walls", "cavity", 96, 1e-3), ("tall", "cavity", 130, 1e-3)])
def test_cycle_matches_oracle_jacobi_smoother(lv, oracle, gpu, kind, ic, K, dt):
    """z = B r on the GPU equals the oracle's cycle (every level in the single-workgroup kernel at
    64^2, five coarse levels at 301^2, singular walled systems, a non-square domain)."""
    _flush("amg cycle %s %d" % (kind, K))
    grid, op, solver, x, dr, d, rp, cols, coef = _system(lv, oracle, kind, K, ic, dt)
    H = _hier(kind, x, dr, d, rp, cols, coef)
    r = np.random.default_rng(11).standard_normal(grid.n)
    z = solver.apply_preconditioner(r, "amg")
    zr = H.apply(r)
    assert np.linalg.norm(z - zr) <= 1e-4 * np.linalg.norm(zr), (np.linalg.norm(z - zr) / np.linalg.norm(zr))
    _closer_to_own_model(z, zr, _hier(kind, x, dr, d, rp, cols, coef, fine_bf16=False).apply(r))
    if K == 301:
        assert len(H.levels) >= 3


@pytest.mark.parametrize("kind,ic,K,dt", [("periodic", "tg", 301, 2e-3), ("walls", "cavity", 96, 1e-3)])
def test_cycle_matches_oracle_in_particle_order(lv, oracle, gpu, kind, ic, K, dt):
    """amg_layout = 1 keeps the operator's own row order and FP32 couplings in the cycle: z = B r
    equals the oracle's cycle without the BF16 rounding of the fine couplings."""
    _flush("amg cycle particle order %s %d" % (kind, K))
    grid, op, solver, x, dr, d, rp, cols, coef = _system(lv, oracle, kind, K, ic, dt)
    b = np.random.default_rng(4).standard_normal(grid.n)
    if not np.any(d > 0):
        b -= b.mean()
    solver.solve(b, np.zeros(grid.n), method="amg", amg_layout=1)  # builds the hierarchy in that layout
    assert max(solver.last_stats["relres"]) <= 1e-8
    H = _hier(kind, x, dr, d, rp, cols, coef, fine_bf16=False)
    r = np.random.default_rng(11).standard_normal(grid.n)
    z = solver.apply_preconditioner(r, "amg")
    zr = H.apply(r)
    assert np.linalg.norm(z - zr) <= 1e-4 * np.linalg.norm(zr), (np.linalg.norm(z - zr) / np.linalg.norm(zr))
    _closer_to_own_model(z, zr, _hier(kind, x, dr, d, rp, cols, coef).apply(r))


def _closer_to_own_model(z, zown, zother):
    """the device's z sits closer to the oracle model of its own layout than to the other layout's
    (the BF16 rounding of the fine couplings moves z by ~1e-5 relative)"""
    e_own, e_other = np.linalg.norm(z - zown), np.linalg.norm(z - zother)
    print("cycle error: own layout %.3e, other layout %.3e" % (e_own / np.linalg.norm(zown),
                                                               e_other / np.linalg.norm(zown)))
    assert e_own < 0.5 * e_other, (e_own, e_other)


@pytest.mark.parametrize("kind,ic,K,dt", [("periodic", "tg", 128, 4e-3), ("walls", "cavity", 96, 1e-3)])
def test_layouts_give_the_same_pressure(lv, oracle, gpu, kind, ic, K, dt):
    """The aggregate order (16-bit offsets, BF16 cycle) and the operator's own order reach the same
    tolerance with iteration counts within 20%, and their pressures agree to the solve's accuracy."""
    _flush("amg layouts %s %d" % (kind, K))
    grid, op, solver, x, dr, d, rp, cols, coef = _system(lv, oracle, kind, K, ic, dt)
    b, _ = solver.rhs(dt, False)
    if not np.any(d > 0):
        b = np.random.default_rng(3).standard_normal(grid.n)
        b -= b.mean()
    x0 = grid.get("P")
    P0 = solver.solve(b, x0, method="amg", amg_layout=0)
    s0 = solver.last_stats
    P1 = solver.solve(b, x0, method="amg", amg_layout=1)
    s1 = solver.last_stats
    assert max(s0["relres"]) <= 1e-8 and max(s1["relres"]) <= 1e-8
    assert abs(s0["iters"][0] - s1["iters"][0]) <= max(2, 0.2 * s1["iters"][0]), (s0["iters"], s1["iters"])
    if not np.any(d > 0):
        P1 = P1 - P1.mean() + P0.mean()
    assert rel_err(P0, P1) <= (1e-4 if not np.any(d > 0) else 1e-5)


@pytest.mark.parametrize("kind,ic,K,dt", [("periodic", "tg", 256, 2e-3), ("walls", "cavity", 256, 1e-3)])
def test_compact_overflow_is_bit_identical(lv, oracle, gpu, kind, ic, K, dt):
    """The aggregate order with the compact copy of the slots past the sixth (layout 0), without it
    (2), and with it only for groups of at most 16 such slots (3, so that both paths run in one
    solve) give the same pressure bit for bit and the same iterations."""
    _flush("amg compact overflow %s %d" % (kind, K))
    grid, op, solver, x, dr, d, rp, cols, coef = _system(lv, oracle, kind, K, ic, dt)
    b, _ = solver.rhs(dt, False)
    if not np.any(d > 0):
        b = np.random.default_rng(5).standard_normal(grid.n)
        b -= b.mean()
    deg = np.diff(rp)
    assert deg.max() >= 8 and deg.min() <= 5  # rows past the sixth slot, and padding below it
    x0 = grid.get("P")
    out = {}
    for layout in (0, 2, 3):
        P = solver.solve(b, x0, method="amg", amg_layout=layout)
        st = solver.last_stats
        assert st["agg_order"] == 1 and max(st["relres"]) <= 1e-8, st
        out[layout] = (P, st["iters"][:st["nsolves"]], st["outer"][:st["nsolves"]])
    for layout in (2, 3):
        assert np.array_equal(out[0][0], out[layout][0]), layout
        assert list(out[0][1]) == list(out[layout][1]) and list(out[0][2]) == list(out[layout][2])


@pytest.mark.parametrize("kind,ic,K,dt", [("periodic", "tg", 512, 2e-3), ("walls", "cavity", 300, 1e-3),
                                          ("tall", "cavity", 200, 1e-3)])
def test_paired_coarse_levels_are_bit_identical(lv, oracle, gpu, kind, ic, K, dt):
    """Pairs of coarse levels above the single-workgroup tail run as one down and one up kernel each
    (amg_level_kernels 0) or every level as its own kernels (1): the same cycle output, pressure and
    iterations bit for bit.  At 512^2 periodic levels 2 and 3 (128^2 and 64^2 cells) are paired, at
    300^2 with walls levels 2 and 3 (75^2 and 38^2 cells: odd widths, partial tiles at the walls), and
    on the walled 1 x 2 domain at 200 x 400 levels 2 and 3 (50 x 100 and 25 x 50 cells: unequal
    axes)."""
    _flush("amg paired coarse levels %s %d" % (kind, K))
    grid, op, solver, x, dr, d, rp, cols, coef = _system(lv, oracle, kind, K, ic, dt)
    b, _ = solver.rhs(dt, False)
    if not np.any(d > 0):
        b = np.random.default_rng(6).standard_normal(grid.n)
        b -= b.mean()
    x0 = grid.get("P")
    r = np.random.default_rng(13).standard_normal(grid.n)
    out = {}
    for lk in (1, 0):
        P = solver.solve(b, x0, method="amg", amg_level_kernels=lk)
        st = solver.last_stats
        assert max(st["relres"]) <= 1e-8, st
        z = solver.apply_preconditioner(r, "amg")  # the cycle with the last solve's kernels
        out[lk] = (P, z, list(st["iters"][:st["nsolves"]]), list(st["outer"][:st["nsolves"]]))
    assert np.all(np.isfinite(out[0][1])) and np.linalg.norm(out[0][1]) > 0
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][2] == out[1][2] and out[0][3] == out[1][3]


def test_find_pressure_compact_copies_are_bit_identical(lv, oracle, gpu):
    """find_pressure! (10 Picard iterations with the window) with the compact copies of the slots past
    the sixth, the operator's and the right-hand side's edge geometry (amg_layout 0), and without them
    (2): the same pressure bit for bit at 256^2 (65536 rows, where the copies are used)."""
    _flush("find_pressure compact copies 256")
    grid, op, x, dr = make_pair(lv, oracle, "periodic", K=256, seed=0x5EED0507, ic="tg")
    lv.stiffened_eos(grid, 1.4, 2500.0 / 1.4)
    P0 = grid.get("P")
    out = {}
    for layout in (0, 2):
        grid.set("P", P0)
        solver = lv.PressureSolver(grid, method="amg", tol=1e-10, itmax=4000, picard_window=4, amg_layout=layout)
        st = lv.find_pressure(solver, 2e-3, 10)
        assert max(st["relres"][:st["nsolves"]]) <= 1e-10, st
        out[layout] = (grid.get("P"), list(st["iters"][:st["nsolves"]]))
    assert np.array_equal(out[0][0], out[2][0])
    assert out[0][1] == out[2][1]


@pytest.mark.parametrize("method", ["pcg64", "mixed"])
def test_own_order_compact_overflow_is_bit_identical(lv, oracle, gpu, method):
    """The operator's own ELL read with the compact copy of the slots past the sixth (systems of at
    least 65536 rows, layout 0) and without it (layout 2): the FP64 and the own-order FP32 solves give
    the same pressure bit for bit and the same iterations."""
    _flush("own-order compact overflow %s" % method)
    grid, op, solver, x, dr, d, rp, cols, coef = _system(lv, oracle, "periodic", 256, "tg", 2e-3)
    assert grid.n >= 65536 and np.diff(rp).max() >= 8
    b, _ = solver.rhs(2e-3, False)
    x0 = grid.get("P")
    out = {}
    for layout in (0, 2):
        P = solver.solve(b, x0, method=method, tol=1e-10, itmax=4000, amg_layout=layout)
        st = solver.last_stats
        assert max(st["relres"]) <= 1e-10, st
        out[layout] = (P, st["iters"], st["outer"])
    assert np.array_equal(out[0][0], out[2][0])
    assert out[0][1] == out[2][1] and out[0][2] == out[2][2]


def _wrapped_offsets(rp, cols, n):
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = cols.astype(np.int64) - rows
    off = np.where(off > n // 2, off - n, off)
    return np.where(off < -(n // 2), off + n, off)


def test_fp64_passes_with_column_offsets_wider_than_16_bits(lv, gpu):
    """The FP64 residual and products read the operator's columns as 16-bit offsets j - i (modulo
    n) when every one fits, and the 32-bit columns otherwise.  The same 300^2 particles (90000 rows)
    in lattice order (every offset fits) and in a random order (most do not): each solve meets its
    tolerance on the residual recomputed on the host from the exported operator, and the two give
    the same pressure to 1e3 tol."""
    _flush("fp64 passes, 16- and 32-bit column offsets")
    K, dt = 300, 2e-3
    dr = 1.0 / K
    x = lv.inputs.jittered_lattice(K, K, dr, seed=0x5EED0508)
    n = K * K
    perm = np.random.default_rng(8).permutation(n)
    orders = {"lattice": x, "random": np.concatenate([x[:n][perm], x[n:][perm]])}
    P_of = {}
    for name, xx in orders.items():
        grid = lv.VoronoiGrid(lv.UnitRectangle(), dr, xperiodic=True, yperiodic=True, positions=xx)
        lv.remesh(grid)
        ic = lv.inputs.taylor_green(xx, grid.get("area"))
        for k in ("v", "rho", "mass", "P", "e", "mu"):
            grid.set_soa(k, ic[k])
        lv.stiffened_eos(grid, ic["gamma"], ic["P0"])
        solver = lv.PressureSolver(grid, method="amg", tol=1e-10, itmax=4000)
        solver.assemble(dt)
        d, rp, cols, coef = solver.operator()
        wide = np.abs(_wrapped_offsets(rp, cols, n)).max() > 32767
        assert wide == (name == "random")
        b, _ = solver.rhs(dt, False)
        P = solver.solve(b, grid.get("P"), method="amg")
        st = solver.last_stats
        assert max(st["relres"]) <= 1e-10 and st["fallbacks"] == 0, st
        A = sp.csr_matrix((coef, cols, rp), shape=(n, n))
        AP = d * P + np.asarray(A.sum(axis=1)).ravel() * P - A @ P
        assert np.linalg.norm(b - AP) <= 1.05e-10 * np.linalg.norm(b)
        if name == "random":
            Pl = np.empty(n)
            Pl[perm] = P
            P = Pl
        P_of[name] = P
    assert rel_err(P_of["random"], P_of["lattice"]) <= 1e-7


def test_cycle_matches_oracle_learned_smoother(lv, oracle, gpu):
    _flush("amg cycle learned")
    grid, op, solver, x, dr, d, rp, cols, coef = _system(lv, oracle, "periodic", 96, "tg", 8e-3)
    th = lv.spai.load_params()
    assert th is not None
    lv.spai.set_params(grid, th)
    oe, on, w, md = lv.spai.outputs(grid, cols.size)
    g = spai_ref.Graph(d, rp, cols, coef)
    s = 1.0 / np.sqrt(g.D)
    Mt = sp.csr_matrix((w.astype(np.float64), (g.src, g.cols)), shape=(g.n, g.n)) + sp.diags(md.astype(np.float64))
    M = (sp.diags(s) @ Mt @ sp.diags(s)).tocsr()
    H = _hier("periodic", x, dr, d, rp, cols, coef, M=M)
    r = np.random.default_rng(12).standard_normal(grid.n)
    z = solver.apply_preconditioner(r, "amg")
    zr = H.apply(r)
    assert np.linalg.norm(z - zr) <= 2e-3 * np.linalg.norm(zr), (np.linalg.norm(z - zr) / np.linalg.norm(zr))


def _host_residual(op, dt, b, P):
    diag, coef = op.pressure_operator(dt)
    r = b - op.apply_A(diag, coef, P)
    return np.linalg.norm(r) / np.linalg.norm(b)


@pytest.mark.parametrize("kind,ic,K,dt,learned", [("periodic", "tg", 128, 4e-3, False), ("periodic", "tg", 128, 4e-3, True),
                                                  ("walls", "cavity", 96, 1e-3, True)])
def test_amg_solve_meets_tolerance_with_few_iterations(lv, oracle, gpu, kind, ic, K, dt, learned):
    """The multilevel solve meets the FP64 tolerance (residual recomputed by the oracle on the host),
    agrees with a 1e-13 oracle solve to 1e3 tol, and needs at most a quarter of Jacobi's inner
    iterations."""
    _flush("amg solve %s %d learned=%s" % (kind, K, learned))
    grid, op, solver, x, dr, d, rp, cols, coef = _system(lv, oracle, kind, K, ic, dt)
    if learned:
        lv.spai.set_params(grid, lv.spai.load_params())
    b, _ = solver.rhs(dt, False)
    if not np.any(d > 0):  # singular: a random consistent right-hand side (the cavity's is zero)
        b = np.random.default_rng(3).standard_normal(grid.n)
        b -= b.mean()
    x0 = grid.get("P")
    P = solver.solve(b, x0, method="amg")
    st = solver.last_stats
    assert max(st["relres"]) <= 1e-8 and st["fallbacks"] == 0, st
    bh = b - b.mean() if not np.any(d > 0) else b
    assert _host_residual(op, dt, bh, P) <= 1.05e-8
    diag, coef_o = op.pressure_operator(dt)
    Po, it, rr = op.pcg(diag, coef_o, bh, x0, 1e-13)
    if not np.any(d > 0):
        Po = Po - Po.mean() + P.mean()
    # 1e3 tol; the singular Laplacian's condition number (~K^2) allows 1e4 tol
    assert rel_err(P, Po) <= (1e-4 if not np.any(d > 0) else 1e-5)
    solver.solve(b, x0, method="mixed")
    it_j = solver.last_stats["iters"][0]
    print("inner iterations: jacobi %d, multilevel %d" % (it_j, st["iters"][0]))
    assert st["iters"][0] <= 0.25 * it_j, (st["iters"][0], it_j)


def test_amg_solve_is_deterministic(lv, oracle, gpu):
    _flush("amg determinism")
    grid, op, solver, x, dr, d, rp, cols, coef = _system(lv, oracle, "periodic", 96, "tg", 4e-3)
    lv.spai.set_params(grid, lv.spai.load_params())
    b, _ = solver.rhs(4e-3, False)
    x0 = grid.get("P")
    P1 = solver.solve(b, x0, method="amg")
    solver.assemble(4e-3)  # rebuilds the hierarchy and re-runs the network
    P2 = solver.solve(b, x0, method="amg")
    assert np.array_equal(P1, P2)


def test_find_pressure_amg_matches_oracle(lv, oracle, gpu):
    """find_pressure! (10 Picard solves, src/pressure.jl:215-224) with the multilevel solver."""
    _flush("amg find_pressure")
    grid, op, x, dr = make_pair(lv, oracle, "periodic", K=48, seed=0x5EED0501, ic="tg")
    lv.stiffened_eos(grid, 1.4, 2500.0 / 1.4)
    op.eos_stiffened(1.4, 2500.0 / 1.4)
    dt = 2e-3
    lv.spai.set_params(grid, lv.spai.load_params())
    solver = lv.PressureSolver(grid, method="amg", tol=1e-12, picard_window=8)
    st = lv.find_pressure(solver, dt)
    op.find_pressure(dt, 10, tol=1e-13)
    assert max(st["relres"]) <= 1e-12
    assert rel_err(grid.get("P"), op["P"]) <= 1e-8


def test_solve_times_are_reported_for_every_solve(lv, oracle, gpu):
    """Each solve's time comes from an event pair that is read when the API call returns (one wait
    per call): all 10 Picard solves of find_pressure! report a positive time, and together no more
    than the call's wall time."""
    import time

    _flush("amg solve times")
    grid, op, x, dr = make_pair(lv, oracle, "periodic", K=48, seed=0x5EED0502, ic="tg")
    lv.stiffened_eos(grid, 1.4, 2500.0 / 1.4)
    solver = lv.PressureSolver(grid, method="amg", tol=1e-8, picard_window=8)
    lv.find_pressure(solver, 2e-3)  # set-up outside the timed call
    t0 = time.perf_counter()
    st = lv.find_pressure(solver, 2e-3)
    wall_ms = (time.perf_counter() - t0) * 1e3
    ms = st["solve_ms"]
    assert len(ms) == 10 and all(0.0 < m < wall_ms for m in ms), (ms, wall_ms)
    assert sum(ms) <= wall_ms, (sum(ms), wall_ms)


def test_solver_launches_stop_at_convergence(lv, oracle, gpu):
    """The host keeps one CG iteration queued ahead of the device and stops launching once the
    device reports convergence in pinned memory: per refinement sweep the SpMV kernel is launched
    at most about one iteration past the last working one.  A sampling timer brackets every k-th
    launch and still counts them all."""
    import ctypes

    _flush("amg launch pacing")
    grid, op, solver, x, dr, d, rp, cols, coef = _system(lv, oracle, "periodic", 301, "tg", 2e-3)
    b, _ = solver.rhs(2e-3, False)
    x0 = grid.get("P")
    L = lv.lib()
    for method, kid, every in (("amg", 0, 1), ("amg", 0, 4), ("pcg64", 1, 1)):
        assert L.lv_timing_enable(grid.ctx, kid, every) == 0
        solver.solve(b, x0, method=method)
        seen, timed, tot, bpl = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        assert L.lv_timing_seen(grid.ctx, kid, ctypes.byref(seen)) == 0
        assert L.lv_timing_read(grid.ctx, kid, ctypes.byref(tot), ctypes.byref(timed), ctypes.byref(bpl)) == 0
        L.lv_timing_enable(grid.ctx, kid, 0)
        st = solver.last_stats
        iters, sweeps = st["iters"][0], max(st["outer"][0], 1)
        assert st["fallbacks"] == 0 and max(st["relres"]) <= 1e-8
        assert iters <= seen.value <= iters + 2 * sweeps, (method, seen.value, iters, sweeps)
        assert timed.value == (seen.value + every - 1) // every and tot.value > 0.0, (every, timed.value, seen.value)


def test_solve_from_converged_guess_skips_the_sweep(lv, oracle, gpu):
    """The tests before the first refinement sweep run on the device (no synchronising read): a solve
    whose initial guess already meets the tolerance queues a sweep that starts done, reports it to
    the host at once (well inside the host's 20 ms no-progress fallback), changes nothing and is not
    counted."""
    _flush("mixed solve from a converged guess")
    grid, op, solver, x, dr, d, rp, cols, coef = _system(lv, oracle, "periodic", 301, "tg", 2e-3)
    b, _ = solver.rhs(2e-3, False)
    x0 = grid.get("P")
    lv.spai.set_params(grid, lv.spai.load_params())
    for method in ("amg", "mixed", "spai"):
        P1 = solver.solve(b, x0, method=method, tol=1e-8)
        st1 = solver.last_stats
        assert st1["iters"][0] > 0 and st1["outer"][0] >= 1 and st1["relres"][0] <= 1e-8, (method, st1)
        P2 = solver.solve(b, P1, method=method, tol=1e-8)
        st2 = solver.last_stats
        assert st2["iters"][0] == 0 and st2["outer"][0] == 0 and st2["fallbacks"] == 0, (method, st2)
        assert st2["relres"][0] <= 1e-8 and st2["relres0"][0] == st2["relres"][0], (method, st2)
        assert st2["solve_ms"][0] < 15.0, (method, st2["solve_ms"])
        assert np.array_equal(P2, P1), method
