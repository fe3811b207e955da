"""Benchmark: full Lagrangian-Voronoi steps on one MI355X.

Workload (the metric's 4M-particle configuration): Taylor-Green vortex on the periodic unit square,
2048 x 2048 particles on a jittered lattice (+-0.1 dr, SplitMix64 seed 0x5EED0002), u = sin 2pi x
cos 2pi y, v = -cos 2pi x sin 2pi y, nu = 1e-3, stiffened gas c0 = 50, dt = 2.5e-4.  One step is the
reference test's step! (tests/taylorgreen.jl:61-71): move! + remesh, stiffened_eos!, find_pressure!
(10 Picard solves, mixed-precision PCG with the multilevel preconditioner to 1e-8 FP64 relative
residual; --smoother learned makes the learned preconditioner its fine-level smoother), pressure_step!, find_D!,
viscous_step!, find_dv!, relaxation_step! + remesh.  Inputs are resident on the device before the
timed region.

Prints one JSON line: particle-steps/s (value), the first Picard solve's time to 1e-8, the roofline
of the dominant kernel (FP32 SpMV of the inner PCG, timed with HIP events on the library's stream)
and the oracle timed on this host's cores (cpu_baseline).
"""
import argparse
import json
import re
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

TIMER_EVERY = 16
HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s at 4M particles; pressure-solve time to 1e-8 rel. residual"


# Configuration families at full size (the benchmark line is tg2048; bench.py --config runs the others).
CONFIGS = {
    "tg2048": None,
    # Taylor-Green 1024^2, the plain method in FP64: Jacobi-PCG to 1e-10, no initial-guess extrapolation
    "tg1024": dict(nx=1024, ny=1024, ymax=1.0, periodic=True, ic="tg", dt=2.5e-4, solver="pcg64", tol=1e-10,
                   window=0, steps=50, warmup=1),
    # Gresho vortex 100^2, incompressible, the plain method in FP64, 100 steps
    "gresho100": dict(nx=100, ny=100, ymax=1.0, periodic=True, ic="gresho", dt=1e-3, solver="pcg64", tol=1e-10,
                      window=0, steps=100, warmup=0, viscous=False),
    # Rayleigh-Taylor 1024 x 2048 walls, density ratio 3, gravity -1, the reference's RT step with
    # multiphase_projection! (dt = 0.1 dr, examples/rayleightaylor.jl:36)
    "rt1024": dict(nx=1024, ny=2048, ymax=2.0, periodic=False, ic="rt", dt=0.1 / 1024, solver="amg", tol=1e-8,
                   window=8, steps=4, warmup=1),
    # lid-driven cavity 4096^2, Re = 1000, incompressible, lid u = 1 (dt = min(0.1 dr, 0.1 Re dr^2),
    # examples/cavity.jl:32)
    "ldc4096": dict(nx=4096, ny=4096, ymax=1.0, periodic=False, ic="cavity", dt=min(0.1 / 4096, 100.0 / 4096 ** 2),
                    solver="amg", tol=1e-8, window=8, steps=200, warmup=1),
    # examples/cavity.jl in full: Re = 100, N = 100, populate_lloyd!, t_end = 0.1 Re, then the
    # centreline profiles by point_value against the published data the example plots against
    "ldc100-profiles": None,
}


def _profile_error(s, ours, ref_pos, ref_val):
    """largest |ours - ref| at the reference's interior abscissae, ours (smoothed by movingavg as the
    example plots it) interpolated linearly on s"""
    keep = (ref_pos > 0.0) & (ref_pos < 1.0)
    return float(np.max(np.abs(np.interp(ref_pos[keep], s, ours) - ref_val[keep])))


def run_ldc_profiles(lv, args, torch):
    """examples/cavity.jl:24-111: populate_lloyd!, 10,000 steps of the incompressible cavity at Re = 100,
    u(0.5, y) and v(x, 0.5) by point_value, compared with Abdelmigid et al. and Ghia et al. (the
    example's reference files)"""
    Re, N = 100.0, 100
    dr = 1.0 / N
    dt = min(0.1 * dr, 0.1 * Re * dr * dr)
    t_end = 0.1 * Re

    def ic(g):
        g.set("rho", 1.0)
        g.set("mass", g.get("area"))
        g.set("mu", 1.0 / Re)
        g.set("c2", np.inf)

    grid = lv.VoronoiGrid(lv.UnitRectangle(), dr)
    t0 = time.perf_counter()
    lv.populate_lloyd(grid, ic=ic)
    torch.cuda.synchronize()
    t_lloyd = time.perf_counter() - t0
    solver = lv.PressureSolver(grid, method=args.solver if args.solver_given else "amg", tol=1e-8, itmax=4000,
                               picard_window=8, picard_carry=args.picard_carry,
                               amg_level_kernels=args.amg_level_kernels)
    # As written, the example's step! (examples/cavity.jl:70-80) never recomputes the density of its
    # incompressible fluid, so the cells' mass drifts away from their area; relaxation_step!'s mass
    # flux then drains a wall cell to negative mass and the run ends in "The Voronoi Mesh has been
    # destroyed" (near step 1140 from these seeds, on the GPU and on the CPU oracle alike; steps 234 to
    # 1204 from other seeds, with the build's PCG or the reference's MINRES at 1e-6).  By default the
    # step therefore adds find_rho! (rho = m/area, src/pressure.jl:36-40) after move!, which keeps the
    # run to t_end = 10; --cavity-as-written runs the example's step unchanged.
    eos = None if args.cavity_as_written else "rho"
    desc = lv.make_step_desc(dt, eos=eos, solver=solver, viscous=True, artificial_viscosity=False,
                             friction={lv.BDARY_UP: (1.0, 0.0)})
    steps = args.steps
    if not args.steps_given:  # run!'s loop, src/simulation.jl:72-93: while t < t_end ... t += dt
        steps, t = 0, 0.0
        while t < t_end:
            steps += 1
            t += dt
    t0 = time.perf_counter()
    failed = None
    for k in range(steps):
        try:
            lv.step_fused(grid, desc)
        except lv.LVError as e:
            failed = dict(step=k + 1, error=str(e))
            break
        if (k + 1) % 1000 == 0:
            print("ldc100 step %d of %d" % (k + 1, steps), file=sys.stderr, flush=True)
        if (k + 1) % 2500 == 0:
            print("ldc100 step %d of %d" % (k + 1, steps), flush=True)
    grid.sync()
    elapsed = time.perf_counter() - t0
    if failed:
        print(json.dumps(dict(config="ldc100-profiles", particles=grid.n, steps_requested=steps, failed=failed,
                              density_update=eos == "rho", ms_per_step=round(elapsed / failed["step"] * 1e3, 3))))
        return
    s = np.linspace(0.0, 1.0, 100)
    u = lv.point_value(grid, np.stack([np.full(100, 0.5), s], 1), "v[1]")
    v = lv.point_value(grid, np.stack([s, np.full(100, 0.5)], 1), "v[2]")
    us, vs = np.array(lv.movingavg(list(u))), np.array(lv.movingavg(list(v)))
    gold = os.path.join(ROOT, "tests", "golden")

    def col(name, pos):
        import csv

        rows = list(csv.DictReader(open(os.path.join(gold, name))))
        return np.array([float(r[pos]) for r in rows]), np.array([float(r["Re100"]) for r in rows])

    err = {}
    for src in ("abdelmigid", "ghia"):
        y, ref_u = col("ldc_y2vx_%s.csv" % src, "y")
        x, ref_v = col("ldc_x2vy_%s.csv" % src, "x")
        err[src] = dict(u_on_x_0p5=round(_profile_error(s, us, y, ref_u), 4),
                        v_on_y_0p5=round(_profile_error(s, vs, x, ref_v), 4))
    out = dict(config="ldc100-profiles", particles=grid.n, steps=steps, t_end_of_the_example=t_end, dt=dt,
               density_update=eos == "rho",
               ms_per_step=round(elapsed / steps * 1e3, 3), lloyd_s=round(t_lloyd, 3),
               max_abs_error_vs_published=err, u_min=round(float(us.min()), 4), v_min=round(float(vs.min()), 4),
               v_max=round(float(vs.max()), 4))
    print(json.dumps(out))


def run_config(lv, args, torch):
    """One configuration family at full size: steps/s, solver iterations and fallbacks, device memory."""
    c = CONFIGS[args.config]
    nx, ny = c["nx"], c["ny"]
    dr = 1.0 / nx
    x = lv.inputs.jittered_lattice(nx, ny, dr, seed=0x5EED0100)
    grid = lv.VoronoiGrid(lv.Rectangle((0.0, 0.0), (1.0, c["ymax"])), dr, xperiodic=c["periodic"],
                          yperiodic=c["periodic"], positions=x)
    lv.remesh(grid)
    area = grid.get("area")
    kw = dict(viscous=c.get("viscous", True))
    if c["ic"] == "tg":
        ic = lv.inputs.taylor_green(x, area)
        keys = ("v", "rho", "mass", "P", "e", "mu")
        kw.update(eos="stiffened", gamma=ic["gamma"], P0=ic["P0"])
    elif c["ic"] == "gresho":
        ic = lv.inputs.gresho(x, area)
        keys = ("v", "rho", "mass", "P", "e", "c2")
    elif c["ic"] == "rt":
        ic = lv.inputs.rayleigh_taylor(x, area)
        keys = ("v", "rho", "mass", "P", "e", "mu", "phase")
        kw.update(eos="ideal", gamma=1.4, gravity=(0.0, -1.0), artificial_viscosity=True,
                  multiphase=lv.MultiphaseSolver(grid))
    else:
        ic = lv.inputs.cavity(x, area)
        keys = ("v", "rho", "mass", "mu", "c2")
        kw.update(friction={lv.BDARY_UP: (1.0, 0.0)})
    for k in keys:
        grid.set_soa(k, ic[k])
    method = args.solver if args.solver_given else c["solver"]
    if method == "spai" or (method == "amg" and args.smoother == "learned"):
        lv.spai.set_params(grid, spai_theta(lv)[0])
    solver = lv.PressureSolver(grid, method=method, tol=c["tol"], itmax=4000, inner_tol=args.inner_tol,
                               picard_window=c["window"], amg_layout=args.amg_layout,
                               picard_carry=args.picard_carry, amg_level_kernels=args.amg_level_kernels)
    desc = lv.make_step_desc(c["dt"], solver=solver, **kw)
    steps = args.steps if args.steps_given else c["steps"]
    warmup = args.warmup if args.warmup_given else c["warmup"]
    for _ in range(warmup):
        lv.step_fused(grid, desc)
    grid.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = [lv.step_fused(grid, desc) for _ in range(steps)]
    grid.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    free, total = torch.cuda.mem_get_info()
    n = nx * ny
    nsolves = sum(s["nsolves"] for s in stats)
    out = dict(metric="particle-steps/s", config=args.config, particles=n, steps=steps, warmup=warmup,
               value=round(n * steps / (t1 - t0), 1), ms_per_step=round((t1 - t0) / steps * 1e3, 3),
               solver=method, smoother=args.smoother if method == "amg" else None, tol=c["tol"],
               inner_iters_per_step=round(sum(s["total_iters"] for s in stats) / steps, 1),
               first_solve_ms=round(sum(s["solve_ms"][0] for s in stats) / steps, 3),
               max_relres=max(max(s["relres"]) for s in stats),
               fallbacks=sum(s["fallbacks"] for s in stats), solves=nsolves,
               fallback_rate=round(sum(s["fallbacks"] for s in stats) / max(nsolves, 1), 4),
               multiphase_iters=[s["multiphase_iters"] for s in stats] if c["ic"] == "rt" else None,
               iters_last_step=list(stats[-1]["iters"][:stats[-1]["nsolves"]]),
               device_memory_in_use_GB=round((total - free) / 1e9, 2), device_memory_GB=round(total / 1e9, 1))
    if c["ic"] in ("rt", "cavity") and not args.no_compare:
        # the pressure-solve comparison at the state the timed steps reached
        eos = (lambda g: lv.ideal_eos(g, 1.4)) if c["ic"] == "rt" else (lambda g: None)
        out["pressure_solve_compare"] = solve_compare(lv, grid, ic, c["dt"], c["tol"], args.inner_tol, eos=eos)
    print(json.dumps(out))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--K", type=int, default=2048, help="lattice side (K*K particles)")
    ap.add_argument("--solver", default="amg", choices=["pcg64", "mixed", "spai", "amg"],
                    help="pcg64: FP64 Jacobi-PCG; mixed: FP32 Jacobi-PCG in FP64 refinement; spai: the same with the "
                         "learned preconditioner (committed trained parameters); amg: the same with the multilevel "
                         "preconditioner whose fine smoother is the learned one")
    ap.add_argument("--smoother", default="jacobi", choices=["learned", "jacobi"],
                    help="fine-level smoother of --solver amg: the learned preconditioner or damped Jacobi")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--inner-tol", type=float, default=1e-4,
                    help="floor of the FP32 sweep's relative tolerance inside the FP64 refinement")
    ap.add_argument("--picard-carry", type=int, default=1, choices=[0, 1],
                    help="carry the first Picard corrections of a step into the next (lv_solve_opts.picard_carry)")
    ap.add_argument("--amg-layout", type=int, default=0, choices=[0, 1, 2, 3],
                    help="fine-level layout of the multilevel solver (lv_solve_opts.amg_layout)")
    ap.add_argument("--amg-level-kernels", type=int, default=0, choices=[0, 1],
                    help="1: every coarse level above the tail as its own kernels (lv_solve_opts.amg_level_kernels)")
    ap.add_argument("--picard-window", type=int, default=8,
                    help="Picard corrections used to extrapolate each solve's initial guess (0: previous P)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP64 / mixed / learned-preconditioner solve comparison")
    ap.add_argument("--cpu-K", type=int, default=2048, help="lattice side of the CPU sample (one full step)")
    ap.add_argument("--err-steps", type=int, default=2,
                    help="steps after the timed ones whose pressure is compared with an FP64 solve of the same "
                         "systems (untimed)")
    ap.add_argument("--config", default="tg2048", choices=sorted(CONFIGS),
                    help="tg2048: the benchmark line; the others run a configuration family at full size and "
                         "print its own line (steps/s, iterations, fallbacks, device memory in use)")
    ap.add_argument("--cavity-as-written", action="store_true",
                    help="ldc100-profiles: the example's step! unchanged (no density update; the run ends in a "
                         "destroyed mesh near t = 1.1)")
    ap.add_argument("--mode", default="steps", choices=["steps", "solve", "copy"],
                    help="steps: the benchmark; solve: one pressure solve (for counter runs); "
                         "copy: a plain device copy of known size (calibrates HBM byte counters)")
    args = ap.parse_args()
    argv = sys.argv[1:]
    args.solver_given = any(a.startswith("--solver") for a in argv)
    args.steps_given = any(a.startswith("--steps") for a in argv)
    args.warmup_given = any(a.startswith("--warmup") for a in argv)
    return args


def mode_copy():
    """Copies 512 MiB device-to-device 5 times with torch: each launch reads and writes 512 MiB."""
    import torch
    n = 128 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    print(json.dumps({"mode": "copy", "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n,
                      "ms_per_launch": round(ms, 4), "GBps_read_plus_write": round(8 * n / (ms * 1e-3) / 1e9, 1)}))


def spai_theta(lv):
    """The committed trained parameters, else the untrained network (zero decoders)."""
    th = lv.spai.load_params()
    return (th, True) if th is not None else (lv.spai.init_params(0), False)


def _best_solve(lv, grid, method, dt, tol, inner_tol, x0, repeats):
    solver = lv.PressureSolver(grid, method=method, tol=tol, itmax=20000, inner_tol=inner_tol)
    best = None
    for _ in range(repeats):
        solver.assemble(dt)
        b, _ = solver.rhs(dt, False)
        solver.solve(b, x0, raise_on_fail=False)
        st = solver.last_stats
        rec = dict(ms=round(st["solve_ms"][0], 3), iters=st["iters"][0], sweeps=st["outer"][0],
                   relres=st["relres"][0], fallbacks=st["fallbacks"])
        if best is None or rec["ms"] < best["ms"]:
            best = rec
    return best


def solve_compare(lv, grid, ic, dt, tol, inner_tol, repeats=3, eos=None):
    """The pressure-solve metric: the first Picard solve of a step (operator and RHS of the current
    state, EOS pressure as initial guess, src/pressure.jl:172) to tol, best of `repeats` solves, each
    timed with HIP events on the library's stream, with
      pcg64                 FP64 Jacobi-PCG (the plain method, the judge);
      mixed                 FP32 Jacobi-PCG sweeps inside FP64 iterative refinement;
      spai                  the same with the learned preconditioner;
      amg                   the same with the multilevel preconditioner (damped-Jacobi smoother);
      amg_learned_smoother  the multilevel preconditioner with the learned preconditioner as its
                            fine-level smoother.
    Set-up is inside the timed solve (network inference, hierarchy); it runs once per assembled
    operator and is forced here for every repeat by re-assembling.  `eos` sets the pressure and
    sound speed of the state (the flow's equation of state; default the stiffened gas of ic)."""
    th, trained = spai_theta(lv)
    if eos is None:
        lv.stiffened_eos(grid, ic["gamma"], ic["P0"])
    else:
        eos(grid)
    x0 = grid.get("P")
    out = {}
    lv.spai.set_params(grid, None)
    for method in ("pcg64", "mixed", "amg"):
        out[method] = _best_solve(lv, grid, method, dt, tol, inner_tol, x0, repeats)
    lv.spai.set_params(grid, th)
    out["spai"] = _best_solve(lv, grid, "spai", dt, tol, inner_tol, x0, repeats)
    out["amg_learned_smoother"] = _best_solve(lv, grid, "amg", dt, tol, inner_tol, x0, repeats)
    out["spai_trained"] = trained
    for k in ("mixed", "spai", "amg", "amg_learned_smoother"):
        out["speedup_%s_vs_pcg64" % k] = round(out["pcg64"]["ms"] / out[k]["ms"], 3)
    return out


def mode_solve(lv, K, method, tol, inner_tol, smoother):
    """One first-Picard pressure solve of the benchmark flow (EOS pressure as initial guess)."""
    dt = 2.5e-4
    grid, x, ic = setup(lv, K)
    if method == "spai" or (method == "amg" and smoother == "learned"):
        lv.spai.set_params(grid, spai_theta(lv)[0])
    lv.stiffened_eos(grid, ic["gamma"], ic["P0"])
    solver = lv.PressureSolver(grid, method=method, tol=tol, itmax=2000, inner_tol=inner_tol)
    solver.assemble(dt)
    b, _ = solver.rhs(dt, False)
    P = solver.solve(b, grid.get("P"))
    print(json.dumps({"mode": "solve", "particles": K * K, "stats": solver.last_stats}))


def setup(lv, K, seed=0x5EED0002):
    dr = 1.0 / K
    x = lv.inputs.jittered_lattice(K, K, dr, seed=seed)
    grid = lv.VoronoiGrid(lv.UnitRectangle(), dr, xperiodic=True, yperiodic=True, positions=x)
    lv.remesh(grid)
    area = grid.get("area")
    ic = lv.inputs.taylor_green(x, area)
    for k in ("v", "rho", "mass", "P", "e", "mu"):
        grid.set_soa(k, ic[k])
    return grid, x, ic


STATE_FIELDS = ("x", "v", "dv", "momentum", "rho", "e", "P", "c2", "mass", "energy", "quality", "mu", "D", "phase")


def fp64_error(lv, grid, desc, steps, tol64=1e-10):
    """Config 3's error against an FP64 solve, per step, outside the timed region: each step runs
    twice from the same state, once with the benchmark's mixed-precision solver and once with FP64
    Jacobi-PCG to tol64 (the plain method), and the pressures after find_pressure! are compared; the
    run continues from the mixed-precision step."""
    solver64 = lv.PressureSolver(grid, method="pcg64", tol=tol64, itmax=20000)
    desc64 = lv.make_step_desc(desc.dt, eos="stiffened", gamma=desc.gamma, P0=desc.P0, solver=solver64,
                               viscous=True, artificial_viscosity=False, alpha=1.0, rusanov=True)
    rows = []
    for _ in range(steps):
        state = {k: grid.get_soa(k) for k in STATE_FIELDS}
        st64 = lv.step_fused(grid, desc64)  # raises if an FP64 solve misses tol64
        P64, v64 = grid.get_soa("P"), grid.get_soa("v")
        for k in STATE_FIELDS:
            grid.set_soa(k, state[k])
        lv.remesh(grid)
        st = lv.step_fused(grid, desc)
        P, v = grid.get_soa("P"), grid.get_soa("v")
        rows.append(dict(P_rel_l2=float(np.linalg.norm(P - P64) / np.linalg.norm(P64)),
                         P_rel_max=float(np.max(np.abs(P - P64)) / np.max(np.abs(P64))),
                         v_rel_max=float(np.max(np.abs(v - v64)) / np.max(np.abs(v64))),
                         max_relres=max(st["relres"]), max_relres_fp64=max(st64["relres"])))
    return dict(fp64_solver="pcg64", fp64_tol=tol64, steps=rows)


def step_compare(lv, grid, desc, steps=2):
    """Full steps (10 Picard solves each, outside the timed region) from the same state with each inner
    solver: the learned preconditioner's network inference runs once per step's assembly and serves
    all its solves, which is where a one-level learned preconditioner has to pay for itself.  ms per
    step from the host clock around synchronised steps, plus the inner iterations per step."""
    import torch

    th, trained = spai_theta(lv)
    state = {k: grid.get_soa(k) for k in STATE_FIELDS}
    out = {}
    for name, method, params in (("mixed", "mixed", None), ("spai", "spai", th), ("amg", "amg", None),
                                 ("amg_learned_smoother", "amg", th)):
        for k in STATE_FIELDS:
            grid.set_soa(k, state[k])
        lv.remesh(grid)
        lv.spai.set_params(grid, params)
        solver = lv.PressureSolver(grid, method=method, tol=desc.solve.tol, itmax=4000,
                                   inner_tol=desc.solve.inner_tol, picard_window=desc.solve.picard_window,
                                   picard_carry=desc.solve.picard_carry)
        d = lv.make_step_desc(desc.dt, eos="stiffened", gamma=desc.gamma, P0=desc.P0, solver=solver,
                              viscous=True, artificial_viscosity=False, alpha=1.0, rusanov=True)
        grid.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = [lv.step_fused(grid, d) for _ in range(steps)]
        grid.sync()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        out[name] = dict(ms_per_step=round((t1 - t0) / steps * 1e3, 2),
                         inner_iters_per_step=sum(s["total_iters"] for s in st) / steps,
                         fallbacks=sum(s["fallbacks"] for s in st), max_relres=max(max(s["relres"]) for s in st))
    for k in STATE_FIELDS:
        grid.set_soa(k, state[k])
    lv.remesh(grid)
    lv.spai.set_params(grid, None)
    out["spai_trained"] = trained
    out["spai_vs_mixed"] = round(out["mixed"]["ms_per_step"] / out["spai"]["ms_per_step"], 3)
    return out


def traffic_from_profiles(kernel):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, copy-kernel
    calibrated), committed under profiles/ as JSON; None if absent."""
    tag = re.sub(r"[^A-Za-z0-9]+", "_", kernel).strip("_")
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % tag)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def cpu_baseline(K, seed, dt, tol):
    """One full oracle step (same algorithm, FP64 Jacobi-PCG) at lattice side K on all host cores."""
    import numpy as np
    import importlib.util
    spec = importlib.util.spec_from_file_location("lvoracle", os.path.join(ROOT, "oracle", "lvoracle.py"))
    O = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(O)
    from __graft_entry__ import load_package
    lv = load_package()
    dr = 1.0 / K
    x = lv.inputs.jittered_lattice(K, K, dr, seed=seed)
    g = O.Grid((0.0, 0.0), (1.0, 1.0), dr, xperiodic=True, yperiodic=True)
    p = O.Particles(g, x)
    p.remesh()
    ic = lv.inputs.taylor_green(x, p.area())
    for k in ("v", "rho", "mass", "P", "e", "mu"):
        p.set(k, ic[k])
    t0 = time.perf_counter()
    p.step(dt, eos="stiffened", gamma=ic["gamma"], P0=ic["P0"], tol=tol)
    t1 = time.perf_counter()
    # one pressure solve (first Picard iteration, warm-started from the EOS pressure) to tol
    p.eos_stiffened(ic["gamma"], ic["P0"])
    diag, coef = p.pressure_operator(dt)
    b, P0, GP = p.rhs(dt, False)
    t2 = time.perf_counter()
    _, iters, _ = p.pcg(diag, coef, b, P0, tol)
    t3 = time.perf_counter()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = K * K
    return dict(value=n / (t1 - t0), unit="particle-steps/s", cores=threads, kind="port",
                sample="one full step of the same flow at %dx%d (%d particles), FP64 Jacobi-PCG to %g; "
                       "C restatement of the reference algorithm with OpenMP static chunks" % (K, K, n, tol),
                step_s=t1 - t0, pressure_solve_ms=(t3 - t2) * 1e3, pressure_solve_iters=int(iters))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.mode == "copy":
        return mode_copy()
    from __graft_entry__ import load_package

    lv = load_package()
    if args.mode == "solve":
        return mode_solve(lv, args.K, args.solver, args.tol, args.inner_tol, args.smoother)
    if args.config == "ldc100-profiles":
        return run_ldc_profiles(lv, args, torch)
    if args.config != "tg2048":
        return run_config(lv, args, torch)
    K = args.K
    n = K * K
    dt = 2.5e-4
    grid, x, ic = setup(lv, K)
    if args.solver == "spai" or (args.solver == "amg" and args.smoother == "learned"):
        lv.spai.set_params(grid, spai_theta(lv)[0])
    solver = lv.PressureSolver(grid, method=args.solver, tol=args.tol, itmax=2000, inner_tol=args.inner_tol,
                               picard_window=args.picard_window, amg_layout=args.amg_layout,
                               picard_carry=args.picard_carry, amg_level_kernels=args.amg_level_kernels)
    desc = lv.make_step_desc(dt, eos="stiffened", gamma=ic["gamma"], P0=ic["P0"], solver=solver,
                             viscous=True, artificial_viscosity=False, alpha=1.0, rusanov=True)
    for _ in range(args.warmup):
        lv.step_fused(grid, desc)
    L = lv.lib()
    dominant = 0  # FP32 SpMV of the inner PCG
    if args.solver == "pcg64":
        dominant = 1
    # HIP events around every 16th launch of the dominant kernel (an event pair adds a few
    # microseconds between the launches it brackets)
    L.lv_timing_enable(grid.ctx, dominant, TIMER_EVERY)
    grid.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(args.steps):
        stats.append(lv.step_fused(grid, desc))
    grid.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    import ctypes
    tot_ms, launches, bpl = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
    L.lv_timing_read(grid.ctx, dominant, ctypes.byref(tot_ms), ctypes.byref(launches), ctypes.byref(bpl))
    seen = ctypes.c_int64()
    L.lv_timing_seen(grid.ctx, dominant, ctypes.byref(seen))
    L.lv_timing_enable(grid.ctx, dominant, 0)
    elapsed = t1 - t0
    value = n * args.steps / elapsed
    first_solve_ms = sum(s["solve_ms"][0] for s in stats) / len(stats)
    first_solve_it = sum(s["iters"][0] for s in stats) / len(stats)
    relres = max(max(s["relres"]) for s in stats)
    avg_launch_ms = tot_ms.value / max(launches.value, 1)
    # a launch after convergence returns at once: the bytes are those of the working launches (one
    # per inner iteration) over the time of all launches (the sampled average times their number)
    working = sum(sum(s["iters"]) for s in stats)
    achieved = bpl.value * working / (avg_launch_ms * 1e-3 * max(seen.value, 1)) / 1e9
    # the FP32 SpMV kernel in the layout the solves used (aggregate order: 8-bit degrees, 16-bit offsets)
    if dominant == 0:
        kname = "k_cg_s<float, lv_ell_agg32>" if stats[-1].get("agg_order") else "k_cg_s<float>"
    else:
        kname = "k_cg_s<double>"
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic_from_profiles(kname),
                    kernel=kname, avg_launch_us=round(avg_launch_ms * 1e3, 2), launches=seen.value,
                    launches_timed=launches.value, working_launches=working,
                    avg_working_launch_us=round(avg_launch_ms * 1e3 * seen.value / max(working, 1), 2),
                    algorithmic_bytes_per_launch=bpl.value)
    out = dict(metric=METRIC, value=round(value, 1), unit="particle-steps/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=round(elapsed / args.steps * 1e3, 3), higher_is_better=True,
               scaling="weak", vs_baseline=None,
               dtype="f64 geometry/stencils/dots/refinement; f32 inner PCG (SpMV/AXPY); bf16 couplings in the "
                     "multilevel cycle's fine smoothing and residuals",
               data="synthetic",
               config=dict(workload="taylor-green %dx%d periodic, jitter +-0.1dr seed 0x5EED0002, dt=2.5e-4, "
                                    "c0=50 stiffened, nu=1e-3, 10 Picard solves/step, %s PCG to %g" %
                                    (K, K, args.solver, args.tol),
                           particles=n, solver=args.solver, tol=args.tol, picard_window=args.picard_window,
                           picard_carry=args.picard_carry, amg_level_kernels=args.amg_level_kernels,
                           smoother=args.smoother if args.solver == "amg" else None),
               pressure_solve_ms=round(first_solve_ms, 3), pressure_solve_iters=first_solve_it,
               max_relres=relres, picard_iters=stats[-1]["iters"], picard_sweeps=stats[-1]["outer"],
               picard_relres0=[float("%.3g" % r) for r in stats[-1]["relres0"]], roofline=roofline)
    if args.err_steps > 0:
        out["error_vs_fp64_solve"] = fp64_error(lv, grid, desc, args.err_steps)
    if not args.no_compare:
        out["pressure_solve_compare"] = solve_compare(lv, grid, ic, dt, args.tol, args.inner_tol)
        out["step_compare"] = step_compare(lv, grid, desc)
    if not args.no_cpu:
        cK = args.cpu_K
        out["cpu_baseline"] = cpu_baseline(cK, 0x5EED0002, dt, args.tol)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
