"""Train the learned SPAI preconditioner on the GPU and save its parameters.

    python lagrangianvoronoi.jl_amd/train_spai.py [--systems 20000] [--lr 2e-4] [--out path.npy]

Training data are 20,000 pressure systems harvested from plain-solver runs: small seeded flows of
every configuration family (Taylor-Green at three c0 dt/dr regimes, lid-driven cavity at two,
Rayleigh-Taylor, Gresho), each stepped with FP64 Jacobi-PCG; after every step the pressure operator
is assembled and Adam takes one step on it (probes: the step's pressure and Rademacher vectors).
Before and after training, held-out pressure solves of the benchmark regime (Taylor-Green, the
2048^2 benchmark's c0 dt/dr, at 128^2 and 256^2, seeds not used in training) are solved to 1e-8 with
Jacobi, with the untrained network and with the trained one, and their inner iterations printed.
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

# (kind, K, dt_scale) of the harvested flows; each flow runs STEPS_PER_FLOW steps
FAMILIES = [("tg", 48, 1.0), ("tg", 48, 0.5), ("tg", 48, 2.0), ("tg", 64, 1.0), ("cavity", 48, 1.0),
            ("cavity", 48, 2.0), ("rt", 32, 1.0), ("gresho", 48, 1.0)]
STEPS_PER_FLOW = 250


def schedule(systems):
    """Flows of every family in turn until `systems` steps are scheduled."""
    out, k = [], 0
    while sum(s[3] for s in out) < systems:
        kind, K, dts = FAMILIES[k % len(FAMILIES)]
        out.append((kind, K, dts, min(STEPS_PER_FLOW, systems - sum(s[3] for s in out))))
        k += 1
    return out


def held_out(lv, thetas, K=128, seed=0x5EED0999):
    """Inner iterations to 1e-8 of one pressure solve of the benchmark regime (Taylor-Green, EOS
    pressure as the initial guess): Jacobi, then the learned preconditioner for each theta."""
    g, desc, dt = lv.spai.flow("tg", K, 1.0, seed=seed)
    lv.stiffened_eos(g, 1.4, 2500.0 / 1.4)
    solver = lv.PressureSolver(g, method="mixed", tol=1e-8, itmax=4000)
    solver.assemble(dt)
    b, _ = solver.rhs(dt, False)
    x0 = g.get("P")
    solver.solve(b, x0, method="mixed")
    out = {"jacobi": solver.last_stats["iters"][0]}
    for name, th in thetas.items():
        lv.spai.set_params(g, th)
        solver.solve(b, x0, method="spai")
        st = solver.last_stats
        assert st["fallbacks"] == 0 and max(st["relres"]) <= 1e-8, st
        out[name] = st["iters"][0]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=20000)
    ap.add_argument("--lr", type=float, default=2e-4)
    ap.add_argument("--nprobe", type=int, default=8)
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    import torch

    assert torch.cuda.is_available()
    from __graft_entry__ import load_package

    lv = load_package()
    th0 = lv.spai.init_params(0)
    t0 = time.time()
    sched = schedule(args.systems)
    th, losses, systems = lv.spai.train(th0, sched, steps_per_system=1, lr=args.lr, nprobe=args.nprobe,
                                        log=lambda m: print(m, flush=True), seed=7, log_every=50)
    L = np.asarray(losses)
    print("trained on %d harvested systems in %.1f s; probe loss: first 500 %.4e, last 500 %.4e"
          % (systems, time.time() - t0, L[:500].mean(), L[-500:].mean()), flush=True)
    for K in (128, 256):
        print("held-out %d^2:" % K, held_out(lv, {"untrained": th0, "trained": th}, K=K), flush=True)
    if args.out:
        np.save(args.out, th.astype(np.float32))
        print("saved", args.out)


if __name__ == "__main__":
    main()
