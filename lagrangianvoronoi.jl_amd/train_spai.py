"""Train the learned SPAI preconditioner on the GPU and save its parameters.

    python lagrangianvoronoi.jl_amd/train_spai.py [--lr 2e-4] [--out path.npy]

Systems are harvested from plain-solver (Jacobi-PCG) runs of small seeded flows of each
configuration family (Taylor-Green at several c0 dt/dr regimes, lid-driven cavity, Rayleigh-Taylor);
after each step the pressure operator is assembled and Adam takes a few steps on it.  Before and
after training, one held-out pressure solve of the benchmark regime is timed with Jacobi and with
the learned preconditioner (inner FP32 iterations, FP64 refinement to 1e-8).
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

SCHEDULE = [("tg", 64, 1.0, 8), ("cavity", 48, 1.0, 6), ("tg", 64, 0.5, 8), ("rt", 32, 1.0, 6), ("tg", 96, 1.0, 8),
            ("tg", 64, 2.0, 8), ("cavity", 64, 2.0, 6), ("tg", 96, 0.7, 8)]


def held_out(lv, theta, K=128, dt_scale=1.0):
    g, desc, dt = lv.spai.flow("tg", K, dt_scale, seed=0x5EED0999)
    lv.spai.set_params(g, theta)
    lv.stiffened_eos(g, 1.4, 2500.0 / 1.4)
    solver = lv.PressureSolver(g, method="mixed", tol=1e-8, itmax=4000)
    solver.assemble(dt)
    b, _ = solver.rhs(dt, False)
    x0 = g.get("P")
    out = {}
    for m in ("mixed", "spai"):
        solver.solve(b, x0, method=m)
        st = solver.last_stats
        out[m] = (st["iters"][0], st["solve_ms"][0], st["relres"][0], st["fallbacks"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lr", type=float, default=2e-4)
    ap.add_argument("--steps-per-system", type=int, default=8)
    ap.add_argument("--nprobe", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=1)
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    import torch

    assert torch.cuda.is_available()
    from __graft_entry__ import load_package

    lv = load_package()
    th = lv.spai.init_params(0)
    print("held-out before:", held_out(lv, th), flush=True)
    t0 = time.time()
    for ep in range(args.epochs):
        th, losses = lv.spai.train(th, SCHEDULE, steps_per_system=args.steps_per_system, lr=args.lr,
                                   nprobe=args.nprobe, log=lambda m: print(m, flush=True), seed=7 + 100 * ep)
        print("epoch %d loss first %.4e last %.4e (%.1f s)" % (ep, np.mean(losses[:8]), np.mean(losses[-8:]),
                                                             time.time() - t0), flush=True)
    print("held-out after:", held_out(lv, th), flush=True)
    if args.out:
        np.save(args.out, th.astype(np.float32))
        print("saved", args.out)


if __name__ == "__main__":
    main()
