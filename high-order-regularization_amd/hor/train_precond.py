"""Trains the learned preconditioner on the GPU and evaluates it.

    python -m hor.train_precond --geoms 300 --h 0.015625 --epochs 12 --out weights.npz --report report.json

Data: (z_j, A z_j) pairs logged by the solver's own flexible GMRES (FP64
direct operator) on seeded sphere/ellipsoid pairs (radii U(0.3, 0.6), gaps
U(0.01, 0.2), random orientations), two right-hand sides per geometry.
Training: Adam on mean |M(A z) - z|^2/|z|^2 with the network's own forward
and backward kernels.  Evaluation: flexible GMRES iterations to 1e-10 with
and without the network on held-out geometries.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def log(msg):
    print(msg, flush=True)


def evaluate(ctx, net, params, alphas, n_geom, h, seed, tol=1e-10):
    """Plain vs learned flexible GMRES (correction scaled by each alpha) on
    held-out geometries: the network on every step (ungated) and under the
    solver's default gate (hor_solver_set_precond_policy)."""
    from . import mpnn as M
    from ._lib import DIRECT, FP64
    from .solver import Solver
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(n_geom):
        DEL, bodies, pd, gap = M.random_pair(rng, h)
        S = Solver(ctx, h, DEL, pd[0], pd[1], method=DIRECT, precision=FP64)
        net.set_geometry(pd, bodies)
        for kind, b in zip(("rigid", "smooth"), M.random_rhs(rng, bodies, pd)):
            _, st0 = M.solve(S, b, net=None, tol=tol, maxit=200)
            row = dict(gap=gap, rhs=kind, n=S.n, plain_iters=st0["iterations"], plain_res=st0["relres"],
                       learned={}, gated={})
            for a in alphas:
                net.set_params(M.scale_output(params, a))
                _, st1 = M.solve(S, b, net=net, tol=tol, maxit=200, gate=False)
                _, st2 = M.solve(S, b, net=net, tol=tol, maxit=200)
                row["learned"][str(a)] = dict(iters=st1["iterations"], res=st1["relres"])
                row["gated"][str(a)] = dict(iters=st2["iterations"], res=st2["relres"],
                                            learned_steps=st2["learned_steps"],
                                            misfit_identity=st2["gate_misfit_identity"],
                                            misfit_network=st2["gate_misfit_network"])
            rows.append(row)
        S.close()
    net.set_params(params)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--geoms", type=int, default=300)
    ap.add_argument("--val-geoms", type=int, default=16)
    ap.add_argument("--h", type=float, default=1.0 / 64.0)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--eval-geoms", type=int, default=8)
    ap.add_argument("--out", default=None, help="weights file (default: the package's precond_weights.npz)")
    ap.add_argument("--report", default=None, help="JSON report path")
    a = ap.parse_args(argv)
    import hor
    from . import mpnn as M
    ctx = hor.Context(0)
    t0 = time.time()
    train = M.make_dataset(ctx, a.geoms, a.h, seed=a.seed, log=log)
    val = M.make_dataset(ctx, a.val_geoms, a.h, seed=a.seed + 1, log=None)
    npairs = sum(s.count for s in train)
    log(f"dataset: {npairs} training pairs on {len(train)} geometries, {sum(s.count for s in val)} validation "
        f"pairs, {time.time() - t0:.1f}s")
    t1 = time.time()
    net, hist = M.train(ctx, train, epochs=a.epochs, lr=a.lr, seed=a.seed, val=val, log=log)
    train_s = time.time() - t1
    p = net.get_params()
    for s in train + val:
        s.free()
    # scale of the learned correction chosen on held-out solves (0 = plain GMRES)
    alphas = [0.25, 0.5, 0.75, 1.0]
    rows = evaluate(ctx, net, p, alphas, a.eval_geoms, a.h, a.seed + 2)
    pi = np.array([r["plain_iters"] for r in rows], float)
    mean_l = {al: float(np.mean([r["learned"][str(al)]["iters"] for r in rows])) for al in alphas}
    mean_g = {al: float(np.mean([r["gated"][str(al)]["iters"] for r in rows])) for al in alphas}
    best = min(alphas, key=lambda al: (mean_l[al], -al))
    log(f"held-out: mean iterations plain {pi.mean():.2f}; learned by alpha {mean_l}; gated by alpha {mean_g}; "
        f"best alpha {best}; max FP64 residual with the network "
        f"{max(r['learned'][str(al)]['res'] for r in rows for al in alphas):.2e}")
    for kind in ("rigid", "smooth"):
        sel = [r for r in rows if r["rhs"] == kind]
        log(f"  {kind}: plain {np.mean([r['plain_iters'] for r in sel]):.2f}, learned "
            f"{np.mean([r['learned'][str(best)]['iters'] for r in sel]):.2f}, gated "
            f"{np.mean([r['gated'][str(best)]['iters'] for r in sel]):.2f}")
    meta = dict(h=a.h, geoms=a.geoms, pairs=npairs, epochs=a.epochs, lr=a.lr, seed=a.seed,
                val_loss_identity=hist[0][2], val_loss_final=hist[-1][2], alpha=best,
                heldout_plain_iters=float(pi.mean()), heldout_learned_iters=mean_l[best],
                heldout_gated_iters=mean_g[best])
    M.save_weights(M.scale_output(p, best), meta, a.out or M.WEIGHTS_PATH)
    rep = dict(meta=meta, history=hist, train_seconds=train_s, eval=rows,
               mean_plain_iters=float(pi.mean()), mean_learned_iters=mean_l, mean_gated_iters=mean_g)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    net.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
