"""Trains the learned preconditioner on the GPU and evaluates it.

    python -m hor.train_precond --geoms 400 --h 0.0416667 --epochs 30 --out weights.npz --report report.json

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
    held-out geometries."""
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
            _, it0, _, r0, _ = S.fgmres(b, m=15, tol=tol, maxit=200)
            row = dict(gap=gap, rhs=kind, n=S.n, plain_iters=it0, plain_res=r0, learned={})
            for a in alphas:
                net.set_params(M.scale_output(params, a))
                _, it1, _, r1, _ = S.fgmres(b, m=15, tol=tol, maxit=200, precond=net)
                row["learned"][str(a)] = dict(iters=it1, res=r1)
            rows.append(row)
        S.close()
    net.set_params(params)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--geoms", type=int, default=560)
    ap.add_argument("--val-geoms", type=int, default=16)
    ap.add_argument("--h", type=float, default=1.0 / 24.0)
    ap.add_argument("--epochs", type=int, default=30)
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
    best = min(alphas, key=lambda al: mean_l[al])
    log(f"held-out: mean iterations plain {pi.mean():.2f}; learned by alpha {mean_l}; best alpha {best}; "
        f"max FP64 residual with the network "
        f"{max(r['learned'][str(al)]['res'] for r in rows for al in alphas):.2e}")
    meta = dict(h=a.h, geoms=a.geoms, pairs=npairs, epochs=a.epochs, lr=a.lr, seed=a.seed,
                val_loss_identity=hist[0][2], val_loss_final=hist[-1][2], alpha=best,
                heldout_plain_iters=float(pi.mean()), heldout_learned_iters=mean_l[best])
    M.save_weights(M.scale_output(p, best), meta, a.out or M.WEIGHTS_PATH)
    rep = dict(meta=meta, history=hist, train_seconds=train_s, eval=rows,
               mean_plain_iters=float(pi.mean()), mean_learned_iters=mean_l)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    net.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
