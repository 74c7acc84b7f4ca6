"""Learned right preconditioner for the two-surface GMRES (no counterpart in
the reference, whose GMRES is unpreconditioned, main.cpp:313-415).

``Network`` wraps the C ABI's point-cloud message-passing network (width 64,
k = 16 nearest neighbours, 3 layers; BF16 matrix-core GEMMs with FP32
accumulation; forward, backward and Adam on the GPU).  ``make_dataset`` logs
(z_j, A z_j) pairs from the solver's own flexible GMRES over seeded
sphere/ellipsoid pairs, ``train`` fits the network to map a residual
r = A z back to its correction z, and ``Solver.fgmres(..., precond=net)``
applies it as a right preconditioner; plain GMRES remains the fallback and
every solve ends on the FP64 true residual.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import _lib
from . import geometry as G
from ._lib import DIRECT, FP64, check, f64, ptr

WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "precond_weights.npz")


def param_count():
    return int(_lib.lib().hor_mpnn_param_count())


def body_size(body):
    """Length scale of a body: the mean semi-axis."""
    return float(np.mean(body.abc))


class Network:
    """Point-cloud network z = M(r) bound to one geometry at a time."""

    def __init__(self, ctx, k=16, params=None):
        self.ctx = ctx
        self.k = int(k)
        h = C.c_void_p()
        check(_lib.lib().hor_mpnn_create(ctx.h, self.k, C.byref(h)), "hor_mpnn_create")
        self.h = h
        ctx.adopt(self)
        self.npts = 0
        if params is not None:
            self.set_params(params)

    def close(self):
        if self.h:
            _lib.lib().hor_mpnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != param_count():
            raise ValueError(f"expected {param_count()} parameters, got {p.size}")
        check(_lib.lib().hor_mpnn_set_params(self.h, ptr(p)), "hor_mpnn_set_params")

    def get_params(self):
        p = np.zeros(param_count(), np.float32)
        check(_lib.lib().hor_mpnn_get_params(self.h, ptr(p)), "hor_mpnn_get_params")
        return p

    def get_grad(self):
        g = np.zeros(param_count(), np.float32)
        check(_lib.lib().hor_mpnn_get_grad(self.h, ptr(g)), "hor_mpnn_get_grad")
        return g

    def set_geometry(self, pd, bodies, batch_max=1):
        """pd: the solver's two surface dicts (x, nrml); bodies: their Body
        objects (centres and sizes).  Returns the neighbour graph (npts, k)."""
        n1, n2 = int(pd[0]["n"]), int(pd[1]["n"])
        x = np.concatenate([f64(pd[0]["x"]), f64(pd[1]["x"])])
        nr = np.concatenate([f64(pd[0]["nrml"]), f64(pd[1]["nrml"])])
        ctr = np.concatenate([bodies[0].ctr, bodies[1].ctr]).astype(np.float64)
        size = np.array([body_size(bodies[0]), body_size(bodies[1])])
        nbr = np.zeros((n1 + n2) * self.k, np.int32)
        check(_lib.lib().hor_mpnn_set_geometry(self.h, n1, n2, ptr(x), ptr(nr), ptr(ctr), ptr(size),
                                               int(batch_max), ptr(nbr)), "hor_mpnn_set_geometry")
        self.npts = n1 + n2
        self.geom = dict(n1=n1, n2=n2, x=x, nrml=nr, ctr=ctr, size=size)
        return nbr.reshape(n1 + n2, self.k)

    def forward_dev(self, batch, r_dev, z_dev):
        check(_lib.lib().hor_mpnn_forward_dev(self.h, int(batch), r_dev, z_dev), "hor_mpnn_forward_dev")

    def forward(self, r):
        """Host convenience: z = M(r) for r of shape (batch, 3*npts)."""
        r = np.ascontiguousarray(np.atleast_2d(r), dtype=np.float64)
        B = r.shape[0]
        rd, zd = self.ctx.alloc(r.nbytes), self.ctx.alloc(r.nbytes)
        try:
            self.ctx.upload(rd, r)
            self.forward_dev(B, rd, zd)
            return self.ctx.download(zd, np.zeros_like(r))
        finally:
            self.ctx.free(rd)
            self.ctx.free(zd)

    def loss_grad_dev(self, batch, r_dev, v_dev):
        loss = C.c_double()
        check(_lib.lib().hor_mpnn_loss_grad_dev(self.h, int(batch), r_dev, v_dev, C.byref(loss)),
              "hor_mpnn_loss_grad_dev")
        return loss.value

    def loss_grad(self, r, v):
        """Host convenience: loss and gradient for r, v of shape (batch, 3*npts)."""
        r = np.ascontiguousarray(np.atleast_2d(r), dtype=np.float64)
        v = np.ascontiguousarray(np.atleast_2d(v), dtype=np.float64)
        rd, vd = self.ctx.alloc(r.nbytes), self.ctx.alloc(v.nbytes)
        try:
            self.ctx.upload(rd, r)
            self.ctx.upload(vd, v)
            loss = self.loss_grad_dev(r.shape[0], rd, vd)
            return loss, self.get_grad()
        finally:
            self.ctx.free(rd)
            self.ctx.free(vd)

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        check(_lib.lib().hor_mpnn_adam_step(self.h, float(lr), float(beta1), float(beta2), float(eps)),
              "hor_mpnn_adam_step")


def init_params(seed=0):
    """Random hidden weights (He-style, from numpy PCG64(seed)); zero output
    layer, so that the untrained network is the identity M(r) = r."""
    from numpy.random import default_rng
    W, INPAD, L, UW = 64, 16, 3, 192
    rng = default_rng(seed)
    p = np.zeros(param_count(), np.float32)
    o = 0
    p[o:o + W * INPAD] = (rng.standard_normal((W, INPAD)) * np.sqrt(2.0 / 9) * (np.arange(INPAD) < 9)).ravel()
    o += W * INPAD + W
    for _ in range(L):
        p[o:o + W * W] = rng.standard_normal(W * W) * np.sqrt(1.0 / W)
        o += W * W
        bd = rng.standard_normal((W, 4)) * 0.3
        bd[:, 3] = 0.0
        p[o:o + W * 4] = bd.ravel()
        o += W * 4 + W
        p[o:o + W * UW] = rng.standard_normal(W * UW) * np.sqrt(1.0 / UW)
        o += W * UW + W
    return p


def save_weights(p, meta=None, path=WEIGHTS_PATH):
    np.savez(path, params=np.asarray(p, np.float32), meta=np.array(repr(meta or {})))


def load_weights(path=WEIGHTS_PATH):
    """Trained parameters committed with the package (None if absent)."""
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as d:
        return np.array(d["params"], np.float32)


# ---------------------------------------------------------------------------
# training data: seeded sphere / ellipsoid pairs


def random_pair(rng, h):
    """Two bodies, each a sphere or a triaxial ellipsoid with semi-axes
    ~U(0.3, 0.6) and a uniformly random orientation, separated along z by a
    gap ~U(0.01, 0.2) between their supporting planes.  Returns DEL, bodies
    and the solver's surface dicts."""
    DEL = G.delta(h)
    shapes = []
    for _ in range(2):
        if rng.uniform() < 0.5:
            abc = np.full(3, rng.uniform(0.3, 0.6))
        else:
            abc = rng.uniform(0.3, 0.6, 3)
        R = G.random_rotation(rng).reshape(3, 3)
        ez = R @ np.array([0.0, 0.0, 1.0])
        shapes.append((abc, R, float(np.sqrt(np.sum((abc * ez) ** 2)))))
    gap = rng.uniform(0.01, 0.2)
    ctrs = [(0.0, 0.0, -(shapes[0][2] + gap / 2)), (0.0, 0.0, shapes[1][2] + gap / 2)]
    bodies = [G.Body(shapes[k][0], ctrs[k], shapes[k][1].ravel()) for k in range(2)]
    surfs = [b.surface(h) for b in bodies]
    pd = []
    for k in range(2):
        t = bodies[1 - k].nearest(h, DEL, surfs[1 - k], surfs[k]["x"], with_force=False)
        pd.append(dict(surfs[k], ctr=bodies[k].ctr, nrst=t["nrst"], x0=t["x0"], n0=t["n0"], b=t["b"],
                       i123=t["i123"]))
    return DEL, bodies, pd, gap


def random_rhs(rng, bodies, pd):
    """A seeded rigid-body field U_k + W_k x (x - c_k) and a smooth
    spherical-harmonic field, each over both surfaces."""
    rigid, smooth = [], []
    for k in range(2):
        x = pd[k]["x"].reshape(-1, 3)
        U, Wv = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        rigid.append((U + np.cross(Wv, x - bodies[k].ctr)).ravel())
        smooth.append(G.sh_density(pd[k]["x"], bodies[k].ctr, rng.uniform(-1, 1, (3, 16))))
    return [np.concatenate(rigid), np.concatenate(smooth)]


class Sample:
    """One geometry's logged pairs, resident on the device."""

    def __init__(self, ctx, DEL, bodies, pd, gap, z_dev, az_dev, count, n):
        self.DEL, self.bodies, self.pd, self.gap = DEL, bodies, pd, gap
        self.z, self.az, self.count, self.n = z_dev, az_dev, count, n
        self.ctx = ctx

    def free(self):
        self.ctx.free(self.z)
        self.ctx.free(self.az)


def make_dataset(ctx, n_geom, h, seed, m=15, tol=1e-10, maxit=30, log=print):
    """Solves A u = b by flexible GMRES without preconditioner for two seeded
    right-hand sides on each of n_geom seeded geometries (direct FP64
    operator) and keeps every Arnoldi pair (z_j, A z_j) on the device."""
    from .solver import Solver
    rng = np.random.default_rng(seed)
    out = []
    t0 = time.time()
    total = 0
    for g in range(n_geom):
        DEL, bodies, pd, gap = random_pair(rng, h)
        S = Solver(ctx, h, DEL, pd[0], pd[1], method=DIRECT, precision=FP64)
        n = S.n
        rhs = random_rhs(rng, bodies, pd)
        cap = len(rhs) * maxit
        zd, azd = ctx.alloc(8 * n * cap), ctx.alloc(8 * n * cap)
        cnt = 0
        for b in rhs:
            zoff = C.c_void_p(zd.value + 8 * n * cnt)
            aoff = C.c_void_p(azd.value + 8 * n * cnt)
            _, it, _, res, nl = S.fgmres(b, m=m, tol=tol, maxit=maxit, log=(zoff, aoff, cap - cnt))
            cnt += nl
        S.close()
        out.append(Sample(ctx, DEL, bodies, pd, gap, zd, azd, cnt, n))
        total += cnt
        if log and (g + 1) % 50 == 0:
            log(f"dataset: {g + 1}/{n_geom} geometries, {total} pairs, {time.time() - t0:.1f}s")
    return out


def train(ctx, samples, epochs=20, lr=1e-3, seed=0, params=None, val=None, log=print, k=16):
    """Adam on mean |M(A z) - z|^2/|z|^2 over each geometry's pairs (one
    batch per geometry and step).  Returns the network and the history of
    (epoch, train loss, validation loss)."""
    net = Network(ctx, k=k, params=init_params(seed) if params is None else params)
    rng = np.random.default_rng(seed)
    hist = []
    bmax = max(s.count for s in samples + (val or []))

    def val_loss():
        if not val:
            return None
        ls = []
        for s in val:
            net.set_geometry(s.pd, s.bodies, batch_max=bmax)
            ls.append(net.loss_grad_dev(s.count, s.az, s.z))
        return float(np.mean(ls))

    v0 = val_loss()
    hist.append((0, None, v0))
    if log:
        log(f"train: epoch 0 validation loss {v0}")
    for ep in range(1, epochs + 1):
        t0 = time.time()
        ls = []
        for gi in rng.permutation(len(samples)):
            s = samples[gi]
            net.set_geometry(s.pd, s.bodies, batch_max=bmax)
            ls.append(net.loss_grad_dev(s.count, s.az, s.z))
            net.adam(lr)
        vl = val_loss()
        hist.append((ep, float(np.mean(ls)), vl))
        if log:
            log(f"train: epoch {ep} loss {np.mean(ls):.5f} validation {vl} ({time.time() - t0:.1f}s)")
    return net, hist


def scale_output(p, alpha):
    """Parameters whose learned correction is alpha times that of p
    (M_alpha(r) = r + alpha (M(r) - r)); alpha = 0 is the identity."""
    q = np.array(p, np.float32, copy=True)
    off_wout = q.size - 4 - 3 * 64     # W_out (3 x 64) and b_out (3 + pad) close the layout
    q[off_wout:] *= alpha
    return q


def solve(S, b, net=None, m=15, tol=1e-8, maxit=200, gate=True, gate_ratio=0.5, force_plain=False):
    """Flexible GMRES with the learned right preconditioner, plain GMRES as
    the fallback.  The solver's policy (hor_solver_set_precond_policy) makes
    the first Arnoldi step of every restart cycle plain and uses the network
    for the rest of the cycle only if, on that step's pair (v0, A v0), it
    inverts A better than the identity does (gate); a non-finite network
    output is replaced by the plain step and ends the network's use in the
    solve; force_plain ignores the network.  Returns x and a dict with the
    iteration count, the learned and plain steps, the gate decisions and the
    final FP64 residual |b - A x|/|b|."""
    b = f64(b, S.n)
    S.set_precond_policy(force_plain=force_plain, gate=gate, gate_ratio=gate_ratio)
    try:
        x, it, _, res, _ = S.fgmres(b, m=m, tol=tol, maxit=maxit, precond=net)
        info = S.last_info()
    finally:
        S.set_precond_policy()
    info.update(iterations=it, relres=res, fallback=bool(net is not None and info["learned_steps"] == 0))
    return x, info
