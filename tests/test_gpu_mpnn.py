"""GPU tests of the learned preconditioner (hand-written HIP forward, backward
and Adam kernels) against the CPU oracle, and of flexible GMRES with it.

Parity bar: the GPU rounds every matrix-core operand to BF16 and accumulates
in FP32; the oracle rounds the same operands the same way and accumulates in
FP64, so the results agree to FP32 accumulation error plus the rare BF16
rounding flip of an intermediate (norm-wise 1e-3 on the network's output and
on each parameter block's gradient).
"""
import numpy as np
import pytest

import mpnn_oracle as MO
from cases import problem_dicts, sphere_pair
from hor import mpnn as M
from hor.solver import Solver
import hor

pytestmark = pytest.mark.gpu


def _setup(N=8, gap=0.1, seed=5):
    h, DEL, bodies, surfs, tgts = sphere_pair(N, gap=gap)
    pd = problem_dicts(bodies, surfs, tgts)
    p = MO.init_params(seed, scale_out=0.5)
    rng = np.random.default_rng(seed)
    p[MO.OFF_BIN:MO.OFF_BIN + MO.W] = 0.1 * rng.standard_normal(MO.W)
    p[MO.OFF_BOUT:MO.OFF_BOUT + 3] = 0.1 * rng.standard_normal(3)
    for l in range(MO.LAYERS):
        b = MO.OFF_L0 + l * MO.L_SIZE
        p[b + MO.L_CC:b + MO.L_CC + MO.W] = 0.1 * rng.standard_normal(MO.W)
        p[b + MO.L_D:b + MO.L_D + MO.W] = 0.1 * rng.standard_normal(MO.W)
    return h, DEL, bodies, pd, p.astype(np.float32)


def _graph(pd, bodies, nbr):
    x = np.concatenate([pd[0]["x"], pd[1]["x"]])
    nr = np.concatenate([pd[0]["nrml"], pd[1]["nrml"]])
    ctr = np.concatenate([bodies[0].ctr, bodies[1].ctr])
    size = np.array([M.body_size(bodies[0]), M.body_size(bodies[1])])
    return MO.Graph(pd[0]["n"], pd[1]["n"], x, nr, ctr, size, nbr)


def _blocks():
    b = [("W_in", MO.OFF_WIN, MO.W * MO.INPAD), ("b_in", MO.OFF_BIN, MO.W), ("W_out", MO.OFF_WOUT, 3 * MO.W),
         ("b_out", MO.OFF_BOUT, 3)]
    for l in range(MO.LAYERS):
        o = MO.OFF_L0 + l * MO.L_SIZE
        b += [(f"W_a{l}", o + MO.L_WA, MO.W * MO.W), (f"B{l}", o + MO.L_BD, MO.W * 4), (f"c{l}", o + MO.L_CC, MO.W),
              (f"W_u{l}", o + MO.L_WU, MO.W * MO.UW), (f"d{l}", o + MO.L_D, MO.W)]
    return b


def test_param_layout():
    assert M.param_count() == MO.NPARAMS


def test_forward_and_gradient_match_oracle(ctx):
    h, DEL, bodies, pd, p = _setup()
    net = M.Network(ctx, k=16, params=p)
    nbr = net.set_geometry(pd, bodies, batch_max=3)
    npts = pd[0]["n"] + pd[1]["n"]
    Gr = _graph(pd, bodies, nbr)
    # the graph is the brute-force k nearest neighbours
    x = Gr.geo  # noqa: F841
    xs = np.concatenate([pd[0]["x"], pd[1]["x"]]).reshape(-1, 3)
    d2 = ((xs[:, None, :] - xs[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    for i in range(0, npts, 37):
        assert np.array_equal(nbr[i], np.lexsort((np.arange(npts), d2[i]))[:16])
    rng = np.random.default_rng(1)
    r = rng.standard_normal((3, 3 * npts))
    v = rng.standard_normal((3, 3 * npts))
    orc = MO.Net(p, bf=True)
    z_orc, _ = orc.forward(Gr, r)
    z_gpu = net.forward(r)
    e = np.linalg.norm(z_gpu - z_orc) / np.linalg.norm(z_orc - r)
    print("forward: |z_gpu - z_oracle| / |M(r) - r| =", e)
    assert e < 1e-3
    loss_g, g_gpu = net.loss_grad(r, v)
    loss_o, g_orc, _ = orc.loss_grad(Gr, r, v)
    print("loss gpu / oracle:", loss_g, loss_o)
    assert abs(loss_g - loss_o) <= 1e-4 * loss_o
    for name, off, n in _blocks():
        a, b = g_gpu[off:off + n].astype(np.float64), g_orc[off:off + n]
        nb = np.linalg.norm(b)
        err = np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)
        print(f"grad {name}: rel err {err:.2e} (|g| = {nb:.3e})")
        assert err < 1e-3, name
    net.close()


def test_identity_network_reproduces_gmres(ctx):
    h, DEL, bodies, pd, _ = _setup(N=8)
    S = Solver(ctx, h, DEL, pd[0], pd[1])
    b = np.random.default_rng(3).uniform(-1, 1, S.n)
    net = M.Network(ctx, k=16, params=np.zeros(M.param_count(), np.float32))
    net.set_geometry(pd, bodies)
    x0, it0, _, res0 = S.gmres(b, m=15, tol=1e-10, maxit=100)
    x1, it1, _, res1, _ = S.fgmres(b, m=15, tol=1e-10, maxit=100, precond=net)
    x2, it2, _, res2, _ = S.fgmres(b, m=15, tol=1e-10, maxit=100)
    print("iterations gmres / fgmres(identity net) / fgmres(none):", it0, it1, it2, res0, res1, res2)
    assert it1 == it2 and abs(it1 - it0) <= 1
    assert res1 <= 1e-10 and res2 <= 1e-10
    assert np.linalg.norm(x1 - x0) <= 1e-9 * np.linalg.norm(x0)
    r = b - S.matvec(x1)
    assert abs(np.linalg.norm(r) / np.linalg.norm(b) - res1) <= 1e-3 * res1 + 1e-15
    net.close()
    S.close()


def test_training_reduces_loss_and_preconditions(ctx):
    h = 1.0 / 16
    train = M.make_dataset(ctx, 6, h, seed=101, log=None)
    val = M.make_dataset(ctx, 2, h, seed=202, log=None)
    assert sum(s.count for s in train) >= 6 * 2 * 5
    net, hist = M.train(ctx, train, epochs=12, lr=2e-3, seed=0, val=val, log=print)
    print("history:", hist)
    # the untrained network is the identity M(r) = r; training must lower the
    # held-out loss |M(A z) - z|^2/|z|^2 below that of the identity
    assert hist[-1][2] < 0.97 * hist[0][2]
    # flexible GMRES with the trained network on a held-out geometry
    rng = np.random.default_rng(7)
    DEL, bodies, pd, gap = M.random_pair(rng, h)
    S = Solver(ctx, h, DEL, pd[0], pd[1])
    net.set_geometry(pd, bodies)
    b = M.random_rhs(rng, bodies, pd)[1]
    _, it0, _, res0, _ = S.fgmres(b, m=15, tol=1e-10, maxit=100)
    x, it1, _, res1, _ = S.fgmres(b, m=15, tol=1e-10, maxit=100, precond=net)
    print(f"held-out geometry: plain {it0} iterations, learned {it1} iterations (FP64 residual {res1:.2e})")
    assert res1 <= 1e-10
    assert np.linalg.norm(b - S.matvec(x)) / np.linalg.norm(b) <= 1.01e-10
    for s in train + val:
        s.free()
    net.close()
    S.close()


def test_nonfinite_network_output_falls_back(ctx):
    """Parameters that overflow FP32 in the hidden layers make the network
    output non-finite; the solver screens every output, takes the plain step
    instead, stops using the network, and ends on the FP64 residual with the
    plain GMRES iteration count."""
    h, DEL, bodies, pd, _ = _setup(N=8)
    S = Solver(ctx, h, DEL, pd[0], pd[1])
    b = np.random.default_rng(9).uniform(-1, 1, S.n)
    x0, st0 = M.solve(S, b, net=None, tol=1e-10, maxit=100)
    net = M.Network(ctx, k=16, params=np.full(M.param_count(), 1e30, np.float32))
    net.set_geometry(pd, bodies)
    z = net.forward(b)
    assert not np.all(np.isfinite(z))          # the injected failure
    for gate in (False, True):
        x1, st1 = M.solve(S, b, net=net, tol=1e-10, maxit=100, gate=gate)
        print(f"gate={gate}: plain {st0['iterations']} iterations; with the overflowing network "
              f"{st1['iterations']} ({st1['nonfinite_outputs']} non-finite outputs, {st1['learned_steps']} learned "
              f"steps), FP64 residual {st1['relres']:.2e}")
        assert st1["nonfinite_outputs"] == 1 and st1["learned_steps"] == 0
        assert st1["iterations"] == st0["iterations"]
        assert st1["relres"] <= 1e-10
        assert np.linalg.norm(b - S.matvec(x1)) / np.linalg.norm(b) <= 1.01e-10
        assert np.linalg.norm(x1 - x0) <= 1e-12 * np.linalg.norm(x0)
    net.close()
    S.close()


def test_forced_plain_policy(ctx):
    """force_plain ignores the network: the solve is plain GMRES exactly."""
    h, DEL, bodies, pd, p = _setup(N=8)
    S = Solver(ctx, h, DEL, pd[0], pd[1])
    b = np.random.default_rng(10).uniform(-1, 1, S.n)
    x0, st0 = M.solve(S, b, net=None, tol=1e-10, maxit=100)
    net = M.Network(ctx, k=16, params=p)
    net.set_geometry(pd, bodies)
    x1, st1 = M.solve(S, b, net=net, tol=1e-10, maxit=100, force_plain=True)
    assert st1["learned_steps"] == 0 and st1["iterations"] == st0["iterations"]
    assert np.array_equal(x1, x0)
    x2, st2 = M.solve(S, b, net=net, tol=1e-10, maxit=100, gate=False)
    print(f"plain {st0['iterations']}, forced plain {st1['iterations']}, ungated network {st2['iterations']} "
          f"({st2['learned_steps']} learned steps)")
    assert st2["learned_steps"] == st2["iterations"] and st2["relres"] <= 1e-10
    net.close()
    S.close()
