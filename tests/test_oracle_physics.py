"""Physical and mathematical identities that pin the oracle independently of
the reference build (CPU).

* SL of the exact traction of a translating spheroid equals U = (1,0,0) on and
  inside the body and the exact exterior velocity (Chwang & Wu) outside, with
  errors falling as h is refined (the method's convergence);
* the double layer of a rigid motion u = U + W x (y - c) is u inside, u/2 as a
  principal value on the surface and 0 outside (Lorentz reciprocal identity),
  which pins the stresslet normalization, sign, jump and subtraction;
* the two-surface operator maps a rigid motion of each surface to 2/3 of it;
* barycentric proxy weights reproduce polynomials of degree <= P exactly;
* GMRES agrees with a dense direct solve of the assembled operator.
"""
import numpy as np
import pytest

from cases import problem_dicts, sphere_pair
from hor import geometry as G

U = np.array([0.3, -0.2, 0.5])
W = np.array([0.1, 0.4, -0.3])


def _case(N, max_targets=6000, seed=4):
    h = 1.0 / N
    DEL = G.delta(h)
    body = G.Body((1.0, 0.5, 0.5), (0.1, 0.2, 0.3))
    s = body.surface(h)
    t = body.band_targets(s, (-1.4, -1.3, -1.2), h, int(round(3 / h)), 3 * DEL)
    if t["n"] > max_targets:
        idx = np.sort(np.random.default_rng(seed).choice(t["n"], max_targets, replace=False))
        t = {k: (v.reshape(t["n"], -1)[idx].ravel() if k in ("x", "x0", "n0", "b") else v) for k, v in t.items()}
        t["n"] = len(idx)
    return body, s, t, DEL


@pytest.fixture(scope="module")
def cases():
    return {N: _case(N) for N in (8, 16, 32)}


def test_single_layer_of_exact_traction(oracle, cases):
    errs = []
    for N in (8, 16, 32):
        body, s, t, DEL = cases[N]
        s = dict(s, f=body.traction(s["x"]))
        t = dict(t, f0=body.traction(t["x0"]))
        on = oracle.sl_on(DEL, s).reshape(-1, 3)
        off = oracle.sl_off(DEL, s, t).reshape(-1, 3)
        inside, outside = t["b"] < -1e-12, t["b"] > 1e-12
        ex = body.spheroid_exact(t["x"]).reshape(-1, 3)
        errs.append(max(np.abs(on - [1, 0, 0]).max(), np.abs(off[inside] - [1, 0, 0]).max(),
                        np.abs(off[outside] - ex[outside]).max()))
    print("SL errors vs exact, h = 1/8, 1/16, 1/32:", errs)
    assert errs[2] < 1e-3 and errs[2] < errs[1] / 4 < errs[0] / 8


def test_double_layer_of_rigid_motion(oracle, cases):
    errs = []
    for N in (8, 16, 32):
        body, s, t, DEL = cases[N]
        xs = s["x"].reshape(-1, 3)
        g = (U + np.cross(W, xs - body.ctr)).ravel()
        s = dict(s, g=g)
        t = dict(t, g0=(U + np.cross(W, t["x0"].reshape(-1, 3) - body.ctr)).ravel())
        on = oracle.dl_on(DEL, s).reshape(-1, 3)
        off = oracle.dl_off(DEL, s, t).reshape(-1, 3)
        tx = t["x"].reshape(-1, 3)
        inside, outside = t["b"] < -1e-12, t["b"] > 1e-12
        e_on = np.abs(on - 0.5 * g.reshape(-1, 3)).max()
        e_in = np.abs(off[inside] - (U + np.cross(W, tx[inside] - body.ctr))).max()
        e_out = np.abs(off[outside]).max()
        errs.append(max(e_on, e_in, e_out))
    print("DL rigid-motion identity errors:", errs)
    assert errs[2] < 1e-3 and errs[2] < errs[0] / 10


def test_matvec_of_rigid_motions(oracle):
    errs = []
    for N in (8, 16):
        h, DEL, bodies, surfs, tgts = sphere_pair(N, gap=0.1)
        pd = problem_dicts(bodies, surfs, tgts)
        pb = oracle.problem(h, DEL, pd)
        us = [((k + 1) * U + np.cross(W, pd[k]["x"].reshape(-1, 3) - bodies[k].ctr)).ravel() for k in range(2)]
        u = np.concatenate(us)
        errs.append(np.abs(oracle.matvec(pb, u) - 2.0 / 3.0 * u).max() / np.abs(u).max())
    print("A(rigid) - 2/3 rigid:", errs)
    assert errs[1] < 1e-4 and errs[1] < errs[0] / 10


@pytest.mark.parametrize("P", [3, 6])
def test_barycentric_moments_reproduce_polynomials(oracle, P):
    body, s, t, DEL = _case(16, max_targets=10)
    tp = oracle.tree_params(P, 200, 0.6, body.ctr)
    rng = np.random.default_rng(9)
    dens = rng.uniform(-1, 1, 3 * s["n"])
    M = oracle.tree_moments(tp, False, s, dens)             # (nodes, 6, (P+1)^3)
    st = oracle.tree_structure(tp, s["x"])
    x = s["x"].reshape(-1, 3)
    coef = rng.uniform(-1, 1, (P + 1, P + 1, P + 1))
    np1 = P + 1
    for q in range(1, len(M)):
        lo, hi = st["members"][q]
        iv = st["interval"][q]
        tcheb = np.cos(np.arange(np1) * np.pi / P)
        t1 = iv[0] + (tcheb + 1) / 2 * (iv[1] - iv[0])
        t2 = iv[2] + (tcheb + 1) / 2 * (iv[3] - iv[2])
        t3 = iv[4] + (tcheb + 1) / 2 * (iv[5] - iv[4])
        c = st["centre"][q]

        def poly(px, py, pz):
            # a polynomial of degree <= P in each variable (tensor Chebyshev basis)
            def cheb(v, lo_, hi_):
                z = (2 * v - lo_ - hi_) / (hi_ - lo_)
                return np.polynomial.chebyshev.chebvander(z, P)
            Vx, Vy, Vz = cheb(px, iv[0], iv[1]), cheb(py, iv[2], iv[3]), cheb(pz, iv[4], iv[5])
            return np.einsum("ni,nj,nk,ijk->n", Vx, Vy, Vz, coef)

        gx, gy, gz = np.meshgrid(t1, t2, t3, indexing="ij")
        pv = poly(gx.ravel(), gy.ravel(), gz.ravel())
        members = st["perm"][lo:hi + 1]
        # the reference snaps a coordinate within 1e-7 of a node onto it
        # (KITC.cpp:412-456); points on grid lines sit 5e-9 from the nodes of
        # the 1e-8-padded boxes, so the snapped coordinates enter the identity
        ys = []
        for d, tn in enumerate((t1, t2, t3)):
            y = x[members, d].copy()
            close = np.abs(y[:, None] - tn[None, :]) < 1e-7
            hit = close.any(axis=1)
            y[hit] = tn[np.argmax(close[hit], axis=1)]
            ys.append(y)
        pm = poly(*ys)
        for comp in range(3):
            lhs = M[q, comp] @ pv
            rhs = np.sum(s["area"][members] * dens.reshape(-1, 3)[members, comp] * pm)
            assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(rhs)), (q, comp, lhs, rhs)
        del c


def test_gmres_matches_dense_solve(oracle):
    h, DEL, bodies, surfs, tgts = sphere_pair(4, gap=0.2)
    pd = problem_dicts(bodies, surfs, tgts)
    pb = oracle.problem(h, DEL, pd)
    n = 3 * (pd[0]["n"] + pd[1]["n"])
    A = np.column_stack([oracle.matvec(pb, e) for e in np.eye(n)])
    b = np.random.default_rng(2).uniform(-1, 1, n)
    x_dense = np.linalg.solve(A, b)
    x, it, hist, res = oracle.gmres(pb, b, m=15, tol=1e-12, iter_max=300, fix_restart=True)
    assert np.linalg.norm(x - x_dense) / np.linalg.norm(x_dense) < 1e-9
    assert np.all(np.diff(hist[:15]) <= 1e-15)   # residual estimates are non-increasing within a cycle
