"""GPU geometry (quadrature, nearest-point data, Monge stencils) against the
host geometry, the reference's fixture and the oracle.

Parity bar: bit-exact for everything built from +, -, *, /, sqrt and
comparisons (points, normals, planes, nearest indices, nearest points, signed
distances, Monge patches, surface force, stencil indices and weights).  The
quadrature weight h^2 PoU/|n_i| takes exp and acos from the device math
library; it agrees to 4e-15 h^2 absolute (weights reach h^2/cos 70deg = 2.9 h^2,
whose ulp is 8.9e-16 h^2).  (Near the edge of a partition of
unity the bump's slope multiplies an ulp of acos by ~1e3, so weights below
1e-20 h^2 differ by up to 1e-12 relative.)
"""
import time

import numpy as np
import pytest

from conftest import golden
from hor import geometry as G

pytestmark = pytest.mark.gpu

AREA_TOL = 4e-15   # times h^2

BODIES = {
    "spheroid h=1/32": ((1.0, 0.5, 0.5), (0.0, 0.0, -0.5), None, 1 / 32),
    "spheroid rotated 30 deg h=1/32": ((1.0, 0.5, 0.5), (0.0, 0.0, 0.5), G.rot_z(30.0), 1 / 32),
    "ellipsoid random rotation h=1/48": ((1.0, 0.8, 0.6), (0.1, -0.2, 0.3),
                                         G.random_rotation(np.random.default_rng(4)), 1 / 48),
    "sphere off-centre h=1/64": ((0.5, 0.5, 0.5), (1.5, 1.5, 1.5), None, 1 / 64),
}


def _assert_surface_equal(a, b):
    assert a["n"] == b["n"]
    assert np.array_equal(a["plane"], b["plane"])
    assert np.array_equal(a["x"], b["x"])
    assert np.array_equal(a["nrml"], b["nrml"])
    err = np.abs(a["area"] - b["area"]).max() / a["h"] ** 2
    assert err <= AREA_TOL, err
    return err


def _assert_nearest_equal(a, b):
    for k in ("nrst", "i123", "x0", "n0", "b", "f0"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", list(BODIES))
def test_surface_matches_host(ctx, name):
    abc, ctr, R, h = BODIES[name]
    body = G.Body(abc, ctr, R)
    host, gpu = body.surface(h), body.surface(h, ctx=ctx)
    e = _assert_surface_equal(gpu, host)
    print(f"{name}: {gpu['n']} points, max |area diff| / h^2 = {e:.1e}")


def test_surface_matches_reference_fixture(ctx):
    """The reference's spheroid pair at h = 1/16 (Generate_Surface output)."""
    g = golden("two_surface_h16.npz")
    h = float(g["h"])
    b1 = G.Body((1.0, 0.5, 0.5), g["ctr1"])
    b2 = G.Body((1.0, 0.5, 0.5), g["ctr2"], G.rot_z(30.0))
    for k, b in ((1, b1), (2, b2)):
        s = b.surface(h, ctx=ctx)
        assert np.array_equal(s["x"], g[f"s{k}_x"])
        assert np.array_equal(s["nrml"], g[f"s{k}_nrml"])
        assert np.array_equal(s["plane"], g[f"s{k}_plane"])
        assert np.abs(s["area"] - g[f"s{k}_area"]).max() <= AREA_TOL * h * h


def test_nearest_matches_reference_fixture(ctx):
    g = golden("two_surface_h16.npz")
    h, DEL = float(g["h"]), float(g["DEL"])
    b1 = G.Body((1.0, 0.5, 0.5), g["ctr1"])
    b2 = G.Body((1.0, 0.5, 0.5), g["ctr2"], G.rot_z(30.0))
    s1, s2 = b1.surface(h), b2.surface(h)
    t2 = b1.nearest(h, DEL, s1, s2["x"], ctx=ctx)
    t1 = b2.nearest(h, DEL, s2, s1["x"], ctx=ctx)
    for k, t in ((1, t1), (2, t2)):
        for key in ("nrst", "i123", "x0", "b", "f0"):
            assert np.array_equal(t[key], g[f"t{k}_{key}"]), (k, key)


@pytest.mark.parametrize("name", list(BODIES))
def test_nearest_matches_host_scattered_targets(ctx, name):
    """Targets scattered around the body: radial factors 0.8 .. 2.5 of the
    quadrature points (inside, near and far) plus jitter."""
    abc, ctr, R, h = BODIES[name]
    body = G.Body(abc, ctr, R)
    s = body.surface(h)
    DEL = G.delta(h)
    rng = np.random.default_rng(17)
    pts = s["x"].reshape(-1, 3)
    idx = rng.choice(len(pts), min(4000, len(pts)), replace=False)
    fac = rng.uniform(0.8, 2.5, len(idx))[:, None]
    tx = body.ctr + fac * (pts[idx] - body.ctr) + rng.normal(0, h, (len(idx), 3))
    host = body.nearest(h, DEL, s, tx.ravel())
    gpu = body.nearest(h, DEL, s, tx.ravel(), ctx=ctx)
    _assert_nearest_equal(gpu, host)
    near = (gpu["nrst"] != 0).sum()
    print(f"{name}: {len(idx)} targets, {near} near")
    assert 0 < near < len(idx)
    # without the force
    g2 = body.nearest(h, DEL, s, tx.ravel(), with_force=False, ctx=ctx)
    for k in ("nrst", "i123", "x0", "n0", "b"):
        assert np.array_equal(g2[k], host[k])


def test_nearest_tie_takes_first_index(ctx):
    """A quadrature point duplicated at the end of the list ties with the
    original: the lower index wins (find_Nearest_grid, surface.cpp:370-390)."""
    body = G.Body((0.5, 0.5, 0.5))
    h = 1 / 16
    s = body.surface(h)
    x = s["x"].reshape(-1, 3)
    DEL = G.delta(h)
    # plane-2 points whose Monge patch (largest normal component) is z
    pick = [i for i in range(s["n"])
            if s["plane"][i] == 2 and x[i, 2] > 1.2 * max(abs(x[i, 0]), abs(x[i, 1]))][5:12]
    assert len(pick) == 7
    for i in pick:
        s2 = dict(s, n=s["n"] + 1, x=np.concatenate([s["x"], x[i]]), plane=np.append(s["plane"], 2))
        t = x[i] * 1.02
        host = body.nearest(h, DEL, s2, t)
        gpu = body.nearest(h, DEL, s2, t, ctx=ctx)
        _assert_nearest_equal(gpu, host)
        assert gpu["nrst"][0] == i


def test_stencils_match_oracle(ctx, oracle):
    """Stencil indices equal the oracle's (the reference's scan); the weights
    interpolate g0 as the oracle does."""
    h = 1 / 32
    DEL = G.delta(h)
    b1 = G.Body((1.0, 0.5, 0.5), (0.0, 0.0, -0.5 - 1.0 / 4096))
    b2 = G.Body((1.0, 0.5, 0.5), (0.0, 0.0, 0.5), G.rot_z(30.0))
    s1, s2 = b1.surface(h), b2.surface(h)
    t2 = b1.nearest(h, DEL, s1, s2["x"])   # points of surface 2, nearest data on surface 1
    t1 = b2.nearest(h, DEL, s2, s1["x"])
    rng = np.random.default_rng(5)
    for src, tgt in ((s1, t2), (s2, t1)):
        st, co = G.stencils(ctx, src, tgt, h)
        ref = oracle.stencil(src["ctr"], h, src["x"], tgt)
        assert np.array_equal(st, ref)
        near = tgt["nrst"] != 0
        assert near.sum() > 100
        assert np.all(st[~near] == 0) and np.all(co[~near] == 0)
        g = rng.uniform(-1, 1, 3 * src["n"])
        G3 = g.reshape(-1, 3)
        a, b = co[:, :6], co[:, 6:]
        vals = G3[st.reshape(-1, 6, 6)]                   # [t, k1, k2, comp]
        g0 = np.einsum("tj,tkjc,tk->tc", b, vals, a)
        g0[~near] = 0
        want = oracle.interp_g0(src["ctr"], h, src["x"], g, tgt).reshape(-1, 3)
        assert np.abs(g0 - want).max() <= 1e-13 * np.abs(want).max()


def test_two_spheres_h256_full_size(ctx, oracle):
    """The GMRES workload's geometry at h = 1/256 (two spheres of radius 0.5,
    gap 0.1): GPU quadrature and nearest data equal the host's; stencils of
    600 sampled near targets equal the oracle's."""
    h = 1 / 256
    DEL = G.delta(h)
    b1 = G.Body((0.5, 0.5, 0.5), (0.0, 0.0, -0.55))
    b2 = G.Body((0.5, 0.5, 0.5), (0.0, 0.0, 0.55))
    t0 = time.perf_counter()
    s1g, s2g = b1.surface(h, ctx=ctx), b2.surface(h, ctx=ctx)
    t1g = b2.nearest(h, DEL, s2g, s1g["x"], ctx=ctx)
    t2g = b1.nearest(h, DEL, s1g, s2g["x"], ctx=ctx)
    tg = time.perf_counter() - t0
    t0 = time.perf_counter()
    s1, s2 = b1.surface(h), b2.surface(h)
    t1 = b2.nearest(h, DEL, s2, s1["x"])
    t2 = b1.nearest(h, DEL, s1, s2["x"])
    th = time.perf_counter() - t0
    print(f"h=1/256: {s1['n']} + {s2['n']} points; GPU {tg:.3f} s, host {th:.3f} s")
    _assert_surface_equal(s1g, s1)
    _assert_surface_equal(s2g, s2)
    _assert_nearest_equal(t1g, t1)
    _assert_nearest_equal(t2g, t2)
    st, co = G.stencils(ctx, s2, t1, h)
    near = np.where(t1["nrst"] != 0)[0]
    assert len(near) > 1000
    pick = np.sort(np.random.default_rng(3).choice(near, 600, replace=False))
    sub = dict(n=len(pick), nrst=t1["nrst"][pick], x0=t1["x0"].reshape(-1, 3)[pick].ravel(),
               i123=t1["i123"].reshape(-1, 3)[pick].ravel())
    ref = oracle.stencil(s2["ctr"], h, s2["x"], sub)
    assert np.array_equal(st[pick], ref)
