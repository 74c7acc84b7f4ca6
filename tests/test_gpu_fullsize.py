"""The headline workload at its full size: ellipsoid (1, .8, .6), h = 1/64,
42,342 sources, 1,314,272 band nodes (|b| <= 0.3), FP64 direct SL_OFF,
DL_OFF and the pressure over every node.

* parity: 1,500 nodes sampled at random (seeded) against the oracle,
  <= 1e-12 relative (2-norm and max-norm);
* size-independent properties over all nodes: linearity in the density
  (2 f1 - 3 f2 against 2 SL[f1] - 3 SL[f2], 1e-12), and for the translating
  body's traction the single layer is U = (1,0,0) at every node inside the
  body (max error 1e-6 at this h) and Oberbeck's exact flow at every node
  outside (max error 1e-4 at this h).
"""
import numpy as np
import pytest

import hor
from conftest import rel_l2, rel_max
from hor import geometry as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def band64():
    h = 1.0 / 64
    body = G.Body((1.0, 0.8, 0.6))
    s = body.surface(h)
    DEL = G.delta(h)
    t = body.band_targets(s, (-1.5, -1.5, -1.5), h, 192, 0.3)
    s["f"] = body.traction(s["x"])
    s["g"] = G.sh_density(s["x"], body.ctr, G.sh_coefficients(20261204))
    t["f0"] = body.traction(t["x0"])
    t["g0"] = G.sh_density(t["x0"], body.ctr, G.sh_coefficients(20261204))
    assert s["n"] == 42342 and t["n"] == 1314272
    return body, s, t, DEL


def _sample(t, idx):
    return dict(n=len(idx), x=t["x"].reshape(-1, 3)[idx].ravel(), b=t["b"][idx],
                x0=t["x0"].reshape(-1, 3)[idx].ravel(), n0=t["n0"].reshape(-1, 3)[idx].ravel(),
                f0=t["f0"].reshape(-1, 3)[idx].ravel(), g0=t["g0"].reshape(-1, 3)[idx].ravel())


def test_full_size_sample_parity(ctx, oracle, band64):
    body, s, t, DEL = band64
    idx = np.sort(np.random.default_rng(64).choice(t["n"], 1500, replace=False))
    sub = _sample(t, idx)
    print("full-size band", s["n"], t["n"], flush=True)
    for kind, dens, d0, orc in ((hor.SL_OFF, s["f"], t["f0"], oracle.sl_off),
                                (hor.DL_OFF, s["g"], t["g0"], oracle.dl_off)):
        P = hor.Plan(ctx, kind, s, t, DEL)
        g = P(dens, d0).reshape(-1, 3)[idx].ravel()
        P.close()
        o = orc(DEL, s, sub)
        assert rel_l2(g, o) <= 1e-12 and rel_max(g, o) <= 1e-12, kind
    P = hor.Plan(ctx, hor.P_OFF, s, t, DEL)
    g = P(s["f"], t["f0"])[idx]
    P.close()
    o = oracle.pressure_off(DEL, s, sub)
    assert rel_l2(g, o) <= 1e-12 and rel_max(g, o) <= 1e-12


def test_full_size_linearity_and_rigid_motion(ctx, band64):
    body, s, t, DEL = band64
    P = hor.Plan(ctx, hor.SL_OFF, s, t, DEL)
    f1, f01 = s["f"], t["f0"]
    f2 = G.sh_density(s["x"], body.ctr, G.sh_coefficients(7))
    f02 = G.sh_density(t["x0"], body.ctr, G.sh_coefficients(7))
    u1, u2 = P(f1, f01), P(f2, f02)
    u3 = P(2 * f1 - 3 * f2, 2 * f01 - 3 * f02)
    assert rel_max(u3, 2 * u1 - 3 * u2) <= 1e-12
    inside, outside = t["b"] < 0, t["b"] > 0
    assert inside.sum() > 100000 and outside.sum() > 100000
    err = np.abs(u1.reshape(-1, 3)[inside] - np.array([1.0, 0.0, 0.0])).max()
    print("SL of the translating traction inside the body: max |u - U| =", err)
    assert err < 1e-6
    # outside: Oberbeck's exact flow of the translating ellipsoid
    ex = body.exact_velocity(t["x"]).reshape(-1, 3)
    e_out = np.abs(u1.reshape(-1, 3)[outside] - ex[outside]).max()
    print(f"outside the body ({outside.sum()} nodes, |b| <= 0.3): max |u - exact| = {e_out:.3e}")
    assert e_out < 1e-4
    P.close()


def test_full_size_fused_matches_separate(ctx, band64):
    """The fused Stokeslet + stresslet operator over all 1,314,272 nodes
    against the separate SL_OFF and DL_OFF plans (1e-15 relative)."""
    body, s, t, DEL = band64
    n = t["n"]
    P = hor.Plan(ctx, hor.SLDL_OFF, s, t, DEL)
    out = P(np.concatenate([s["f"], s["g"]]), np.concatenate([t["f0"], t["g0"]]))
    P.close()
    for half, kind, dens, d0 in ((out[:3 * n], hor.SL_OFF, s["f"], t["f0"]),
                                 (out[3 * n:], hor.DL_OFF, s["g"], t["g0"])):
        P = hor.Plan(ctx, kind, s, t, DEL)
        ref = P(dens, d0)
        P.close()
        print("fused vs separate", kind, rel_max(half, ref), flush=True)
        assert rel_l2(half, ref) <= 1e-15 and rel_max(half, ref) <= 1e-15


def test_full_size_mixed_precision(ctx, oracle, band64):
    """The mixed-precision fused pass over all 1,314,272 nodes: within 1e-6 of
    the FP64 pass (2-norm), and its error against the oracle on the seeded
    1,500-node sample stated."""
    body, s, t, DEL = band64
    n = t["n"]
    fg, fg0 = np.concatenate([s["f"], s["g"]]), np.concatenate([t["f0"], t["g0"]])
    P64 = hor.Plan(ctx, hor.SLDL_OFF, s, t, DEL)
    out64 = P64(fg, fg0)
    P64.close()
    P32 = hor.Plan(ctx, hor.SLDL_OFF, s, t, DEL, precision=hor.MIXED)
    out32 = P32(fg, fg0)
    P32.close()
    idx = np.sort(np.random.default_rng(64).choice(n, 1500, replace=False))
    sub = _sample(t, idx)
    for name, half, orc in (("SL", slice(0, 3 * n), oracle.sl_off(DEL, s, sub)),
                            ("DL", slice(3 * n, 6 * n), oracle.dl_off(DEL, s, sub))):
        e64 = rel_l2(out32[half], out64[half])
        eo = rel_l2(out32[half].reshape(-1, 3)[idx].ravel(), orc)
        print(f"mixed {name}: vs FP64 over all nodes {e64:.2e}, vs the oracle on 1500 nodes {eo:.2e}", flush=True)
        assert e64 <= 1e-6 and eo <= 1e-6
