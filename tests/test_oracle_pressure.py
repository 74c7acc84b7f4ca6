"""CPU pins of the oracle's regularized pressure (stokes_pressure_7ord,
stokes_pressure_away, Lapl_DL_7ord; highreg_stokes_extend_to_grid/stokes.cpp:99-203,
utilities.cpp:190-211).

* against the reference's own outputs (golden fixture written from the
  compiled reference: band nodes of the grid program at h = 1/16, inside
  nodes included, and two box faces);
* Lapl_DL_7ord equals the s2 kernel of Stokes_SL_7ord (the same function
  written twice in the reference), itself pinned to 50-digit closed forms;
* the translating spheroid's exact pressure (Chwang & Wu, p_exact): the
  regularized sums converge to it at band nodes outside the body.
"""
import numpy as np

from conftest import golden
from hor import geometry as G


def _grid_surface():
    g = golden("grid_h16.npz")
    s = dict(n=len(g["s_area"]), x=g["s_x"], nrml=g["s_nrml"], area=g["s_area"], f=g["f"])
    return g, s


def test_pressure_matches_reference_fixture(oracle):
    g, s = _grid_surface()
    gp = golden("grid_pressure_h16.npz")
    assert int(gp["N"]) == int(g["N"]) and float(gp["h"]) == float(g["h"])
    h = float(gp["h"])
    tgt = dict(n=len(gp["b"]), x=(gp["ijk"] * h).astype(np.float64).ravel(), b=gp["b"], n0=gp["n0"],
               f0=gp["f0"])
    p = oracle.pressure_off(float(gp["DEL"]), s, tgt)
    ref = gp["p_band"]
    err = np.abs(p - ref).max() / np.abs(ref).max()
    print("band nodes:", tgt["n"], "inside:", int((gp["b"] < 0).sum()), "max rel diff", err)
    assert (gp["b"] < -1e-12).sum() > 100   # the inside jump is exercised
    assert err < 1e-13
    pa = oracle.pressure_away(s, gp["face"])
    assert np.abs(pa - gp["p_face"]).max() / np.abs(gp["p_face"]).max() < 1e-13


def test_lapl_dl_kernel_is_sl_s2(oracle):
    r = np.concatenate([[0.0, 5e-15], np.linspace(1e-3, 1.0, 300)])
    for d, lam in ((0.1, 0.37), (0.05, -1.3), (0.2, 0.0)):
        c = oracle.reg_coeff(lam)
        a = oracle.kernel("lapl_dl", r, d, c)[:, 1]
        b = oracle.kernel("sl", r, d, c)[:, 1]
        assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()


def test_pressure_converges_to_exact_translating_spheroid(oracle):
    """Outside band nodes (0 < b <= 2h) of the reference spheroid (1,.5,.5)
    in the grid program's box: the error against p_exact falls with h."""
    errs = []
    for N in (16, 32):
        h = 1.0 / N
        DEL = G.delta(h)
        body = G.Body((1.0, 0.5, 0.5), (1.5, 1.5, 1.5))
        s = body.surface(h)
        s["f"] = body.traction(s["x"])
        t = body.band_targets(s, (0.0, 0.0, 0.0), h, 3 * N, 2 * h)
        keep = t["b"] > 0
        tgt = dict(n=int(keep.sum()), x=t["x"].reshape(-1, 3)[keep].ravel(), b=t["b"][keep],
                   n0=t["n0"].reshape(-1, 3)[keep].ravel())
        tgt["f0"] = body.traction(t["x0"].reshape(-1, 3)[keep].ravel())
        p = oracle.pressure_off(DEL, s, tgt)
        pe = _p_exact(tgt["x"], (1.5, 1.5, 1.5))
        errs.append(np.abs(p - pe).max() / np.abs(pe).max())
    print("max rel error vs exact, h = 1/16, 1/32:", errs)
    assert errs[0] < 0.05 and errs[1] < errs[0] / 3


def _p_exact(x, c, a=1.0, b=0.5):
    """Exact pressure of a prolate spheroid (a > b = c) translating with
    U = (1, 0, 0): p = -2 alpha1 (1/R1 - 1/R2), R1,2 the distances to the
    foci (x = -+ c_f), alpha1 = e^2 / ((1 + e^2) L - 2e), L = log((1+e)/(1-e))
    (Chwang & Wu 1975, eq. 29a)."""
    x = np.asarray(x).reshape(-1, 3) - np.asarray(c)
    cf = np.sqrt(a * a - b * b)
    r2 = x[:, 1] ** 2 + x[:, 2] ** 2
    R1 = np.sqrt((x[:, 0] + cf) ** 2 + r2)
    R2 = np.sqrt((x[:, 0] - cf) ** 2 + r2)
    e = cf / a
    L = np.log((1 + e) / (1 - e))
    a1 = e * e / ((1 + e * e) * L - 2 * e)
    return -2.0 * a1 * (1.0 / R1 - 1.0 / R2)


def test_exact_pressure_matches_reference(refgrid):
    g = golden("grid_pressure_h16.npz")
    x = g["face"]
    assert np.abs(_p_exact(x, (1.5, 1.5, 1.5)) - g["p_face_exact"]).max() <= 1e-14
    assert np.abs(refgrid.p_exact(x) - g["p_face_exact"]).max() == 0.0
