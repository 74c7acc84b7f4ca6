"""GPU parity of the direct regularized sums against the CPU oracle.

Tolerance (FP64 path): ||gpu - oracle||_2 / ||oracle||_2 <= 1e-12 and
max|gpu - oracle| / max|oracle| <= 1e-12.  Mixed path (FP32 far pairs, FP64
accumulation): <= 1e-6 against the FP64 path.
"""
import numpy as np
import pytest

import hor
from cases import ellipsoid_case, spheroid_pair
from conftest import rel_l2, rel_max

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def ell():
    return ellipsoid_case(h=1 / 16, max_band=3000)


def _check(a, b, tol=TOL):
    assert np.all(np.isfinite(a))
    assert rel_l2(a, b) <= tol, rel_l2(a, b)
    assert rel_max(a, b) <= tol, rel_max(a, b)


def test_sl_offsurf_band(ctx, oracle, ell):
    body, s, t, DEL = ell
    print("SL_OFF band", s["n"], t["n"], flush=True)
    g = hor.ops.eval_Stokes_SL_offSurf(ctx, 0.0, DEL, s, t)
    _check(g, oracle.sl_off(DEL, s, t))


def test_dl_offsurf_band(ctx, oracle, ell):
    body, s, t, DEL = ell
    print("DL_OFF band", s["n"], t["n"], flush=True)
    g = hor.ops.eval_Stokes_DL_offSurf(ctx, 0.0, DEL, s, t)
    _check(g, oracle.dl_off(DEL, s, t))


def test_sl_onsurf(ctx, oracle, ell):
    body, s, t, DEL = ell
    g = hor.ops.eval_Stokes_SL_onSurf(ctx, DEL, s)
    _check(g, oracle.sl_on(DEL, s))


def test_dl_onsurf(ctx, oracle, ell):
    body, s, t, DEL = ell
    g = hor.ops.eval_Stokes_DL_onSurf(ctx, DEL, s)
    _check(g, oracle.dl_on(DEL, s))


def test_two_surface_offsurf(ctx, oracle):
    """The reference's spheroid pair: far targets (|b|/delta > 8, nrst = 0) and
    near targets in one call."""
    h, DEL, bodies, (s1, s2), (t1, t2) = spheroid_pair(N=32)
    rng = np.random.default_rng(7)
    s1["g"] = rng.uniform(-1, 1, 3 * s1["n"])
    t2["g0"] = np.where(np.repeat(t2["nrst"] != 0, 3), rng.uniform(-1, 1, 3 * t2["n"]), 0.0)
    assert (t2["nrst"] == 0).any() and (t2["nrst"] != 0).any()
    _check(hor.ops.eval_Stokes_SL_offSurf(ctx, h, DEL, s1, t2), oracle.sl_off(DEL, s1, t2))
    _check(hor.ops.eval_Stokes_DL_offSurf(ctx, h, DEL, s1, t2), oracle.dl_off(DEL, s1, t2))


def test_ragged_sizes_and_coincident_targets(ctx, oracle):
    """Sizes that are not multiples of the tile/wave widths; off-surface targets
    placed exactly on sources (r = 0 self limit) with b = 0, and inside (b < 0)."""
    body, s, t, DEL = ellipsoid_case(h=1 / 8, max_band=300)
    for ns in (1, 63, 65, 257):
        ss = {k: (v[: 3 * ns] if k in ("x", "nrml", "f", "g") else v[:ns] if k in ("area", "plane") else v)
              for k, v in s.items()}
        ss["n"] = ns
        for nt in (1, 5, 257):
            tt = {k: (v[: 3 * nt] if k in ("x", "x0", "n0", "f0", "g0") else v[:nt] if k == "b" else v)
                  for k, v in t.items() if k != "ijk"}
            tt["n"] = nt
            _check(hor.ops.eval_Stokes_SL_offSurf(ctx, 0, DEL, ss, tt), oracle.sl_off(DEL, ss, tt))
            _check(hor.ops.eval_Stokes_DL_offSurf(ctx, 0, DEL, ss, tt), oracle.dl_off(DEL, ss, tt))
    # targets on the quadrature points themselves, treated as off-surface with b = 0
    nt = 40
    tt = dict(n=nt, x=s["x"][: 3 * nt].copy(), b=np.zeros(nt), x0=s["x"][: 3 * nt].copy(),
              n0=s["nrml"][: 3 * nt].copy(), f0=s["f"][: 3 * nt].copy(), g0=s["g"][: 3 * nt].copy())
    tt["b"][::3] = -0.5 * DEL     # some inside: jump term g0
    _check(hor.ops.eval_Stokes_SL_offSurf(ctx, 0, DEL, s, tt), oracle.sl_off(DEL, s, tt))
    _check(hor.ops.eval_Stokes_DL_offSurf(ctx, 0, DEL, s, tt), oracle.dl_off(DEL, s, tt))


def test_mixed_precision_error(ctx, ell):
    """FP32 far pairs with FP64 accumulation stay within 1e-6 of FP64."""
    body, s, t, DEL = ell
    for kind, dens, d0 in ((hor.SL_OFF, s["f"], t["f0"]), (hor.DL_OFF, s["g"], t["g0"])):
        p64 = hor.Plan(ctx, kind, s, t, DEL, precision=hor.FP64)
        p32 = hor.Plan(ctx, kind, s, t, DEL, precision=hor.MIXED)
        a, b = p32(dens, d0), p64(dens, d0)
        print("mixed vs fp64", kind, rel_l2(a, b), flush=True)
        assert rel_l2(a, b) <= 1e-6


def test_plan_reuse_is_deterministic(ctx, ell):
    body, s, t, DEL = ell
    p = hor.Plan(ctx, hor.DL_OFF, s, t, DEL)
    a = p(s["g"], t["g0"])
    b = p(s["g"], t["g0"])
    assert np.array_equal(a, b)


def test_nan_density_rejected(ctx, ell):
    body, s, t, DEL = ell
    p = hor.Plan(ctx, hor.SL_ON, s, None, DEL)
    f = s["f"].copy()
    f[5] = np.nan
    with pytest.raises(hor.HorError):
        p(f)
