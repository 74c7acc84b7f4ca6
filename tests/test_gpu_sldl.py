"""The fused Stokeslet + stresslet operator (HOR_SLDL_OFF): eval_Stokes_SL_offSurf
(stokeslets.cpp:12-86) and eval_Stokes_DL_offSurf (stresslets.cpp:13-96) at the
same off-surface targets in one pass.

* parity: both halves against the CPU oracle, <= 1e-12 relative (2-norm and
  max-norm), the bar of the separate sums;
* the halves agree with the separate SL_OFF and DL_OFF plans to 1e-15
  relative (the fused pair functions keep each kernel's expressions; only
  the shared separation and 1/r are computed once);
* targets on both sides of the surface and beyond 8 delta (plain class), a
  target count that is not a multiple of the 256-target block;
* mixed precision: within 1e-6 of FP64 and 1e-7 (FP32 rounding) of the
  separate mixed operators;
* the operator is a direct sum only.
"""
import numpy as np
import pytest

import hor
from cases import ellipsoid_case
from conftest import rel_l2, rel_max

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _check(a, b, tol=TOL):
    assert np.all(np.isfinite(a))
    assert rel_l2(a, b) <= tol, rel_l2(a, b)
    assert rel_max(a, b) <= tol, rel_max(a, b)


def _fused_and_separate(ctx, s, t, DEL):
    fsl, fdl = hor.ops.eval_Stokes_SL_DL_offSurf(ctx, 0, DEL, s, t)
    sl = hor.ops.eval_Stokes_SL_offSurf(ctx, 0, DEL, s, t)
    dl = hor.ops.eval_Stokes_DL_offSurf(ctx, 0, DEL, s, t)
    return fsl, fdl, sl, dl


def test_fused_matches_oracle_and_separate_band(ctx, oracle):
    body, s, t, DEL = ellipsoid_case(h=1 / 32, max_band=6001)
    assert (t["b"] < 0).sum() > 500 and (t["b"] > 0).sum() > 500
    print("fused band", s["n"], t["n"], flush=True)
    fsl, fdl, sl, dl = _fused_and_separate(ctx, s, t, DEL)
    _check(fsl, oracle.sl_off(DEL, s, t))
    _check(fdl, oracle.dl_off(DEL, s, t))
    _check(fsl, sl, 1e-15)
    _check(fdl, dl, 1e-15)


def test_fused_far_targets_and_ragged_count(ctx, oracle):
    """Grid nodes out to 12 delta: a share of the targets is in the plain
    class (|b|/delta > 8), with no jump term; 999 targets."""
    body, s, t, DEL = ellipsoid_case(h=1 / 16, band_factor=12.0, max_band=999)
    far = np.abs(t["b"]) / DEL > 8
    assert far.sum() > 20 and (~far).sum() > 500
    fsl, fdl, sl, dl = _fused_and_separate(ctx, s, t, DEL)
    _check(fsl, oracle.sl_off(DEL, s, t))
    _check(fdl, oracle.dl_off(DEL, s, t))
    _check(fsl, sl, 1e-15)
    _check(fdl, dl, 1e-15)


def test_fused_mixed_precision(ctx):
    """Mixed precision (FP32 far pairs and 4-8 delta regularized pairs, FP64
    accumulation): within 1e-6 of FP64, like the separate mixed operators,
    and within FP32 rounding (1e-7) of them: the same FP32 expressions, which
    the compiler may contract differently."""
    body, s, t, DEL = ellipsoid_case(h=1 / 32, max_band=6001)
    n = t["n"]
    fg, fg0 = np.concatenate([s["f"], s["g"]]), np.concatenate([t["f0"], t["g0"]])
    out64 = hor.Plan(ctx, hor.SLDL_OFF, s, t, DEL)(fg, fg0)
    out32 = hor.Plan(ctx, hor.SLDL_OFF, s, t, DEL, precision=hor.MIXED)(fg, fg0)
    for half, kind, dens, d0 in ((slice(0, 3 * n), hor.SL_OFF, s["f"], t["f0"]),
                                 (slice(3 * n, 6 * n), hor.DL_OFF, s["g"], t["g0"])):
        sep32 = hor.Plan(ctx, kind, s, t, DEL, precision=hor.MIXED)(dens, d0)
        print("fused mixed", kind, "vs fp64", rel_l2(out32[half], out64[half]), "vs separate mixed",
              rel_l2(out32[half], sep32), flush=True)
        assert rel_l2(out32[half], out64[half]) <= 1e-6
        assert rel_l2(out32[half], sep32) <= 1e-7


def test_fused_is_direct_only(ctx):
    body, s, t, DEL = ellipsoid_case(h=1 / 8, max_band=200)
    with pytest.raises(hor.HorError):
        hor.Plan(ctx, hor.SLDL_OFF, s, t, DEL, method=hor.TREE, ctr=body.ctr, tree=hor.TreeParams(4, 100, 0.6))
