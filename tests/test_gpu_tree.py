"""GPU parity of the KITC treecode against the CPU oracle's treecode.

Same tree, same MAC decisions, same leaf test: the FP64 GPU path agrees with
the oracle to ||.||_2 relative 1e-12 (max-norm relative 1e-12).  The mixed
path (FP32 proxies and far pairs) is held to 1e-6 against FP64.
"""
import numpy as np
import pytest

import hor
from cases import ellipsoid_case, spheroid_pair
from conftest import rel_l2, rel_max

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _check(a, b, tol=TOL):
    assert np.all(np.isfinite(a))
    assert rel_l2(a, b) <= tol, rel_l2(a, b)
    assert rel_max(a, b) <= tol, rel_max(a, b)


@pytest.fixture(scope="module")
def ell():
    return ellipsoid_case(h=1 / 16, max_band=2000)


# P = 12 is the largest order a plan accepts: 13 x 13 proxy rows (the moment
# GEMM's third row block has a wave past the last row, which only stages
# members) and 117 weight columns (eight column tiles).
@pytest.mark.parametrize("P,N0,theta", [(6, 200, 0.6), (8, 300, 0.6), (4, 150, 0.8), (12, 300, 0.6)])
def test_tree_four_ops(ctx, oracle, ell, P, N0, theta):
    body, s, t, DEL = ell
    ctr = body.ctr
    tp_o = oracle.tree_params(P, N0, theta, ctr)
    tp_g = hor.TreeParams(P, N0, theta)
    print("tree", P, N0, theta, s["n"], t["n"], flush=True)
    _check(hor.ops.eval_Stokes_SL_onSurf_tree(ctx, ctr, 0, DEL, s, tree=tp_g), oracle.sl_on_tree(tp_o, DEL, s))
    _check(hor.ops.eval_Stokes_DL_onSurf_tree(ctx, ctr, 0, DEL, s, tree=tp_g), oracle.dl_on_tree(tp_o, DEL, s))
    _check(hor.ops.eval_Stokes_SL_offSurf_tree(ctx, ctr, 0, DEL, s, t, tree=tp_g),
           oracle.sl_off_tree(tp_o, DEL, s, t))
    _check(hor.ops.eval_Stokes_DL_offSurf_tree(ctx, ctr, 0, DEL, s, t, tree=tp_g),
           oracle.dl_off_tree(tp_o, DEL, s, t))


def test_tree_reference_defaults_spheroid_pair(ctx, oracle):
    """P=6, N0=1000, theta=0.6 (KITC.h:14-17) on the reference's geometry at h=1/32,
    including far targets (nrst = 0) and the treecode's jump rule."""
    h, DEL, (b1, b2), (s1, s2), (t1, t2) = spheroid_pair(N=32)
    rng = np.random.default_rng(11)
    s1["g"] = rng.uniform(-1, 1, 3 * s1["n"])
    t2["g0"] = np.where(np.repeat(t2["nrst"] != 0, 3), rng.uniform(-1, 1, 3 * t2["n"]), 0.0)
    tp_o = oracle.tree_params(6, 1000, 0.6, b1.ctr)
    _check(hor.ops.eval_Stokes_SL_offSurf_tree(ctx, b1.ctr, h, DEL, s1, t2), oracle.sl_off_tree(tp_o, DEL, s1, t2))
    _check(hor.ops.eval_Stokes_DL_offSurf_tree(ctx, b1.ctr, h, DEL, s1, t2), oracle.dl_off_tree(tp_o, DEL, s1, t2))
    _check(hor.ops.eval_Stokes_SL_onSurf_tree(ctx, b1.ctr, h, DEL, s1), oracle.sl_on_tree(tp_o, DEL, s1))
    _check(hor.ops.eval_Stokes_DL_onSurf_tree(ctx, b1.ctr, h, DEL, s1), oracle.dl_on_tree(tp_o, DEL, s1))


def test_tree_dl_reused_plan_subtraction_coefficients(ctx, oracle, ell):
    """A double-layer treecode plan evaluated again and again: the first
    evaluation subtracts g0 inside its kernels, the second computes the plan's
    subtraction coefficients C_i and sums on g alone (D[g] - C g0), and so do
    the later ones.  Every evaluation, on new seeded densities, against the
    oracle's treecode.  The off-surface targets include 20 placed on sources
    (b = 0, a self pair in the leaf sums) and 20 at delta/2 inside the surface
    below a source (x0 = y, b = -delta/2)."""
    body, s, t, DEL = ell
    ctr = body.ctr
    P8 = hor.TreeParams(8, 300, 0.6)
    tp_o = oracle.tree_params(8, 300, 0.6, ctr)
    nt = t["n"]
    t = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in t.items()}
    y, n = s["x"][:120].reshape(-1, 3), s["nrml"][:120].reshape(-1, 3)
    b = np.where(np.arange(40) < 20, 0.0, -0.5 * DEL)
    t["x"][:120] = (y + b[:, None] * n).ravel()
    t["x0"][:120] = y.ravel()
    t["n0"][:120] = n.ravel()
    t["b"][:40] = b
    rng = np.random.default_rng(23)
    p_on = hor.Plan(ctx, hor.DL_ON, s, None, DEL, method=hor.TREE, ctr=ctr, tree=P8)
    p_off = hor.Plan(ctx, hor.DL_OFF, s, t, DEL, method=hor.TREE, ctr=ctr, tree=P8)
    for it in range(3):
        s["g"] = rng.uniform(-1, 1, 3 * s["n"])
        t["g0"] = rng.uniform(-1, 1, 3 * nt)
        on, off = p_on(s["g"]), p_off(s["g"], t["g0"])
        e_on, e_off = rel_l2(on, oracle.dl_on_tree(tp_o, DEL, s)), rel_l2(off, oracle.dl_off_tree(tp_o, DEL, s, t))
        print(f"evaluation {it + 1}: DL_ON {e_on:.2e}, DL_OFF {e_off:.2e}", flush=True)
        _check(on, oracle.dl_on_tree(tp_o, DEL, s))
        _check(off, oracle.dl_off_tree(tp_o, DEL, s, t))
    p_on.close()
    p_off.close()


def test_tree_single_leaf_equals_direct(ctx, ell):
    """With N0 > N the root is one leaf: the treecode is the direct sum."""
    body, s, t, DEL = ell
    tp = hor.TreeParams(6, 10 * s["n"], 0.6)
    a = hor.ops.eval_Stokes_DL_offSurf_tree(ctx, body.ctr, 0, DEL, s, t, tree=tp)
    b = hor.ops.eval_Stokes_DL_offSurf(ctx, 0, DEL, s, t)
    # the direct DL adds g0 only in its regularized branch, the treecode for b<0 always
    _check(a, b)


def test_tree_mixed_vs_fp64(ctx, ell):
    body, s, t, DEL = ell
    tp = hor.TreeParams(8, 300, 0.6)
    for kind, dens, d0 in ((hor.SL_OFF, s["f"], t["f0"]), (hor.DL_OFF, s["g"], t["g0"]),
                           (hor.DL_ON, s["g"], None)):
        a = hor.Plan(ctx, kind, s, None if kind == hor.DL_ON else t, DEL, hor.TREE, hor.MIXED, body.ctr, tp)(dens, d0)
        b = hor.Plan(ctx, kind, s, None if kind == hor.DL_ON else t, DEL, hor.TREE, hor.FP64, body.ctr, tp)(dens, d0)
        print("tree mixed vs fp64", kind, rel_l2(a, b), flush=True)
        assert rel_l2(a, b) <= 1e-6


@pytest.mark.parametrize("P,N0", [(8, 300), (10, 400)])
def test_tree_sl_far_mfma(ctx, oracle, ell, P, N0):
    """The mixed single-layer far field on FP32 matrix cores (FAR_MFMA) against
    the oracle's FP64 treecode, on-surface and off-surface: within the mixed
    path's 1e-6, repeatable, and close to the FP32 vector kernel (FAR_VALU).
    (P + 1)^3 = 729 and 1331 proxies are not multiples of four, so the padded
    proxy rows are exercised.  Other plans refuse the matrix-core far field."""
    body, s, t, DEL = ell
    tp_o = oracle.tree_params(P, N0, 0.6, body.ctr)
    tp = hor.TreeParams(P, N0, 0.6)
    cases = ((hor.SL_ON, None, None, oracle.sl_on_tree(tp_o, DEL, s)),
             (hor.SL_OFF, t, t["f0"], oracle.sl_off_tree(tp_o, DEL, s, t)))
    for kind, tgt, d0, ref in cases:
        valu = hor.Plan(ctx, kind, s, tgt, DEL, hor.TREE, hor.MIXED, body.ctr, tp)(s["f"], d0)
        plan = hor.Plan(ctx, kind, s, tgt, DEL, hor.TREE, hor.MIXED, body.ctr, tp).set_far_kernel(hor.FAR_MFMA)
        assert plan.stats()["proxy_pairs"] > 0
        mfma = plan(s["f"], d0)
        again = plan(s["f"], d0)
        e_m, e_v, d_mv = rel_l2(mfma, ref), rel_l2(valu, ref), rel_l2(mfma, valu)
        print(f"SL kind {kind} P={P}: FAR_MFMA vs oracle {e_m:.2e} (max {rel_max(mfma, ref):.2e}), "
              f"FAR_VALU vs oracle {e_v:.2e}, MFMA vs VALU {d_mv:.2e}", flush=True)
        assert np.all(np.isfinite(mfma))
        assert np.array_equal(mfma, again)
        assert e_m <= 1e-6
        assert d_mv <= 1e-6
    plan = hor.Plan(ctx, hor.DL_ON, s, None, DEL, hor.TREE, hor.MIXED, body.ctr, tp)
    with pytest.raises(hor.HorError):
        plan.set_far_kernel(hor.FAR_MFMA)
    plan = hor.Plan(ctx, hor.SL_ON, s, None, DEL, hor.TREE, hor.FP64, body.ctr, tp)
    with pytest.raises(hor.HorError):
        plan.set_far_kernel(hor.FAR_MFMA)
    plan.set_far_kernel(hor.FAR_VALU)


def test_tree_converges_in_P(ctx, ell):
    """The treecode converges as P grows (self-convergence against P = 10).
    Its limit is not the regularized direct sum: the far field always uses the
    plain kernel (KITC.cpp:1089-1101), also for accepted clusters within
    8 delta, which leaves a P-independent gap at this coarse h."""
    body, s, t, DEL = ell
    ref = hor.ops.eval_Stokes_SL_offSurf_tree(ctx, body.ctr, 0, DEL, s, t, tree=hor.TreeParams(10, 150, 0.6))
    errs = [rel_l2(hor.ops.eval_Stokes_SL_offSurf_tree(ctx, body.ctr, 0, DEL, s, t, tree=hor.TreeParams(P, 150, 0.6)), ref)
            for P in (2, 4, 6, 8)]
    print("tree vs P=10 by P", errs, flush=True)
    assert all(errs[k + 1] < errs[k] for k in range(3))
    assert errs[-1] < 1e-5
