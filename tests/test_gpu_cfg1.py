"""The direct sums of the CPU-runnable ellipsoid workload in full, GPU against
the oracle: ellipsoid (1, 0.8, 0.6) at h = 1/32 (10,618 quadrature points),
seeded spherical-harmonic densities f and g (coefficients ~U(-1,1)), targets
= every quadrature point (on-surface SL and DL, stokeslets.cpp:141-180,
stresslets.cpp:186-228) and every grid node within 3 delta of the surface
(about 170k nodes, off-surface SL and DL, stokeslets.cpp:12-86,
stresslets.cpp:13-96).  Every target is compared, <= 1e-12 relative
(2-norm and max-norm)."""
import numpy as np
import pytest

import hor
from cases import ellipsoid_case
from conftest import rel_l2, rel_max

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg1():
    body, s, t, DEL = ellipsoid_case(h=1 / 32)
    assert s["n"] == 10618 and t["n"] > 150000
    return body, s, t, DEL


def _check(name, a, b):
    e2, em = rel_l2(a, b), rel_max(a, b)
    print(f"{name}: {a.size // 3} targets, rel l2 {e2:.2e}, rel max {em:.2e}", flush=True)
    assert np.all(np.isfinite(a)) and e2 <= 1e-12 and em <= 1e-12


def test_cfg1_on_surface(ctx, oracle, cfg1):
    body, s, t, DEL = cfg1
    _check("SL on-surface", hor.ops.eval_Stokes_SL_onSurf(ctx, DEL, s), oracle.sl_on(DEL, s))
    _check("DL on-surface", hor.ops.eval_Stokes_DL_onSurf(ctx, DEL, s), oracle.dl_on(DEL, s))


def test_cfg1_band(ctx, oracle, cfg1):
    body, s, t, DEL = cfg1
    print(f"band: {t['n']} grid nodes within 3 delta, {(t['b'] < 0).sum()} inside", flush=True)
    osl, odl = oracle.sl_off(DEL, s, t), oracle.dl_off(DEL, s, t)
    _check("SL off-surface", hor.ops.eval_Stokes_SL_offSurf(ctx, 0, DEL, s, t), osl)
    _check("DL off-surface", hor.ops.eval_Stokes_DL_offSurf(ctx, 0, DEL, s, t), odl)
    fsl, fdl = hor.ops.eval_Stokes_SL_DL_offSurf(ctx, 0, DEL, s, t)
    _check("fused SL half", fsl, osl)
    _check("fused DL half", fdl, odl)
