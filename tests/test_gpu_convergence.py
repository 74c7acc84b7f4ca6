"""Convergence of the GPU sums on the analytic translating-body solution.

A prolate spheroid (1, 0.5, 0.5) translating with U = (1, 0, 0) along its axis:
the single layer of the exact surface traction is U on and inside the body and
the Chwang & Wu velocity outside (extend_to_grid/utilities.cpp:267-285,
360-411).  The same for the triaxial ellipsoid (1, 0.8, 0.6) of the band
workload translating along x, and rotated and translating in a seeded
direction, against Oberbeck's exterior flow (hor_translating_exact_velocity).
The FP64 direct SL sums (on-surface at the quadrature points, off-surface at
grid nodes within 3 delta of the surface) are compared with it
at h = 1/32, 1/64, 1/128 (delta = 4 (1/64)^(2/7) h^(5/7)); the maximum error
must fall at an observed order of at least 3 between the two finest levels
(the regularization error is O(delta^7) = O(h^5), the quadrature error
high order for this smooth integrand).
"""
import numpy as np
import pytest

import hor
from hor import geometry as G

pytestmark = pytest.mark.gpu


def _errors(ctx, N, body=None, U=(1.0, 0.0, 0.0), max_targets=30000, seed=3):
    h = 1.0 / N
    DEL = G.delta(h)
    prolate = body is None
    body = G.Body((1.0, 0.5, 0.5)) if prolate else body
    U = np.asarray(U, dtype=np.float64)
    s = body.surface(h)
    s["f"] = body.traction(s["x"], U)
    t = body.band_targets(s, (-1.5, -1.5, -1.5), h, int(round(3.0 / h)), 3 * DEL)
    if t["n"] > max_targets:
        idx = np.sort(np.random.default_rng(seed).choice(t["n"], max_targets, replace=False))
        t = dict(n=len(idx), x=t["x"].reshape(-1, 3)[idx].ravel(), b=t["b"][idx],
                 x0=t["x0"].reshape(-1, 3)[idx].ravel(), n0=t["n0"].reshape(-1, 3)[idx].ravel())
    t["f0"] = body.traction(t["x0"], U)
    on = hor.ops.eval_Stokes_SL_onSurf(ctx, DEL, s).reshape(-1, 3)
    off = hor.ops.eval_Stokes_SL_offSurf(ctx, h, DEL, s, t).reshape(-1, 3)
    inside, outside = t["b"] < -1e-12, t["b"] > 1e-12
    ex = (body.spheroid_exact(t["x"]) if prolate else body.exact_velocity(t["x"], U)).reshape(-1, 3)
    e_on = np.abs(on - U).max()
    e_in = np.abs(off[inside] - U).max() if inside.any() else 0.0
    e_out = np.abs(off[outside] - ex[outside]).max() if outside.any() else 0.0
    return max(e_on, e_in, e_out), (e_on, e_in, e_out), s["n"], t["n"]


_BODIES = {
    "prolate": (None, (1.0, 0.0, 0.0)),
    "triaxial": (G.Body((1.0, 0.8, 0.6)), (1.0, 0.0, 0.0)),
    "triaxial-rotated": (G.Body((1.0, 0.8, 0.6), (0.03, -0.02, 0.01), G.random_rotation(np.random.default_rng(5))),
                         tuple(np.random.default_rng(6).uniform(-1, 1, 3))),
}


@pytest.mark.parametrize("name", list(_BODIES))
def test_sl_converges_to_exact_translating_body(ctx, name):
    body, U = _BODIES[name]
    print(f"{name}: semi-axes {(1.0, 0.5, 0.5) if body is None else tuple(body.abc)}, U = {np.round(U, 3)}")
    rows = [(N, *_errors(ctx, N, body, U)) for N in (32, 64, 128)]
    for N, e, parts, ns, nt in rows:
        print(f"h = 1/{N}: {ns} sources, {nt} band targets, max error {e:.3e} "
              f"(on {parts[0]:.2e}, inside {parts[1]:.2e}, outside {parts[2]:.2e})")
    orders = [np.log2(rows[k][1] / rows[k + 1][1]) for k in range(2)]
    print("observed orders:", orders)
    assert rows[2][1] < rows[1][1] < rows[0][1]
    assert orders[1] >= 3.0
