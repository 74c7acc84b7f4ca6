"""The exact velocity of a translating ellipsoid (Oberbeck's exterior flow,
hor_translating_exact_velocity), pinned without the GPU:

* prolate spheroid translating along its axis: the reference's own SL_exact
  (extend_to_grid/utilities.cpp:360-411) at the exterior grid nodes of the
  golden fixture grid_h16.npz;
* sphere, any direction: Stokes' closed form;
* triaxial body (1, 0.8, 0.6), rotated, any direction:
  - it is U on the surface (continuity of the exterior field at lambda -> 0);
  - it is divergence-free (central differences);
  - it is the single layer of the body's traction (the oracle's plain sum at
    h = 1/32 on points well off the surface, where the quadrature is accurate
    to near rounding): the double layer of a rigid motion vanishes outside,
    so the exterior flow is the single layer alone.  A wrong constant in the
    drag, a swapped axis or a wrong sign fails this.
"""
import numpy as np

from conftest import golden
from hor import geometry as G


def _exterior(x, body):
    R = np.eye(3) if body.R is None else np.asarray(body.R).reshape(3, 3)
    xb = (x - body.ctr) @ R.T
    return ((xb / np.asarray(body.abc)) ** 2).sum(axis=1) > 1.0


def test_prolate_matches_reference_sl_exact():
    g = golden("grid_h16.npz")
    c = float(g["ctr"])
    body = G.Body((1.0, 0.5, 0.5), (c, c, c))
    tx = (g["ijk"] * float(g["h"])).astype(np.float64).reshape(-1, 3)
    out = _exterior(tx, body)
    assert out.sum() > 1000
    u = body.exact_velocity(tx[out].ravel()).reshape(-1, 3)
    ref = g["exact"].reshape(-1, 3)[out]
    err = np.abs(u - ref).max()
    print(f"{out.sum()} exterior nodes: max |Oberbeck - SL_exact| = {err:.2e}")
    # two FP64 evaluations of one function through different formulas (Chwang & Wu's
    # logarithms, the ellipsoidal integrals): rounding-level agreement, |U| = 1
    assert err <= 2e-12
    # inside and on the body the velocity is U
    assert np.array_equal(body.exact_velocity(tx[~out].ravel()).reshape(-1, 3),
                          np.tile([1.0, 0.0, 0.0], ((~out).sum(), 1)))


def test_sphere_is_stokes_solution():
    rng = np.random.default_rng(1)
    R = 0.7
    body = G.Body((R, R, R), (0.1, -0.2, 0.3), G.random_rotation(rng))
    U = rng.uniform(-1, 1, 3)
    d = rng.standard_normal((500, 3))
    d *= (R * rng.uniform(1.0001, 20.0, 500) / np.linalg.norm(d, axis=1))[:, None]
    u = body.exact_velocity((d + body.ctr).ravel(), U).reshape(-1, 3)
    r = np.linalg.norm(d, axis=1)[:, None]
    Ux = (d @ U)[:, None]
    want = 0.75 * R * (U / r + Ux * d / r ** 3) + 0.25 * R ** 3 * (U / r ** 3 - 3 * Ux * d / r ** 5)
    err = np.abs(u - want).max() / np.abs(U).max()
    print(f"sphere: max error {err:.2e}")
    assert err <= 1e-13


def _triaxial(seed=2):
    rng = np.random.default_rng(seed)
    body = G.Body((1.0, 0.8, 0.6), (0.05, -0.1, 0.02), G.random_rotation(rng))
    return rng, body, rng.uniform(-1, 1, 3)


def test_triaxial_boundary_value_and_divergence():
    rng, body, U = _triaxial()
    s = body.surface(1.0 / 16)
    x = s["x"].reshape(-1, 3)
    for eps in (1e-6, 1e-9):
        xo = body.ctr + (x - body.ctr) * (1.0 + eps)
        v = body.exact_velocity(xo.ravel(), U).reshape(-1, 3)
        err = np.abs(v - U).max() / np.abs(U).max()
        print(f"surface scaled by 1+{eps:g}: max |v - U| / |U| = {err:.2e}")
        assert err <= 20 * eps
    # divergence by central differences at points 0.05 .. 2 off the surface
    n = s["nrml"].reshape(-1, 3)
    idx = rng.choice(len(x), 200, replace=False)
    p = x[idx] + n[idx] * rng.uniform(0.05, 2.0, (200, 1))
    hstep = 1e-4
    div = np.zeros(200)
    for k in range(3):
        e = np.zeros(3)
        e[k] = hstep
        vp = body.exact_velocity((p + e).ravel(), U).reshape(-1, 3)[:, k]
        vm = body.exact_velocity((p - e).ravel(), U).reshape(-1, 3)[:, k]
        div += (vp - vm) / (2 * hstep)
    scale = np.abs(body.exact_velocity(p.ravel(), U)).max()
    print(f"max |div v| = {np.abs(div).max():.2e} (|v| up to {scale:.2f})")
    assert np.abs(div).max() <= 1e-6 * scale


def test_triaxial_equals_single_layer_of_traction(oracle):
    """The plain single-layer sum of the traction at points 2 to 4 away from
    the body converges to the exact velocity as h falls (a wrong constant
    would leave the error at a plateau)."""
    rng, body, U = _triaxial(3)
    d = rng.standard_normal((300, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    p = body.ctr + d * rng.uniform(2.0, 4.0, (300, 1))
    ex = body.exact_velocity(p.ravel(), U).reshape(-1, 3)
    # every target flagged beyond 8 delta: the plain Stokeslet sum, no subtraction
    tgt = dict(n=300, x=p.ravel(), b=np.full(300, 1e3), n0=np.zeros(900), f0=np.zeros(900))
    errs = []
    for h in (1.0 / 32, 1.0 / 64):
        s = body.surface(h)
        s["f"] = body.traction(s["x"], U)
        sl = oracle.sl_off(G.delta(h), s, tgt).reshape(-1, 3)
        errs.append(np.abs(sl - ex).max() / np.abs(ex).max())
        print(f"h = {h:g}: single layer of the traction vs exact at 300 exterior points, max rel error {errs[-1]:.2e}")
    assert errs[1] <= 1e-7
    assert errs[0] / errs[1] >= 64.0      # order >= 6 between h = 1/32 and 1/64
