"""Host geometry: bodies, quadrature, nearest data, band targets, densities.

Surfaces and target sets are plain dicts of numpy arrays (xyz interleaved):

* surface: ``n, x, nrml, area, plane, ctr`` (+ ``f``/``g`` densities)
* targets: ``n, x, b, x0, n0, nrst, i123`` (+ ``f0``/``g0``), nearest data
  relative to a surface (Compute_nearest_data, surface.cpp:450-526).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import check, f64, i32, ptr

PI = 3.14159265358979323846


def delta(h, kappa0=4.0, q=5.0 / 7.0):
    """delta = kappa0 (1/64)^(1-q) h^q (highreg_stokes_2surf/main.cpp:147-157)."""
    kappa = kappa0 * math.pow(1.0 / 64.0, 1.0 - q)
    return kappa * math.pow(h, q)


def rot_z(gamma_deg):
    """The reference's un-rotation matrix for a rotation by gamma about z
    (utilities.cpp:10-40 with alpha = beta = 0), row-major."""
    al, be, ga = 0.0, 0.0, gamma_deg * PI / 180.0
    ca, sa, cb, sb, cg, sg = math.cos(al), math.sin(al), math.cos(be), math.sin(be), math.cos(ga), math.sin(ga)
    return np.array([cb * cg, cb * sg, -sb,
                     sa * sb * cg - ca * sg, sa * sb * sg + ca * cg, sa * cb,
                     ca * sb * cg + sa * sg, ca * sb * sg - sa * cg, ca * cb])


def random_rotation(rng):
    """Uniform random SO(3) matrix (row-major), from a QR of a Gaussian matrix."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q.ravel()


class Body:
    """Ellipsoid with semi-axes abc, centre ctr, un-rotation R (x_body = R (x - ctr))."""

    def __init__(self, abc, ctr=(0.0, 0.0, 0.0), R=None):
        self.abc = np.asarray(abc, dtype=np.float64)
        self.ctr = np.asarray(ctr, dtype=np.float64)
        self.R = None if R is None else np.asarray(R, dtype=np.float64).ravel()
        self.c = _lib.Body()
        a, c = f64(self.abc, 3), f64(self.ctr, 3)
        r = None if self.R is None else f64(self.R, 9)
        check(_lib.lib().hor_body_init(C.byref(self.c), ptr(a), ptr(c), ptr(r)), "hor_body_init")

    def surface(self, h, pou_deg=70.0, ctx=None):
        """Generate_Surface (surface.cpp:13-71), translated to the centre; on
        the GPU of ctx when given (hor_surface_generate_gpu)."""
        L = _lib.lib()
        if ctx is None:
            fn, head, name = L.hor_surface_generate, (), "hor_surface_generate"
        else:
            fn, head, name = L.hor_surface_generate_gpu, (ctx.h,), "hor_surface_generate_gpu"
        n = C.c_int()
        check(fn(*head, C.byref(self.c), h, pou_deg, 0, C.byref(n), None, None, None, None), name)
        n = n.value
        x, nr, a, pl = np.zeros(3 * n), np.zeros(3 * n), np.zeros(n), np.zeros(n, np.int32)
        m = C.c_int()
        check(fn(*head, C.byref(self.c), h, pou_deg, n, C.byref(m), ptr(x), ptr(nr), ptr(a), ptr(pl)), name)
        return dict(n=n, x=x, nrml=nr, area=a, plane=pl, ctr=self.ctr.copy(), h=h)

    def nearest(self, h, DEL, surf, tx, with_force=True, ctx=None):
        """Nearest data of targets tx relative to this body's surface surf; on
        the GPU of ctx when given (hor_nearest_data_gpu)."""
        tx = f64(tx)
        nt = tx.size // 3
        nrst = np.zeros(nt, np.int32)
        x0, n0, b = np.zeros(3 * nt), np.zeros(3 * nt), np.zeros(nt)
        i123 = np.zeros(3 * nt, np.int32)
        f0 = np.zeros(3 * nt) if with_force else None
        args = (C.byref(self.c), h, DEL, surf["n"], ptr(surf["x"]), ptr(i32(surf["plane"])), nt, ptr(tx),
                ptr(nrst), ptr(x0), ptr(n0), ptr(b), ptr(i123), ptr(f0))
        if ctx is None:
            check(_lib.lib().hor_nearest_data(*args), "hor_nearest_data")
        else:
            check(_lib.lib().hor_nearest_data_gpu(ctx.h, *args), "hor_nearest_data_gpu")
        d = dict(n=nt, x=tx, nrst=nrst, x0=x0, n0=n0, b=b, i123=i123)
        if with_force:
            d["f0"] = f0
        return d

    def force(self, x, nrml):
        """surfc_force (surface.cpp:530-559)."""
        x, nrml = f64(x), f64(nrml)
        f = np.zeros_like(x)
        check(_lib.lib().hor_surface_force(C.byref(self.c), x.size // 3, ptr(x), ptr(nrml), ptr(f)),
              "hor_surface_force")
        return f

    def band_targets(self, surf, origin, h, ngrid, band):
        """Grid nodes origin + h*ijk with |b| <= band (find_nrst_surfc)."""
        L = _lib.lib()
        org = f64(origin, 3)
        cnt = C.c_int()
        args = (C.byref(self.c), ptr(org), h, int(ngrid), band, surf["n"], ptr(surf["x"]))
        check(L.hor_band_targets(*args, 0, C.byref(cnt), None, None, None, None, None), "hor_band_targets")
        n = cnt.value
        ijk = np.zeros(3 * n, np.int32)
        x, b, x0, n0 = np.zeros(3 * n), np.zeros(n), np.zeros(3 * n), np.zeros(3 * n)
        check(L.hor_band_targets(*args, n, C.byref(cnt), ptr(ijk), ptr(x), ptr(b), ptr(x0), ptr(n0)),
              "hor_band_targets")
        return dict(n=n, ijk=ijk.reshape(n, 3), x=x, b=b, x0=x0, n0=n0)

    def traction(self, x, U=(1.0, 0.0, 0.0)):
        """Single-layer density of this body translating with velocity U."""
        x, u = f64(x), f64(np.asarray(U), 3)
        f = np.zeros_like(x)
        check(_lib.lib().hor_traction_translating(C.byref(self.c), ptr(u), x.size // 3, ptr(x), ptr(f)),
              "hor_traction_translating")
        return f

    def exact_velocity(self, x, U=(1.0, 0.0, 0.0)):
        """Exact velocity of this ellipsoid translating with U (U inside and on
        the body, Oberbeck's exterior flow outside)."""
        x, u0 = f64(x), f64(np.asarray(U), 3)
        u = np.zeros_like(x)
        check(_lib.lib().hor_translating_exact_velocity(C.byref(self.c), ptr(u0), x.size // 3, ptr(x), ptr(u)),
              "hor_translating_exact_velocity")
        return u

    def spheroid_exact(self, x):
        """Exact exterior velocity of the prolate spheroid translating with U=(1,0,0) in its body frame."""
        x = f64(x)
        u = np.zeros_like(x)
        check(_lib.lib().hor_spheroid_exact_velocity(C.byref(self.c), x.size // 3, ptr(x), ptr(u)),
              "hor_spheroid_exact_velocity")
        return u


def stencils(ctx, surf, tgt, h):
    """6x6 Monge stencils of targets tgt (nearest data relative to surf) on
    the GPU (hor_stencil_build): (nt, 36) source indices [k1][k2] and (nt, 12)
    Lagrange weights along i1 then i2."""
    nt = tgt["n"]
    st, co = np.zeros(36 * nt, np.int32), np.zeros(12 * nt)
    c = f64(np.asarray(surf["ctr"]), 3)
    check(_lib.lib().hor_stencil_build(ctx.h, ptr(c), h, surf["n"], ptr(f64(surf["x"])), nt, ptr(i32(tgt["nrst"])),
                                       ptr(f64(tgt["x0"])), ptr(i32(tgt["i123"])), ptr(st), ptr(co)),
          "hor_stencil_build")
    return st.reshape(nt, 36), co.reshape(nt, 12)


# ---------------------------------------------------------------------------
# Seeded smooth densities: real spherical harmonics of degree <= 3 of the unit
# vector from the body centre (16 basis functions), with coefficients
# ~U(-1,1) per component.

def _sh_basis(u):
    x, y, z = u[:, 0], u[:, 1], u[:, 2]
    one = np.ones_like(x)
    return np.stack([
        0.5 * math.sqrt(1 / PI) * one,
        math.sqrt(3 / (4 * PI)) * y, math.sqrt(3 / (4 * PI)) * z, math.sqrt(3 / (4 * PI)) * x,
        0.5 * math.sqrt(15 / PI) * x * y, 0.5 * math.sqrt(15 / PI) * y * z,
        0.25 * math.sqrt(5 / PI) * (3 * z * z - 1), 0.5 * math.sqrt(15 / PI) * x * z,
        0.25 * math.sqrt(15 / PI) * (x * x - y * y),
        0.25 * math.sqrt(35 / (2 * PI)) * y * (3 * x * x - y * y), 0.5 * math.sqrt(105 / PI) * x * y * z,
        0.25 * math.sqrt(21 / (2 * PI)) * y * (5 * z * z - 1), 0.25 * math.sqrt(7 / PI) * z * (5 * z * z - 3),
        0.25 * math.sqrt(21 / (2 * PI)) * x * (5 * z * z - 1), 0.25 * math.sqrt(105 / PI) * z * (x * x - y * y),
        0.25 * math.sqrt(35 / (2 * PI)) * x * (x * x - 3 * y * y),
    ], axis=1)


def sh_coefficients(seed):
    """(3, 16) coefficients ~U(-1,1) from numpy PCG64(seed)."""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, size=(3, 16))


def sh_density(x, ctr, coef):
    """Density sum_l coef[k, l] Y_l(x/|x|) with x relative to ctr; 3n interleaved."""
    p = f64(x).reshape(-1, 3) - np.asarray(ctr)[None, :]
    u = p / np.linalg.norm(p, axis=1)[:, None]
    Y = _sh_basis(u)
    return np.ascontiguousarray((Y @ np.asarray(coef).T).ravel())
