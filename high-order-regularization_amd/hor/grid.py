"""The grid program (highreg_stokes_extend_to_grid/main.cpp:107-968): the
velocity and pressure of a translating body on the whole grid [0, L]^3.

Boundary-integral values on the box faces (plain sums) and at the band nodes
(regularized sums with subtraction) come from the GPU plans; the extension to
the rest of the grid (transfinite interpolation, 4th-order FD Laplacian,
DST-I Poisson solve) runs on the host (libhor.so, OpenMP).

Grids are numpy arrays of shape (N+1, N+1, N+1), node (i, j, k) at
(i, j, k) * L / N.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from . import geometry as G
from . import ops
from ._lib import check, f64, ptr

__all__ = ["level_set", "exact_pressure", "dst_poisson3d", "solve_poisson", "d4", "gradient_error",
           "grid_program"]


def level_set(body, x):
    """phi of the body at points x (3n) (utilities.cpp:58-86)."""
    x = f64(x)
    out = np.zeros(x.size // 3)
    check(_lib.lib().hor_level_set(C.byref(body.c), out.size, ptr(x), ptr(out)), "hor_level_set")
    return out


def exact_pressure(body, x):
    """p_exact (extend_to_grid/utilities.cpp:294-321) of the prolate spheroid body."""
    x = f64(x)
    out = np.zeros(x.size // 3)
    check(_lib.lib().hor_spheroid_exact_pressure(C.byref(body.c), out.size, ptr(x), ptr(out)),
          "hor_spheroid_exact_pressure")
    return out


def dst_poisson3d(f):
    """fftw_DST_Poisson3d_FD (main.cpp:1423-1491) on an (n, n, n) array of h^2 rhs."""
    f = np.array(f, dtype=np.float64, order="C")
    check(_lib.lib().hor_dst_poisson3d(f.shape[0], ptr(f)), "hor_dst_poisson3d")
    return f


def solve_poisson(L, band, outside, p, u, variable):
    """solve_Poisson (main.cpp:969-1044): returns the extension of u (face and
    band values) to the whole grid.  band: 0/1/2 per node (2 = BAND2);
    outside: phi > 0; variable 0 (pressure) or 1..3 (velocity component,
    forced by dp/dx_variable)."""
    u = np.array(u, dtype=np.float64, order="C")
    N = u.shape[0] - 1
    b = np.ascontiguousarray(band, dtype=np.int8)
    o = np.ascontiguousarray(outside, dtype=np.uint8)
    pp = None if p is None else f64(p)
    check(_lib.lib().hor_grid_solve_poisson(N, float(L), b.ctypes.data_as(C.c_void_p),
                                            o.ctypes.data_as(C.c_void_p), ptr(pp), ptr(u), int(variable)),
          "hor_grid_solve_poisson")
    return u


def d4(f, h, axis):
    """4th-order first difference along axis (main.cpp:628-900): central
    inside, one-sided at the first and last two nodes."""
    F = np.moveaxis(np.asarray(f, dtype=np.float64), axis, 0)
    N = F.shape[0] - 1
    out = np.empty_like(F)
    out[2:N - 1] = (-F[4:] + 8.0 * F[3:-1] - 8.0 * F[1:-3] + F[:-4]) / (12.0 * h)
    for i in (0, 1):
        out[i] = (-(25.0 / 12.0) * F[i] + 4.0 * F[i + 1] - 3.0 * F[i + 2] + (4.0 / 3.0) * F[i + 3]
                  - 0.25 * F[i + 4]) / h
    for i in (N - 1, N):
        out[i] = ((25.0 / 12.0) * F[i] - 4.0 * F[i - 1] + 3.0 * F[i - 2] - (4.0 / 3.0) * F[i - 3]
                  + 0.25 * F[i - 4]) / h
    return np.moveaxis(out, 0, axis)


def gradient_error(u, u_ex, h):
    """|grad u - grad u_exact| per node, both gradients by d4 (main.cpp:615-927):
    sqrt of the sum over the 9 components."""
    e = np.zeros(u.shape[1:])
    for c in range(3):
        for ax in range(3):
            e += (d4(u[c], h, ax) - d4(u_ex[c], h, ax)) ** 2
    return np.sqrt(e)


def _faces(N):
    i = np.arange(N + 1)
    I, J, K = np.meshgrid(i, i, i, indexing="ij")
    return (I == 0) | (I == N) | (J == 0) | (J == N) | (K == 0) | (K == N)


def grid_program(ctx, N=96, L=3.0, abc=(1.0, 0.5, 0.5), band1=4.0, band2=2.0, ctr=1.5, log=None):
    """The reference's test_Stokes_Poisson3d_FD (main.cpp:107-968) for the
    prolate spheroid abc centred at (ctr, ctr, ctr) translating with U = (1,0,0):
    returns the grid pressure and velocity, their exact values and the errors
    the reference prints (max and l2 over the grid, over outside + on, over
    outside only)."""
    h = L / N
    DEL = G.delta(h)
    t = {}
    t0 = time.perf_counter()
    body = G.Body(abc, (ctr, ctr, ctr))
    s = body.surface(h, ctx=ctx)
    s["f"] = body.traction(s["x"])                 # stokeslet_density (utilities.cpp:267-285)
    t["surface_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    bt = body.band_targets(s, (0.0, 0.0, 0.0), h, N, band1 * h)   # |b| <= BAND1 h + 1e-12
    t["band_s"] = time.perf_counter() - t0
    n1 = N + 1
    flat = (bt["ijk"][:, 0] * n1 + bt["ijk"][:, 1]) * n1 + bt["ijk"][:, 2]
    band = np.zeros(n1 ** 3, np.int8)
    band[flat] = 1
    band[flat[np.abs(bt["b"]) <= band2 * h + 1e-12]] = 2
    g = np.arange(n1) * h
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    phi = level_set(body, X.ravel())
    outside = (phi > 0).astype(np.uint8)
    faces = _faces(N).ravel()
    # band nodes outside or on the surface (main.cpp:1140, 1323)
    sel = (bt["b"] > 0.0) | (np.abs(bt["b"]) < 1.e-12)
    bx = bt["x"].reshape(-1, 3)[sel].ravel()
    tb = dict(n=int(sel.sum()), x=bx, b=bt["b"][sel], x0=bt["x0"].reshape(-1, 3)[sel].ravel(),
              n0=bt["n0"].reshape(-1, 3)[sel].ravel())
    tb["f0"] = body.traction(tb["x0"])
    fx = X[faces].ravel()
    # stage 1: pressure
    t0 = time.perf_counter()
    p = np.zeros(n1 ** 3)
    p[faces] = ops.stokes_pressure_away(ctx, s, fx)
    p[flat[sel]] = ops.stokes_pressure_7ord(ctx, DEL, s, tb)
    t["pressure_bi_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    p = solve_poisson(L, band, outside, None, p.reshape(n1, n1, n1), 0).ravel()
    t["pressure_poisson_s"] = time.perf_counter() - t0
    # stage 2: velocity; inside the body U (main.cpp:1236-1238)
    t0 = time.perf_counter()
    u = np.zeros((3, n1 ** 3))
    u[0] = 1.0
    far = dict(n=int(faces.sum()), x=fx, b=np.full(int(faces.sum()), 1e300),
               n0=np.zeros(fx.size), f0=np.zeros(fx.size))
    uf = ops.Plan(ctx, ops.SL_OFF, s, far, 1.0)(s["f"], far["f0"]).reshape(-1, 3)     # stokes_SL_away
    ub = ops.Plan(ctx, ops.SL_OFF | _lib.ALL_NEAR, s, tb, DEL)(s["f"], tb["f0"]).reshape(-1, 3)  # stokes_SL_7ord
    for c in range(3):
        u[c, faces] = uf[:, c]
        u[c, flat[sel]] = ub[:, c]
    t["velocity_bi_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    P3 = p.reshape(n1, n1, n1)
    for c in range(3):
        u[c] = solve_poisson(L, band, outside, P3, u[c].reshape(n1, n1, n1), c + 1).ravel()
    t["velocity_poisson_s"] = time.perf_counter() - t0
    # exact solution: outside and on the surface the Chwang-Wu flow, inside U and p = 0
    on = np.abs(phi) < 1.e-12
    keep = (phi > 0) | on
    p_ex = np.zeros(n1 ** 3)
    p_ex[keep] = exact_pressure(body, X[keep].ravel())
    u_ex = np.zeros((3, n1 ** 3))
    u_ex[0] = 1.0
    ue = np.zeros(3 * int(keep.sum()))
    check(_lib.lib().hor_spheroid_exact_velocity(C.byref(body.c), int(keep.sum()), ptr(f64(X[keep].ravel())),
                                                 ptr(ue)), "hor_spheroid_exact_velocity")
    u_ex[:, keep] = ue.reshape(-1, 3).T
    ep = np.abs(p - p_ex)
    eu = np.sqrt(((u - u_ex) ** 2).sum(0))
    out_only = phi > 0
    err = {
        "p_max_grid": float(ep.max()), "p_l2_grid": float(np.sqrt((ep ** 2).mean())),
        "p_max_outside_on": float(ep[keep].max()), "p_l2_outside_on": float(np.sqrt((ep[keep] ** 2).mean())),
        "p_max_outside": float(ep[out_only].max()),
        "u_max_grid": float(eu.max()), "u_l2_grid": float(np.sqrt((eu ** 2).mean())),
        "u_max_outside_on": float(eu[keep].max()), "u_l2_outside_on": float(np.sqrt((eu[keep] ** 2).mean())),
        "p_exact_max": float(np.abs(p_ex).max()), "u_exact_max": float(np.sqrt((u_ex ** 2).sum(0)).max()),
    }
    eg = gradient_error(u.reshape(3, n1, n1, n1), u_ex.reshape(3, n1, n1, n1), h)
    err["gradu_max_grid"] = float(eg.max())
    err["gradu_l2_grid"] = float(np.sqrt((eg ** 2).mean()))
    if log:
        log(f"grid N={N}: sources {s['n']}, band nodes {bt['n']}, times {t}, errors {err}")
    return dict(N=N, h=h, DEL=DEL, p=p.reshape(n1, n1, n1), u=u.reshape(3, n1, n1, n1),
                p_exact=p_ex.reshape(n1, n1, n1), u_exact=u_ex.reshape(3, n1, n1, n1),
                phi=phi.reshape(n1, n1, n1), band=band.reshape(n1, n1, n1), errors=err, times=t,
                sources=s["n"], band_nodes=bt["n"])
