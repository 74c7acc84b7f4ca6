"""Two-surface second-kind solve (highreg_stokes_2surf/main.cpp:140-677) on the GPU.

``Solver`` binds two surfaces (dicts with x, nrml, area, ctr and the nearest
data of each surface's points relative to the other surface: nrst, x0, n0, b,
i123) and provides the matvec (solve_IE_1step), the right-hand side
(eval_SL), the stencil interpolation (interpolate_DL_density) and restarted
GMRES with every vector resident in HBM.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DIRECT, FP64, check, f64, i32, ptr


def _desc(s, keep):
    d = _lib.SurfaceDesc()
    n = int(s["n"])
    d.n = n
    arrs = dict(x=f64(s["x"], 3 * n), nrml=f64(s["nrml"], 3 * n), area=f64(s["area"], n),
                nrst=i32(s["nrst"], n), x0=f64(s["x0"], 3 * n), n0=f64(s["n0"], 3 * n),
                b=f64(s["b"], n), i123=i32(s["i123"], 3 * n))
    keep.extend(arrs.values())
    for k, v in arrs.items():
        setattr(d, k, ptr(v))
    for k in range(3):
        d.ctr[k] = float(s["ctr"][k])
    return d


class Solver:
    def __init__(self, ctx, h, DEL, s1, s2, method=DIRECT, precision=FP64, tree1=None, tree2=None):
        self._keep = []
        self.d1, self.d2 = _desc(s1, self._keep), _desc(s2, self._keep)
        t1 = None if tree1 is None else C.byref(tree1)
        t2 = None if tree2 is None else C.byref(tree2)
        h_ = C.c_void_p()
        check(_lib.lib().hor_solver_create(ctx.h, float(h), float(DEL), int(method), int(precision), t1, t2,
                                           C.byref(self.d1), C.byref(self.d2), C.byref(h_)),
              "hor_solver_create")
        self.h, self.ctx = h_, ctx
        ctx.adopt(self)
        n = C.c_int()
        check(_lib.lib().hor_solver_size(self.h, C.byref(n)), "hor_solver_size")
        self.n = n.value
        self.n1, self.n2 = int(s1["n"]), int(s2["n"])

    def close(self):
        if self.h:
            _lib.lib().hor_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        """Work of one matvec (hor_solver_stats)."""
        v = [C.c_double() for _ in range(4)]
        check(_lib.lib().hor_solver_stats(self.h, *[C.byref(x) for x in v]), "hor_solver_stats")
        return dict(direct_pairs=v[0].value, near_pairs=v[1].value, proxy_pairs=v[2].value,
                    moment_terms=v[3].value)

    def plan_stats(self, which):
        """Work of one DL plan: 0/1 on-surface (DL11, DL22), 2/3 cross-surface
        (DL12, DL21) (hor_solver_plan_stats)."""
        v = [C.c_double() for _ in range(4)]
        check(_lib.lib().hor_solver_plan_stats(self.h, int(which), *[C.byref(x) for x in v]),
              "hor_solver_plan_stats")
        return dict(direct_pairs=v[0].value, near_pairs=v[1].value, proxy_pairs=v[2].value,
                    moment_terms=v[3].value)

    def phase_timing(self):
        """Kernel ms of the last treecode matvec by phase: moments, near sums,
        far field (hor_solver_phase_timing)."""
        v = [C.c_double() for _ in range(3)]
        check(_lib.lib().hor_solver_phase_timing(self.h, *[C.byref(x) for x in v]), "hor_solver_phase_timing")
        return dict(moments_ms=v[0].value, near_ms=v[1].value, far_ms=v[2].value)

    def matvec(self, u):
        u = f64(u, self.n)
        out = np.zeros(self.n)
        check(_lib.lib().hor_solver_matvec(self.h, ptr(u), ptr(out)), "hor_solver_matvec")
        return out

    def matvec_dev(self, u_dev, out_dev):
        check(_lib.lib().hor_solver_matvec_dev(self.h, u_dev, out_dev), "hor_solver_matvec_dev")

    def interp_g0(self, k, u_other):
        """g0 at the points of surface k (1 or 2) from the other surface's density."""
        n_other = self.n2 if k == 1 else self.n1
        n_me = self.n1 if k == 1 else self.n2
        u = f64(u_other, 3 * n_other)
        out = np.zeros(3 * n_me)
        check(_lib.lib().hor_solver_interp_g0(self.h, int(k), ptr(u), ptr(out)), "hor_solver_interp_g0")
        return out

    def eval_sl(self, f1, f2, f01, f02):
        out = np.zeros(self.n)
        a = [f64(f1, 3 * self.n1), f64(f2, 3 * self.n2), f64(f01, 3 * self.n1), f64(f02, 3 * self.n2)]
        check(_lib.lib().hor_solver_eval_sl(self.h, *[ptr(x) for x in a], ptr(out)), "hor_solver_eval_sl")
        return out

    def fgmres(self, b, m=15, tol=1e-10, maxit=100, precond=None, log=None):
        """Flexible GMRES(m) with an optional right preconditioner (a
        ``hor.mpnn.Network`` bound to this geometry).  The returned residual is
        the FP64 true residual |b - A x|/|b|.  ``log`` = (z_dev, az_dev,
        log_max) device buffers receive the Arnoldi pairs (z_j, A z_j).
        Returns x, iterations, residual-estimate history, residual, n_logged."""
        b = f64(b, self.n)
        x = np.zeros(self.n)
        hist = np.zeros(max(maxit, 1))
        it, res, nl = C.c_int(), C.c_double(), C.c_int()
        lz, laz, lmax = (None, None, 0) if log is None else log
        check(_lib.lib().hor_solver_fgmres(self.h, ptr(b), ptr(x), int(m), float(tol), int(maxit),
                                           None if precond is None else precond.h, C.byref(it), C.byref(res),
                                           ptr(hist), lz, laz, int(lmax), C.byref(nl)), "hor_solver_fgmres")
        return x, it.value, hist[: it.value], res.value, nl.value

    def set_precond_policy(self, force_plain=False, gate=True, gate_ratio=0.5):
        """How fgmres uses a network (hor_solver_set_precond_policy):
        force_plain ignores it; gate decides per restart cycle from the
        cycle's first (plain) Arnoldi pair whether the network is used."""
        p = _lib.PrecondPolicy(int(bool(force_plain)), int(bool(gate)), float(gate_ratio))
        check(_lib.lib().hor_solver_set_precond_policy(self.h, C.byref(p)), "hor_solver_set_precond_policy")

    def last_info(self):
        """Record of the last gmres/fgmres call (hor_solver_fgmres_info)."""
        info = _lib.FgmresInfo()
        check(_lib.lib().hor_solver_fgmres_info(self.h, C.byref(info)), "hor_solver_fgmres_info")
        return {k: getattr(info, k) for k, _ in _lib.FgmresInfo._fields_}

    def gmres(self, b, m=15, tol=1e-10, maxit=100):
        b = f64(b, self.n)
        x = np.zeros(self.n)
        hist = np.zeros(max(maxit, 1))
        it, res = C.c_int(), C.c_double()
        check(_lib.lib().hor_solver_gmres(self.h, ptr(b), ptr(x), int(m), float(tol), int(maxit), C.byref(it),
                                          C.byref(res), ptr(hist)), "hor_solver_gmres")
        return x, it.value, hist[: it.value], res.value


SEED = 20261203


def two_sphere_problem(h, gap, seed=SEED, ctx=None, radius=0.5):
    """The two-sphere workload of the second-kind solve: spheres of the given
    radius with centres (0, 0, -+(2 radius + gap)/2), the solver's surface
    dicts (nearest data of each surface's points relative to the other
    surface, the surface force f and its value f0 at each nearest point,
    surface.cpp:450-559) and a seeded rigid-body right-hand side
    b = U_k + W_k x (x - c_k) with U, W ~ U(-1, 1) per sphere (numpy PCG64).
    The reference's own right-hand side is b = -2/3 SL[f] (main.cpp:199-207),
    Solver.eval_sl(f1, f2, f01, f02).  Quadrature and nearest data are built
    on the GPU of ctx when given, else on the host (identical outputs).
    Returns DEL, the two bodies, the two surface dicts and the rigid b."""
    from . import geometry as G
    DEL = G.delta(h)
    zc = radius + gap / 2
    bodies = [G.Body((radius,) * 3, (0.0, 0.0, -zc)), G.Body((radius,) * 3, (0.0, 0.0, zc))]
    surfs = [b.surface(h, ctx=ctx) for b in bodies]
    pd = []
    for k in range(2):
        t = bodies[1 - k].nearest(h, DEL, surfs[1 - k], surfs[k]["x"], with_force=True, ctx=ctx)
        pd.append(dict(surfs[k], ctr=bodies[k].ctr, nrst=t["nrst"], x0=t["x0"], n0=t["n0"], b=t["b"],
                       i123=t["i123"], f=bodies[k].force(surfs[k]["x"], surfs[k]["nrml"]), f0=t["f0"]))
    rng = np.random.default_rng(seed)
    rigid = []
    for k in range(2):
        U, Wv = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        x = pd[k]["x"].reshape(-1, 3)
        rigid.append((U + np.cross(Wv, x - bodies[k].ctr)).ravel())
    return DEL, bodies, pd, np.concatenate(rigid)
