"""Operator API over libhor.so: context, plans, and the reference's functions.

The module-level functions carry the reference's names and argument meaning
(stokeslets.h:9-29, stresslets.h:9-30, KITC.h:99-118, and the grid program's
stokes.h:20-28 for the pressure) with surfaces/targets as dicts (see
hor.geometry).  Everything runs on the GPU; a missing library or
device raises instead of falling back to the CPU.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib
from ._lib import (DIRECT, DL_OFF, DL_ON, FAR_MFMA, FAR_VALU, FP64, MIXED, P_OFF, SL_OFF, SLDL_OFF, SL_ON, TREE,  # noqa: F401
                   check, f64, ptr)

__all__ = ["Context", "Plan", "TreeParams", "SL_ON", "SL_OFF", "DL_ON", "DL_OFF", "P_OFF", "SLDL_OFF", "DIRECT", "TREE",
           "FP64", "MIXED", "eval_Stokes_SL_offSurf", "eval_Stokes_SL_onSurf",
           "eval_Stokes_DL_offSurf", "eval_Stokes_DL_onSurf", "eval_Stokes_SL_onSurf_tree",
           "eval_Stokes_SL_offSurf_tree", "eval_Stokes_DL_onSurf_tree", "eval_Stokes_DL_offSurf_tree",
           "stokes_pressure_7ord", "stokes_pressure_away", "device_count"]


def device_count():
    n = C.c_int()
    check(_lib.lib().hor_device_count(C.byref(n)), "hor_device_count")
    return n.value


def TreeParams(P=6, N0=1000, theta=0.6):
    """Treecode parameters (KITC.h:14-17 defaults)."""
    t = _lib.TreeParams()
    t.P, t.N0, t.theta = int(P), int(N0), float(theta)
    return t


class Context:
    """One GPU, one HIP stream.  Objects created on the context (plans,
    solvers, networks) are closed before the context is destroyed."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(_lib.lib().hor_ctx_create(int(device), C.byref(h)), "hor_ctx_create")
        self.h = h
        self._children = weakref.WeakSet()

    def adopt(self, obj):
        """Registers obj (with a close() method) to be closed with the context."""
        self._children.add(obj)
        return obj

    def close(self):
        if self.h:
            for obj in list(self._children):
                obj.close()
            _lib.lib().hor_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_timing(self):
        """(total ms, dominant-kernel ms) of the last evaluation (HIP events)."""
        a, b = C.c_double(), C.c_double()
        check(_lib.lib().hor_ctx_last_timing(self.h, C.byref(a), C.byref(b)), "hor_ctx_last_timing")
        return a.value, b.value

    # device memory helpers
    def alloc(self, nbytes):
        p = C.c_void_p()
        check(_lib.lib().hor_dev_alloc(self.h, int(nbytes), C.byref(p)), "hor_dev_alloc")
        return p

    def free(self, p):
        check(_lib.lib().hor_dev_free(self.h, p), "hor_dev_free")

    def upload(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        check(_lib.lib().hor_dev_upload(self.h, dst, ptr(arr), arr.nbytes), "hor_dev_upload")

    def download(self, src, out):
        check(_lib.lib().hor_dev_download(self.h, ptr(out), src, out.nbytes), "hor_dev_download")
        return out


class Plan:
    """Geometry-bound operator (hor_plan_create)."""

    def __init__(self, ctx, kind, src, tgt=None, DEL=None, method=DIRECT, precision=FP64,
                 ctr=None, tree=None):
        ns = src["n"]
        on = kind in (SL_ON, DL_ON)
        if on:
            nt = ns if tgt is None else tgt["n"]
            tx = src["x"] if tgt is None else tgt["x"]
            tb = tx0 = tn0 = None
        else:
            nt = tgt["n"]
            tx, tb, tx0, tn0 = tgt["x"], tgt["b"], tgt.get("x0"), tgt["n0"]
        self._keep = [f64(src["x"], 3 * ns), f64(src["nrml"], 3 * ns), f64(src["area"], ns),
                      f64(tx, 3 * nt), f64(tb), f64(tx0), f64(tn0)]
        k = self._keep
        c = None
        if ctr is not None:
            c = f64(np.asarray(ctr), 3)
            self._keep.append(c)
        tp = None if tree is None else C.byref(tree)
        h = C.c_void_p()
        check(_lib.lib().hor_plan_create(ctx.h, int(kind), int(method), int(precision), float(DEL), ns,
                                         ptr(k[0]), ptr(k[1]), ptr(k[2]), nt, ptr(k[3]), ptr(k[4]),
                                         ptr(k[5]), ptr(k[6]), ptr(c), tp, C.byref(h)), "hor_plan_create")
        self.h, self.ctx, self.kind, self.ns, self.nt = h, ctx, kind, ns, nt
        ctx.adopt(self)

    def close(self):
        if self.h:
            _lib.lib().hor_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __call__(self, dens, d0=None):
        """SLDL_OFF: dens = [f | g] (6 ns), d0 = [f0 | g0] (6 nt); returns
        [SL | DL] (6 nt)."""
        per = 6 if self.kind == SLDL_OFF else 3
        dens = f64(dens, per * self.ns)
        d0 = None if d0 is None else f64(d0, per * self.nt)
        out = np.zeros((1 if self.kind == P_OFF else per) * self.nt)
        check(_lib.lib().hor_plan_eval(self.h, ptr(dens), ptr(d0), ptr(out)), "hor_plan_eval")
        return out

    def eval_dev(self, dens_dev, d0_dev, out_dev):
        check(_lib.lib().hor_plan_eval_dev(self.h, dens_dev, d0_dev, out_dev), "hor_plan_eval_dev")

    def set_far_kernel(self, kernel):
        """FAR_VALU or FAR_MFMA (hor_plan_set_far_kernel): the far-field kernel
        of a mixed-precision single-layer treecode plan."""
        check(_lib.lib().hor_plan_set_far_kernel(self.h, int(kernel)), "hor_plan_set_far_kernel")
        return self

    def phase_timing(self):
        """Kernel ms of the last evaluation's moments, near sums and far field
        (treecode plans; hor_plan_phase_timing)."""
        v = [C.c_double() for _ in range(3)]
        check(_lib.lib().hor_plan_phase_timing(self.h, *[C.byref(x) for x in v]), "hor_plan_phase_timing")
        return dict(moments_ms=v[0].value, near_ms=v[1].value, far_ms=v[2].value)

    def stats(self):
        v = [C.c_double() for _ in range(4)]
        check(_lib.lib().hor_plan_stats(self.h, *[C.byref(x) for x in v]), "hor_plan_stats")
        return dict(direct_pairs=v[0].value, near_pairs=v[1].value, proxy_pairs=v[2].value,
                    moment_terms=v[3].value)


# ---------------------------------------------------------------------------
# The reference's entry points (one-shot: plan, evaluate, release).

def eval_Stokes_SL_offSurf(ctx, h, DEL, surf, targets):
    """stokeslets.cpp:12-86; density surf['f'], nearest force targets['f0']."""
    return Plan(ctx, SL_OFF, surf, targets, DEL)(surf["f"], targets["f0"])


def eval_Stokes_SL_onSurf(ctx, DEL, surf, targets=None):
    """stokeslets.cpp:141-180; target i is quadrature point i."""
    return Plan(ctx, SL_ON, surf, targets, DEL)(surf["f"])


def eval_Stokes_DL_offSurf(ctx, h, DEL, surf, targets):
    """stresslets.cpp:13-96; density surf['g'], nearest density targets['g0']."""
    return Plan(ctx, DL_OFF, surf, targets, DEL)(surf["g"], targets["g0"])


def eval_Stokes_SL_DL_offSurf(ctx, h, DEL, surf, targets, precision=FP64):
    """eval_Stokes_SL_offSurf and eval_Stokes_DL_offSurf at the same targets in
    one pass (HOR_SLDL_OFF); returns (SL, DL), each 3*nt."""
    n = targets["n"]
    out = Plan(ctx, SLDL_OFF, surf, targets, DEL, precision=precision)(
        np.concatenate([surf["f"], surf["g"]]), np.concatenate([targets["f0"], targets["g0"]]))
    return out[:3 * n], out[3 * n:]


def eval_Stokes_DL_onSurf(ctx, DEL, surf, targets=None):
    """stresslets.cpp:186-228."""
    return Plan(ctx, DL_ON, surf, targets, DEL)(surf["g"])


def eval_Stokes_SL_onSurf_tree(ctx, ctr, h, DEL, surf, targets=None, tree=None):
    """KITC.cpp:1389-1492; ctr = centre of the source surface's root box."""
    return Plan(ctx, SL_ON, surf, targets, DEL, TREE, ctr=ctr, tree=tree or TreeParams())(surf["f"])


def eval_Stokes_SL_offSurf_tree(ctx, ctr, h, DEL, surf, targets, tree=None):
    """KITC.cpp:1496-1602."""
    return Plan(ctx, SL_OFF, surf, targets, DEL, TREE, ctr=ctr, tree=tree or TreeParams())(surf["f"], targets["f0"])


def eval_Stokes_DL_onSurf_tree(ctx, ctr, h, DEL, surf, targets=None, tree=None):
    """KITC.cpp:1606-1709."""
    return Plan(ctx, DL_ON, surf, targets, DEL, TREE, ctr=ctr, tree=tree or TreeParams())(surf["g"])


def eval_Stokes_DL_offSurf_tree(ctx, ctr, h, DEL, surf, targets, tree=None):
    """KITC.cpp:1712-1830."""
    return Plan(ctx, DL_OFF, surf, targets, DEL, TREE, ctr=ctr, tree=tree or TreeParams())(surf["g"], targets["g0"])


def stokes_pressure_7ord(ctx, DEL, surf, targets):
    """stokes_pressure_7ord (extend_to_grid/stokes.cpp:134-203) at every target:
    surf['f'] the surface force, targets['f0'] the force at the nearest point;
    one pressure per target."""
    return Plan(ctx, P_OFF, surf, targets, DEL)(surf["f"], targets["f0"])


def stokes_pressure_away(ctx, surf, tx):
    """stokes_pressure_away (extend_to_grid/stokes.cpp:102-127): plain sum at
    targets away from the surface."""
    tx = f64(tx)
    nt = tx.size // 3
    out = np.zeros(nt)
    f, x, nr, a = f64(surf["f"]), f64(surf["x"]), f64(surf["nrml"]), f64(surf["area"])
    check(_lib.lib().hor_pressure_away(ctx.h, surf["n"], ptr(x), ptr(nr), ptr(a), ptr(f), nt, ptr(tx), ptr(out)),
          "hor_pressure_away")
    return out
