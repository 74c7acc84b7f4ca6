"""hor_dl_onsurf_dev_noalloc (the allocation-free on-surface stresslet sum)
against hor_dl_onsurf through the C ABI, to the FP64 parity bar (1e-12
relative, 2-norm and max-norm): two calls in a row on the same GPU, the
second on a smaller input, so the second reuses the device array the first
left behind; sizes beyond the array are rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

import hor
from cases import ellipsoid_case
from conftest import rel_l2, rel_max
from hor import _lib

pytestmark = pytest.mark.gpu


def _noalloc(ctx, DEL, s, nt):
    L = _lib.lib()
    dev = {}
    try:
        for k in ("x", "nrml", "area", "g"):
            a = np.ascontiguousarray(s[k], dtype=np.float64)
            dev[k] = ctx.alloc(a.nbytes)
            ctx.upload(dev[k], a)
        out = ctx.alloc(24 * nt)
        dev["out"] = out
        rc = L.hor_dl_onsurf_dev_noalloc(ctx.h, C.c_double(DEL), s["n"], dev["x"], dev["nrml"], dev["area"],
                                         dev["g"], nt, dev["x"], out)
        _lib.check(rc, "hor_dl_onsurf_dev_noalloc")
        return ctx.download(out, np.zeros(3 * nt))
    finally:
        for p in dev.values():
            ctx.free(p)


def _reference(ctx, DEL, s, nt):
    out = np.zeros(3 * nt)
    x = np.ascontiguousarray(s["x"], dtype=np.float64)
    _lib.check(_lib.lib().hor_dl_onsurf(ctx.h, C.c_double(DEL), s["n"], _lib.ptr(x), _lib.ptr(s["nrml"]),
                                        _lib.ptr(s["area"]), _lib.ptr(s["g"]), nt, _lib.ptr(x), _lib.ptr(out)),
               "hor_dl_onsurf")
    return out


def test_noalloc_matches_dl_onsurf_twice(ctx):
    for h, frac in ((1.0 / 32, 1.0), (1.0 / 16, 0.6)):
        body, s, t, DEL = ellipsoid_case(h=h, max_band=10)
        nt = max(1, int(frac * s["n"]))
        print(f"noalloc DL_ON: ellipsoid h=1/{round(1 / h)}, ns={s['n']}, nt={nt}", flush=True)
        got = _noalloc(ctx, DEL, s, nt)
        want = _reference(ctx, DEL, s, nt)
        e2, em = rel_l2(got, want), rel_max(got, want)
        print(f"  rel l2 {e2:.2e}, rel max {em:.2e}", flush=True)
        assert np.all(np.isfinite(got))
        assert e2 <= 1e-12 and em <= 1e-12


def test_noalloc_rejects_oversize(ctx):
    L = _lib.lib()
    one = ctx.alloc(8)
    try:
        rc = L.hor_dl_onsurf_dev_noalloc(ctx.h, C.c_double(0.1), 200000, one, one, one, one, 200000, one, one)
        assert rc == -1
        assert b"scratch" in L.hor_last_error()
        rc = L.hor_dl_onsurf_dev_noalloc(ctx.h, C.c_double(0.1), 100, one, one, one, one, 101, one, one)
        assert rc == -1
    finally:
        ctx.free(one)
