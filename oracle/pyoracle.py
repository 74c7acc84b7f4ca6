"""ctypes bindings for the checker libraries (TEST INFRASTRUCTURE ONLY).

* ``Oracle``  - liboracle.so, the C restatement of the reference algorithms.
* ``Ref2``    - _ref/libref2surf.so, the reference's highreg_stokes_2surf sources
                compiled unmodified (with a packing harness).
* ``RefGrid`` - _ref/librefgrid.so, highreg_stokes_extend_to_grid's utilities,
                surface and stokes sources.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this module.  Arrays are float64 / int32, xyz interleaved (3*n).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_D = C.c_double
_I = C.c_int
_P = C.c_void_p


def _p(a):
    """Pointer to a contiguous numpy array (or NULL for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(_P)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


class _TreeParams(C.Structure):
    _fields_ = [("P", _I), ("N0", _I), ("theta", _D), ("ctr", _D * 3)]


class _Surface(C.Structure):
    _fields_ = [
        ("n", _I), ("x", _P), ("nrml", _P), ("area", _P), ("ctr", _D * 3),
        ("nrst", _P), ("x0", _P), ("n0", _P), ("b", _P), ("i123", _P),
        ("f", _P), ("f0", _P),
    ]


class _Problem(C.Structure):
    _fields_ = [("h", _D), ("DEL", _D), ("use_tree", _I),
                ("tree", _TreeParams * 2), ("s", _Surface * 2)]


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run __graft_entry__.build())")
    return C.CDLL(path)


class Oracle:
    """The C restatement (oracle/hor_oracle.c)."""

    def __init__(self, path=None):
        self.lib = _load(path or os.path.join(HERE, "liboracle.so"))
        L = self.lib
        for name in ("orc_sl_off", "orc_sl_on", "orc_dl_off", "orc_dl_on",
                     "orc_sl_on_tree", "orc_sl_off_tree", "orc_dl_on_tree",
                     "orc_dl_off_tree", "orc_stencil", "orc_interp_g0",
                     "orc_matvec", "orc_eval_sl", "orc_reg_coeff", "orc_In_024",
                     "orc_pressure_off", "orc_pressure_away", "orc_lapl_dl_kernel",
                     "orc_dst_poisson3d", "orc_extend_into_cube", "orc_lapl_in_cube",
                     "orc_lapl_near_surface", "orc_dp_4ord", "orc_solve_poisson"):
            getattr(L, name).restype = None
        L.orc_gmres.restype = _I
        L.orc_tree_structure.restype = _I
        L.orc_tree_moments.restype = _I

    # ---- kernels
    def In_024(self, lam):
        out = np.zeros(3)
        self.lib.orc_In_024(_D(lam), _p(out))
        return out

    def reg_coeff(self, lam):
        out = np.zeros(3)
        self.lib.orc_reg_coeff(_D(lam), _p(out))
        return out

    def kernel(self, kind, r, d, c=(0.0, 0.0, 0.0)):
        """kind: 'sl', 'sl_on', 'dl', 'dl_on', 'lapl_dl' -> array (len(r), 2)."""
        L = self.lib
        cc = _f64(np.asarray(c))
        out = np.zeros((len(r), 2))
        a, b = _D(), _D()
        for i, rv in enumerate(np.asarray(r, dtype=np.float64)):
            if kind == "sl":
                L.orc_sl_kernel(_D(rv), _D(d), _p(cc), C.byref(a), C.byref(b))
            elif kind == "sl_on":
                L.orc_sl_kernel_on(_D(rv), _D(d), C.byref(a), C.byref(b))
            elif kind == "dl":
                L.orc_dl_kernel(_D(rv), _D(d), _p(cc), C.byref(a), C.byref(b))
            elif kind == "lapl_dl":
                a.value = 0.0
                L.orc_lapl_dl_kernel(_D(rv), _D(d), _p(cc), C.byref(b))
            else:
                a.value = 0.0
                L.orc_dl_kernel_on(_D(rv), _D(d), C.byref(b))
            out[i] = (a.value, b.value)
        return out

    # ---- direct sums
    def sl_off(self, DEL, src, tgt):
        out = np.zeros(3 * tgt["n"])
        self.lib.orc_sl_off(_D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]), _p(src["area"]),
                            _p(src["f"]), tgt["n"], _p(tgt["x"]), _p(tgt["b"]), _p(tgt["n0"]),
                            _p(tgt["f0"]), _p(out))
        return out

    def sl_on(self, DEL, src, nt=None):
        nt = src["n"] if nt is None else nt
        out = np.zeros(3 * nt)
        self.lib.orc_sl_on(_D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]), _p(src["area"]),
                           _p(src["f"]), nt, _p(src["x"]), _p(out))
        return out

    def dl_off(self, DEL, src, tgt):
        out = np.zeros(3 * tgt["n"])
        self.lib.orc_dl_off(_D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]), _p(src["area"]),
                            _p(src["g"]), tgt["n"], _p(tgt["x"]), _p(tgt["b"]), _p(tgt["x0"]),
                            _p(tgt["n0"]), _p(tgt["g0"]), _p(out))
        return out

    def dl_on(self, DEL, src, nt=None):
        nt = src["n"] if nt is None else nt
        out = np.zeros(3 * nt)
        self.lib.orc_dl_on(_D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]), _p(src["area"]),
                           _p(src["g"]), nt, _p(src["x"]), _p(out))
        return out

    # ---- pressure
    def pressure_off(self, DEL, src, tgt):
        """stokes_pressure_7ord at each target; src['f'] the force, tgt['f0'] at x0."""
        out = np.zeros(tgt["n"])
        self.lib.orc_pressure_off(_D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]), _p(src["area"]),
                                  _p(src["f"]), tgt["n"], _p(tgt["x"]), _p(tgt["b"]), _p(tgt["n0"]),
                                  _p(tgt["f0"]), _p(out))
        return out

    def pressure_away(self, src, tx):
        tx = _f64(tx)
        out = np.zeros(len(tx) // 3)
        self.lib.orc_pressure_away(src["n"], _p(src["x"]), _p(src["area"]), _p(src["f"]), len(out), _p(tx),
                                   _p(out))
        return out

    # ---- grid extension (arrays are flat numpy, (N+1)^3 or (N-1)^3)
    def dst_poisson3d(self, f):
        f = np.array(f, dtype=np.float64)
        n = round(len(f) ** (1 / 3))
        self.lib.orc_dst_poisson3d(int(n), _p(f))
        return f

    def extend_into_cube(self, N, L, g):
        G = np.zeros((N + 1) ** 3)
        self.lib.orc_extend_into_cube(int(N), _D(L), _D(L / N), _p(_f64(g)), _p(G))
        return G

    def lapl_in_cube(self, N, G):
        rhs = np.zeros((N - 1) ** 3)
        self.lib.orc_lapl_in_cube(int(N), _p(_f64(G)), _p(rhs))
        return rhs

    def lapl_near_surface(self, N, band, u, rhs):
        rhs = np.array(rhs, dtype=np.float64)
        b = np.ascontiguousarray(band, dtype=np.int8)
        self.lib.orc_lapl_near_surface(int(N), b.ctypes.data_as(C.c_void_p), _p(_f64(u)), _p(rhs))
        return rhs

    def dp_4ord(self, N, L, axis, band, outside, p, rhs):
        rhs = np.array(rhs, dtype=np.float64)
        b = np.ascontiguousarray(band, dtype=np.int8)
        o = np.ascontiguousarray(outside, dtype=np.uint8)
        self.lib.orc_dp_4ord(int(N), _D(L / N), int(axis), b.ctypes.data_as(C.c_void_p),
                             o.ctypes.data_as(C.c_void_p), _p(_f64(p)), _p(rhs))
        return rhs

    def solve_poisson(self, N, L, band, outside, p, u, variable):
        u = np.array(u, dtype=np.float64)
        b = np.ascontiguousarray(band, dtype=np.int8)
        o = np.ascontiguousarray(outside, dtype=np.uint8)
        pp = _f64(p) if p is not None else np.zeros(1)
        self.lib.orc_solve_poisson(int(N), _D(L), _D(L / N), b.ctypes.data_as(C.c_void_p),
                                   o.ctypes.data_as(C.c_void_p), _p(pp), _p(u), int(variable))
        return u

    # ---- treecode
    @staticmethod
    def tree_params(P, N0, theta, ctr):
        tp = _TreeParams()
        tp.P, tp.N0, tp.theta = int(P), int(N0), float(theta)
        for k in range(3):
            tp.ctr[k] = float(ctr[k])
        return tp

    def sl_on_tree(self, tp, DEL, src, nt=None):
        """On-surface treecode SL at the first nt source points (all by default)."""
        nt = src["n"] if nt is None else int(nt)
        out = np.zeros(3 * nt)
        self.lib.orc_sl_on_tree(C.byref(tp), _D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]),
                                _p(src["area"]), _p(src["f"]), nt, _p(src["x"]), _p(out))
        return out

    def sl_off_tree(self, tp, DEL, src, tgt):
        out = np.zeros(3 * tgt["n"])
        self.lib.orc_sl_off_tree(C.byref(tp), _D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]),
                                 _p(src["area"]), _p(src["f"]), tgt["n"], _p(tgt["x"]),
                                 _p(tgt["b"]), _p(tgt["n0"]), _p(tgt["f0"]), _p(out))
        return out

    def dl_on_tree(self, tp, DEL, src, nt=None):
        """On-surface treecode DL at the first nt source points (all by default)."""
        nt = src["n"] if nt is None else int(nt)
        out = np.zeros(3 * nt)
        self.lib.orc_dl_on_tree(C.byref(tp), _D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]),
                                _p(src["area"]), _p(src["g"]), nt, _p(src["x"]), _p(out))
        return out

    def dl_off_tree(self, tp, DEL, src, tgt):
        out = np.zeros(3 * tgt["n"])
        self.lib.orc_dl_off_tree(C.byref(tp), _D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]),
                                 _p(src["area"]), _p(src["g"]), tgt["n"], _p(tgt["x"]),
                                 _p(tgt["b"]), _p(tgt["x0"]), _p(tgt["n0"]), _p(tgt["g0"]),
                                 _p(out))
        return out

    def tree_structure(self, tp, x):
        x = _f64(x)
        ns = len(x) // 3
        nn = self.lib.orc_tree_structure(C.byref(tp), ns, _p(x), 0, None, None, None, None, None, None)
        mem = np.zeros(2 * nn, np.int32)
        iv = np.zeros(6 * nn)
        ce = np.zeros(3 * nn)
        mac = np.zeros(nn)
        par = np.zeros(nn, np.int32)
        perm = np.zeros(ns, np.int32)
        self.lib.orc_tree_structure(C.byref(tp), ns, _p(x), nn, _p(mem), _p(iv), _p(ce), _p(mac),
                                    _p(par), _p(perm))
        return dict(members=mem.reshape(nn, 2), interval=iv.reshape(nn, 6),
                    centre=ce.reshape(nn, 3), mac=mac, parent=par, perm=perm)

    def tree_moments(self, tp, dl, src, dens):
        ns = src["n"]
        nn = self.lib.orc_tree_moments(C.byref(tp), int(dl), ns, _p(src["x"]), _p(src["nrml"]),
                                       _p(src["area"]), _p(dens), 0, None)
        ncomp = 12 if dl else 6
        pf = (tp.P + 1) ** 3
        out = np.zeros(nn * ncomp * pf)
        self.lib.orc_tree_moments(C.byref(tp), int(dl), ns, _p(src["x"]), _p(src["nrml"]),
                                  _p(src["area"]), _p(dens), nn, _p(out))
        return out.reshape(nn, ncomp, pf)

    # ---- stencil / interpolation
    def stencil(self, ctr, h, src_x, tgt):
        nt = tgt["n"]
        out = np.zeros(36 * nt, np.int32)
        c = _f64(np.asarray(ctr))
        self.lib.orc_stencil(_p(c), _D(h), len(src_x) // 3, _p(src_x), nt, _p(tgt["nrst"]),
                             _p(tgt["x0"]), _p(tgt["i123"]), _p(out))
        return out.reshape(nt, 36)

    def interp_g0(self, ctr, h, src_x, g, tgt):
        nt = tgt["n"]
        out = np.zeros(3 * nt)
        c = _f64(np.asarray(ctr))
        self.lib.orc_interp_g0(_p(c), _D(h), len(src_x) // 3, _p(src_x), _p(_f64(g)), nt,
                               _p(tgt["nrst"]), _p(tgt["x0"]), _p(tgt["i123"]), _p(out))
        return out

    # ---- two-surface problem
    def problem(self, h, DEL, surfs, use_tree=False, tree=((6, 1000, 0.6), (6, 1000, 0.6))):
        """surfs: two dicts with x, nrml, area, ctr, nrst, x0, n0, b, i123, f, f0
        (nearest data relative to the other surface)."""
        pb = _Problem()
        pb.h, pb.DEL, pb.use_tree = float(h), float(DEL), int(bool(use_tree))
        keep = []
        for k in range(2):
            s = surfs[k]
            P, N0, th = tree[k]
            pb.tree[k] = self.tree_params(P, N0, th, s["ctr"])
            S = pb.s[k]
            S.n = s["n"]
            for nm in ("x", "nrml", "area", "x0", "n0", "b", "f", "f0"):
                a = _f64(s.get(nm))
                keep.append(a)
                setattr(S, nm, _p(a))
            for nm in ("nrst", "i123"):
                a = _i32(s.get(nm))
                keep.append(a)
                setattr(S, nm, _p(a))
            for d in range(3):
                S.ctr[d] = float(s["ctr"][d])
        pb._keep = keep  # keep arrays alive
        return pb

    def matvec(self, pb, u):
        u = _f64(u)
        out = np.zeros_like(u)
        self.lib.orc_matvec(C.byref(pb), _p(u), _p(out))
        return out

    def eval_sl(self, pb):
        n = 3 * (pb.s[0].n + pb.s[1].n)
        out = np.zeros(n)
        self.lib.orc_eval_sl(C.byref(pb), _p(out))
        return out

    def gmres(self, pb, b, m=15, tol=1e-10, iter_max=100, fix_restart=True):
        b = _f64(b)
        x = np.zeros_like(b)
        hist = np.zeros(iter_max)
        fr = _D()
        it = self.lib.orc_gmres(C.byref(pb), len(b), _p(b), _p(x), int(m), _D(tol),
                                int(iter_max), int(bool(fix_restart)), _p(hist), C.byref(fr))
        return x, it, hist[:it], fr.value


class Ref2:
    """The compiled reference (highreg_stokes_2surf), via ref_shim_2surf.cpp."""

    def __init__(self, path=None):
        self.lib = _load(path or os.path.join(HERE, "_ref", "libref2surf.so"))
        self.lib.ref2_generate_surface.restype = _I

    def constants(self):
        out = np.zeros(14)
        self.lib.ref2_constants(_p(out))
        return dict(abc=out[0:3].copy(), ctr1=out[3:6].copy(), ctr2=out[6:9].copy(),
                    use_tree=int(out[9]), P=int(out[10]), N0=int(out[11]),
                    sq_theta=out[12], pou_theta=out[13])

    def generate_surface(self, rotated, N, h):
        L = self.lib
        n = L.ref2_generate_surface(int(rotated), int(N), _D(h), 0, None, None, None, None)
        x, nr, a, pl = np.zeros(3 * n), np.zeros(3 * n), np.zeros(n), np.zeros(n, np.int32)
        L.ref2_generate_surface(int(rotated), int(N), _D(h), n, _p(x), _p(nr), _p(a), _p(pl))
        return dict(n=n, x=x, nrml=nr, area=a, plane=pl)

    def reg_coeff(self, lam):
        lam = _f64(np.atleast_1d(lam))
        out = np.zeros(3 * len(lam))
        self.lib.ref2_reg_coeff(len(lam), _p(lam), _p(out))
        return out.reshape(-1, 3)

    def kernel(self, kind, r, d, c=(0.0, 0.0, 0.0)):
        k = {"sl": 0, "sl_on": 1, "dl": 2, "dl_on": 3}[kind]
        r = _f64(np.atleast_1d(r))
        cc = _f64(np.asarray(c))
        out = np.zeros(2 * len(r))
        self.lib.ref2_kernel(k, len(r), _p(r), _D(d), _p(cc), _p(out))
        return out.reshape(-1, 2)

    def nearest(self, surf_num, rotated, h, DEL, surf, tx):
        tx = _f64(tx)
        nt = len(tx) // 3
        nrst = np.zeros(nt, np.int32)
        x0, n0, f0 = np.zeros(3 * nt), np.zeros(3 * nt), np.zeros(3 * nt)
        b = np.zeros(nt)
        i123 = np.zeros(3 * nt, np.int32)
        self.lib.ref2_nearest(int(surf_num), int(rotated), _D(h), _D(DEL), surf["n"],
                              _p(surf["x"]), _p(surf["nrml"]), _p(surf["area"]),
                              _p(_i32(surf["plane"])), nt, _p(tx), _p(nrst), _p(x0), _p(n0),
                              _p(f0), _p(b), _p(i123))
        return dict(n=nt, x=tx, nrst=nrst, x0=x0, n0=n0, f0=f0, b=b, i123=i123)

    def surface_force(self, surf_num, x, nrml):
        x, nrml = _f64(x), _f64(nrml)
        f = np.zeros_like(x)
        self.lib.ref2_surface_force(int(surf_num), len(x) // 3, _p(x), _p(nrml), _p(f))
        return f

    def sl_off(self, DEL, src, tgt):
        out = np.zeros(3 * tgt["n"])
        self.lib.ref2_sl_off(_D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]), _p(src["area"]),
                             _p(src["f"]), tgt["n"], _p(tgt["x"]), _p(tgt["b"]), _p(tgt["n0"]),
                             _p(tgt["f0"]), _p(out))
        return out

    def sl_on(self, DEL, src):
        out = np.zeros(3 * src["n"])
        self.lib.ref2_sl_on(_D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]), _p(src["area"]),
                            _p(src["f"]), src["n"], _p(src["x"]), _p(out))
        return out

    def dl_off(self, DEL, src, tgt):
        out = np.zeros(3 * tgt["n"])
        self.lib.ref2_dl_off(_D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]), _p(src["area"]),
                             _p(src["g"]), tgt["n"], _p(tgt["x"]), _p(tgt["b"]), _p(tgt["x0"]),
                             _p(tgt["n0"]), _p(tgt["g0"]), _p(out))
        return out

    def dl_on(self, DEL, src):
        out = np.zeros(3 * src["n"])
        self.lib.ref2_dl_on(_D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]), _p(src["area"]),
                            _p(src["g"]), src["n"], _p(src["x"]), _p(out))
        return out

    def sl_on_tree(self, surf_num, DEL, src):
        out = np.zeros(3 * src["n"])
        self.lib.ref2_sl_on_tree(int(surf_num), _D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]),
                                 _p(src["area"]), _p(src["f"]), src["n"], _p(src["x"]), _p(out))
        return out

    def sl_off_tree(self, surf_num, DEL, src, tgt):
        out = np.zeros(3 * tgt["n"])
        self.lib.ref2_sl_off_tree(int(surf_num), _D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]),
                                  _p(src["area"]), _p(src["f"]), tgt["n"], _p(tgt["x"]),
                                  _p(tgt["b"]), _p(tgt["n0"]), _p(tgt["f0"]), _p(out))
        return out

    def dl_on_tree(self, surf_num, DEL, src):
        out = np.zeros(3 * src["n"])
        self.lib.ref2_dl_on_tree(int(surf_num), _D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]),
                                 _p(src["area"]), _p(src["g"]), src["n"], _p(src["x"]), _p(out))
        return out

    def dl_off_tree(self, surf_num, DEL, src, tgt):
        out = np.zeros(3 * tgt["n"])
        self.lib.ref2_dl_off_tree(int(surf_num), _D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]),
                                  _p(src["area"]), _p(src["g"]), tgt["n"], _p(tgt["x"]),
                                  _p(tgt["b"]), _p(tgt["x0"]), _p(tgt["n0"]), _p(tgt["g0"]),
                                  _p(out))
        return out

    def interp_g0(self, surf_num, h, src_x, g, tgt):
        nt = tgt["n"]
        out = np.zeros(3 * nt)
        self.lib.ref2_interp_g0(int(surf_num), _D(h), len(src_x) // 3, _p(src_x), _p(_f64(g)), nt,
                                _p(tgt["x"]), _p(tgt["nrst"]), _p(tgt["x0"]), _p(tgt["n0"]),
                                _p(tgt["i123"]), _p(out))
        return out

    # two-surface problem (state lives in the harness)
    def problem_setup(self, h, DEL, s1, s2):
        self._keep = (s1, s2)
        self.lib.ref2_problem_setup(_D(h), _D(DEL), s1["n"], _p(s1["x"]), _p(s1["nrml"]),
                                    _p(s1["area"]), _p(_i32(s1["plane"])), _p(s1.get("f")),
                                    s2["n"], _p(s2["x"]), _p(s2["nrml"]), _p(s2["area"]),
                                    _p(_i32(s2["plane"])), _p(s2.get("f")))

    def problem_nearest(self, k, n):
        nrst = np.zeros(n, np.int32)
        x0, n0, f0 = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        b = np.zeros(n)
        i123 = np.zeros(3 * n, np.int32)
        self.lib.ref2_problem_nearest(int(k), _p(nrst), _p(x0), _p(n0), _p(f0), _p(b), _p(i123))
        return dict(nrst=nrst, x0=x0, n0=n0, f0=f0, b=b, i123=i123)

    def problem_matvec(self, u):
        u = _f64(u)
        out = np.zeros_like(u)
        self.lib.ref2_problem_matvec(_p(u), _p(out))
        return out

    def problem_eval_sl(self, n1, n2):
        a, b = np.zeros(3 * n1), np.zeros(3 * n2)
        self.lib.ref2_problem_eval_sl(_p(a), _p(b))
        return np.concatenate([a, b])

    def full_solve(self, h, n1, n2):
        a, b = np.zeros(3 * n1), np.zeros(3 * n2)
        self.lib.ref2_full_solve(_D(h), _p(a), _p(b))
        return np.concatenate([a, b])


class RefGrid:
    """The compiled reference (highreg_stokes_extend_to_grid), via ref_shim_grid.cpp."""

    def __init__(self, path=None):
        self.lib = _load(path or os.path.join(HERE, "_ref", "librefgrid.so"))
        self.lib.refg_generate_surface.restype = _I
        self.lib.refg_band_targets.restype = _I

    def constants(self):
        out = np.zeros(4)
        self.lib.refg_constants(_p(out))
        return dict(abc=out[:3].copy(), ctr=out[3])

    def generate_surface(self, N, h):
        L = self.lib
        n = L.refg_generate_surface(int(N), _D(h), 0, None, None, None)
        x, nr, a = np.zeros(3 * n), np.zeros(3 * n), np.zeros(n)
        L.refg_generate_surface(int(N), _D(h), n, _p(x), _p(nr), _p(a))
        return dict(n=n, x=x, nrml=nr, area=a)

    def band_targets(self, N, h, band1, band2, surf):
        L = self.lib
        args = (int(N), _D(h), _D(band1), _D(band2), surf["n"], _p(surf["x"]), _p(surf["nrml"]),
                _p(surf["area"]))
        n = L.refg_band_targets(*args, 0, None, None, None, None, None)
        ijk = np.zeros(3 * n, np.int32)
        b, x0, n0 = np.zeros(n), np.zeros(3 * n), np.zeros(3 * n)
        band2a = np.zeros(n, np.int32)
        L.refg_band_targets(*args, n, _p(ijk), _p(b), _p(x0), _p(n0), _p(band2a))
        return dict(n=n, ijk=ijk.reshape(n, 3), b=b, x0=x0, n0=n0, band2=band2a)

    def sl_7ord(self, DEL, src, tx, tb, f0n0):
        tx, tb, f0n0 = _f64(tx), _f64(tb), _f64(f0n0)
        nt = len(tb)
        out = np.zeros(3 * nt)
        self.lib.refg_sl_7ord(_D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]), _p(src["area"]),
                              _p(src["f"]), nt, _p(tx), _p(tb), _p(f0n0), _p(out))
        return out

    def pressure_7ord(self, DEL, src, tgt):
        out = np.zeros(tgt["n"])
        self.lib.refg_pressure_7ord(_D(DEL), src["n"], _p(src["x"]), _p(src["nrml"]), _p(src["area"]),
                                    _p(src["f"]), tgt["n"], _p(_f64(tgt["x"])), _p(_f64(tgt["b"])),
                                    _p(_f64(tgt["n0"])), _p(_f64(tgt["f0"])), _p(out))
        return out

    def pressure_away(self, src, tx):
        tx = _f64(tx)
        out = np.zeros(len(tx) // 3)
        self.lib.refg_pressure_away(src["n"], _p(src["x"]), _p(src["area"]), _p(src["f"]), len(out), _p(tx),
                                    _p(out))
        return out

    def p_exact(self, x):
        x = _f64(x)
        p = np.zeros(len(x) // 3)
        self.lib.refg_p_exact(len(p), _p(x), _p(p))
        return p

    def density(self, x):
        x = _f64(x)
        f = np.zeros_like(x)
        self.lib.refg_density(len(x) // 3, _p(x), _p(f))
        return f

    def sl_exact(self, x):
        x = _f64(x)
        u = np.zeros_like(x)
        self.lib.refg_sl_exact(len(x) // 3, _p(x), _p(u))
        return u
