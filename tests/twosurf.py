"""Rows of the two-surface operators evaluated by the oracle at sampled points.

At benchmark sizes the oracle cannot afford every row of a matvec, so the
tests compare sampled rows: for the points ``rows[k]`` of surface k the row of

  * the matvec of solve_IE_1step (highreg_stokes_2surf/main.cpp:541-677):
    u_k - 2/3 [DL_kk(u_k) + DL_ok(u_o)], with g0 of the off-surface term
    interpolated from u_o on the 6x6 Monge stencil (main.cpp:916-1046);
  * the right-hand side of eval_SL (main.cpp:421-537): SL_kk(f_k) + SL_ok(f_o).

The on-surface sums evaluate their first nt source points, so the sampled
points are moved to the front of the source list.  The tree's node
membership depends on the point coordinates only (midpoint splits), so the
reordering changes only the order of the sums: the rows equal those of the
full operator to rounding.
"""
import numpy as np


def _sel3(a, idx):
    return np.ascontiguousarray(np.asarray(a).reshape(-1, 3)[idx].ravel())


def _front(s, idx, key, dens):
    n = s["n"]
    order = np.concatenate([idx, np.setdiff1d(np.arange(n), idx)])
    return {"n": n, "x": _sel3(s["x"], order), "nrml": _sel3(s["nrml"], order),
            "area": np.ascontiguousarray(np.asarray(s["area"])[order]), key: _sel3(dens, order)}


def sample_rows(pd, count, seed):
    """Per surface: count//2 points near the other surface (nrst != 0) and the
    rest drawn from all points, sorted, without repeats."""
    rng = np.random.default_rng(seed)
    rows = []
    for k in range(2):
        near = np.flatnonzero(np.asarray(pd[k]["nrst"]) != 0)
        a = rng.choice(near, min(count // 2, near.size), replace=False) if near.size else np.zeros(0, int)
        b = rng.choice(pd[k]["n"], count - a.size, replace=False)
        rows.append(np.unique(np.concatenate([a, b])).astype(np.int64))
    return rows


def operator_rows(O, h, DEL, pd, rows, u=None, tree=None, rhs=False):
    """Rows rows[k] of A u (rhs=False) or of the SL of the surface forces
    pd[k]["f"] (rhs=True) by the oracle's treecode (tree = (P, N0, theta)) or
    its direct sums (tree = None).  Returns [surface-1 rows, surface-2 rows],
    3 values per row, xyz interleaved."""
    n1 = pd[0]["n"]
    key = "f" if rhs else "g"
    dens = [pd[0]["f"], pd[1]["f"]] if rhs else [u[:3 * n1], u[3 * n1:]]
    out = []
    for k in range(2):
        o = 1 - k
        idx = np.asarray(rows[k])
        nt = idx.size
        sk = _front(pd[k], idx, key, dens[k])
        so = {"n": pd[o]["n"], "x": pd[o]["x"], "nrml": pd[o]["nrml"], "area": pd[o]["area"], key: dens[o]}
        tgt = {"n": nt, "x": _sel3(pd[k]["x"], idx), "b": np.ascontiguousarray(np.asarray(pd[k]["b"])[idx]),
               "x0": _sel3(pd[k]["x0"], idx), "n0": _sel3(pd[k]["n0"], idx)}
        if rhs:
            tgt["f0"] = _sel3(pd[k]["f0"], idx)
        else:
            st = {"n": nt, "nrst": np.ascontiguousarray(np.asarray(pd[k]["nrst"])[idx]), "x0": tgt["x0"],
                  "i123": np.ascontiguousarray(np.asarray(pd[k]["i123"]).reshape(-1, 3)[idx].ravel())}
            tgt["g0"] = O.interp_g0(pd[o]["ctr"], h, pd[o]["x"], dens[o], st)
        if tree is not None:
            tpk = O.tree_params(*tree, pd[k]["ctr"])
            tpo = O.tree_params(*tree, pd[o]["ctr"])
            if rhs:
                v = O.sl_on_tree(tpk, DEL, sk, nt=nt) + O.sl_off_tree(tpo, DEL, so, tgt)
            else:
                v = O.dl_on_tree(tpk, DEL, sk, nt=nt) + O.dl_off_tree(tpo, DEL, so, tgt)
        else:
            if rhs:
                v = O.sl_on(DEL, sk, nt=nt) + O.sl_off(DEL, so, tgt)
            else:
                v = O.dl_on(DEL, sk, nt=nt) + O.dl_off(DEL, so, tgt)
        out.append(v if rhs else _sel3(dens[k], idx) - (2.0 / 3.0) * v)
    return out


def gpu_rows(full, pd, rows):
    """The same rows of a full GPU result vector [surface 1 | surface 2]."""
    n1 = pd[0]["n"]
    return [_sel3(full[:3 * n1], rows[0]), _sel3(full[3 * n1:], rows[1])]
