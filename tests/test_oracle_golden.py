"""The oracle against the compiled reference's outputs (golden fixtures, CPU).

two_surface_h16.npz holds the reference's spheroid pair at h = 1/16 (its own
Generate_Surface, Compute_nearest_data, set_surface_force), seeded densities,
and the outputs of the eight sum functions, interpolate_DL_density,
solve_IE_1step, eval_SL and the full GMRES solve.
"""
import numpy as np
import pytest

from conftest import golden, rel_l2, rel_max


@pytest.fixture(scope="module")
def G():
    return golden("two_surface_h16.npz")


def _surf(G, k):
    return dict(n=len(G[f"s{k}_area"]), x=G[f"s{k}_x"], nrml=G[f"s{k}_nrml"], area=G[f"s{k}_area"],
                plane=G[f"s{k}_plane"], f=G[f"s{k}_f"])


def _tgt(G, k):
    s = _surf(G, k)
    return dict(n=s["n"], x=s["x"], nrst=G[f"t{k}_nrst"], x0=G[f"t{k}_x0"], n0=G[f"t{k}_n0"],
                f0=G[f"t{k}_f0"], b=G[f"t{k}_b"], i123=G[f"t{k}_i123"])


def test_direct_sums(oracle, G):
    DEL = float(G["DEL"])
    src = dict(_surf(G, 1), g=G["g1"])
    tgt = dict(_tgt(G, 2), g0=G["t2_g0_interp"])
    assert np.array_equal(oracle.sl_off(DEL, src, tgt), G["sl_off"])
    assert np.array_equal(oracle.dl_off(DEL, src, tgt), G["dl_off"])
    assert np.array_equal(oracle.sl_on(DEL, src), G["sl_on"])
    assert np.array_equal(oracle.dl_on(DEL, src), G["dl_on"])


def test_treecode(oracle, G):
    DEL = float(G["DEL"])
    src = dict(_surf(G, 1), g=G["g1"])
    tgt = dict(_tgt(G, 2), g0=G["t2_g0_interp"])
    tp = oracle.tree_params(int(G["P"]), int(G["N0"]), float(np.sqrt(G["sq_theta"])), G["ctr1"])
    for name, got in (("sl_off_tree", oracle.sl_off_tree(tp, DEL, src, tgt)),
                      ("dl_off_tree", oracle.dl_off_tree(tp, DEL, src, tgt)),
                      ("sl_on_tree", oracle.sl_on_tree(tp, DEL, src)),
                      ("dl_on_tree", oracle.dl_on_tree(tp, DEL, src))):
        # SL is bit-exact; DL differs by one reordered add per pair (KITC.cpp:1047)
        assert rel_l2(got, G[name]) <= 1e-14 and rel_max(got, G[name]) <= 1e-14, name


def test_interpolation(oracle, G):
    t = _tgt(G, 2)
    got = oracle.interp_g0(G["ctr1"], float(G["h"]), G["s1_x"], G["g1"], t)
    assert np.array_equal(got, G["t2_g0_interp"])


def _problem(oracle, G, use_tree=True):
    pd = []
    for k in (1, 2):
        s, t = _surf(G, k), _tgt(G, k)
        pd.append(dict(n=s["n"], x=s["x"], nrml=s["nrml"], area=s["area"], ctr=G[f"ctr{k}"],
                       nrst=t["nrst"], x0=t["x0"], n0=t["n0"], b=t["b"], i123=t["i123"], f=s["f"], f0=t["f0"]))
    P, N0, th = int(G["P"]), int(G["N0"]), float(np.sqrt(G["sq_theta"]))
    return oracle.problem(float(G["h"]), float(G["DEL"]), pd, use_tree=use_tree, tree=((P, N0, th), (P, N0, th)))


def test_matvec_and_rhs(oracle, G):
    pb = _problem(oracle, G)
    got = oracle.matvec(pb, G["u"])
    assert rel_l2(got, G["matvec"]) <= 1e-14 and rel_max(got, G["matvec"]) <= 1e-14
    assert rel_l2(oracle.eval_sl(pb), G["eval_sl"]) <= 1e-15


def test_gmres_reproduces_reference_solve(oracle, G):
    pb = _problem(oracle, G)
    b = -(2.0 / 3.0) * oracle.eval_sl(pb)
    x, it, hist, res = oracle.gmres(pb, b, m=15, tol=1e-10, iter_max=100, fix_restart=False)
    assert it == 14                       # the reference prints 14 iterations at h = 1/16
    assert rel_l2(x, G["gmres_solution"]) <= 1e-13
    x2, it2, _, _ = oracle.gmres(pb, b, m=15, tol=1e-10, iter_max=100, fix_restart=True)
    assert it2 == it and rel_l2(x2, x) == 0.0   # no restart happens within 15 iterations
