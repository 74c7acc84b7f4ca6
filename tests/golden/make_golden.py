"""Writes the golden fixtures of tests/golden from the compiled reference.

Runs only where the reference sources were built into oracle/_ref (this
needs /root/reference at build time); the GPU tests only read the .npz files.
Inputs come from the reference's own geometry routines (Generate_Surface,
Compute_nearest_data, set_surface_force, find_nrst_surfc) and seeded
densities; outputs are the reference functions' results.

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py pressure   # grid_pressure_h16.npz only
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from pyoracle import Ref2, RefGrid  # noqa: E402


def delta(h, kappa0=4.0, q=5.0 / 7.0):
    return kappa0 * (1.0 / 64.0) ** (1.0 - q) * h ** q


def two_surface(R, N, seed):
    K = R.constants()
    h = 1.0 / N
    DEL = delta(h)
    s = []
    for rot, ctr in ((0, K["ctr1"]), (1, K["ctr2"])):
        d = R.generate_surface(rot, N, h)
        d["x"] = (d["x"].reshape(-1, 3) + ctr).ravel()
        s.append(d)
    s1, s2 = s
    s1["f"] = R.surface_force(1, s1["x"], s1["nrml"])
    s2["f"] = R.surface_force(2, s2["x"], s2["nrml"])
    R.problem_setup(h, DEL, s1, s2)
    t1 = R.problem_nearest(1, s1["n"])
    t2 = R.problem_nearest(2, s2["n"])
    rng = np.random.default_rng(seed)
    out = dict(h=h, DEL=DEL, ctr1=K["ctr1"], ctr2=K["ctr2"])
    for k, (sk, tk) in enumerate(((s1, t1), (s2, t2)), start=1):
        for nm in ("x", "nrml", "area", "plane", "f"):
            out[f"s{k}_{nm}"] = sk[nm]
        for nm in ("nrst", "x0", "n0", "f0", "b", "i123"):
            out[f"t{k}_{nm}"] = tk[nm]
    # operators on surface 1 -> points of surface 2 (nearest data t2) and on-surface
    g1 = rng.uniform(-1, 1, 3 * s1["n"])
    src = dict(s1, g=g1)
    tgt = dict(n=s2["n"], x=s2["x"], b=t2["b"], x0=t2["x0"], n0=t2["n0"], f0=t2["f0"],
               nrst=t2["nrst"], i123=t2["i123"])
    tgt["g0"] = R.interp_g0(1, h, s1["x"], g1, tgt)
    out["g1"] = g1
    out["t2_g0_interp"] = tgt["g0"]
    out["sl_off"] = R.sl_off(DEL, src, tgt)
    out["dl_off"] = R.dl_off(DEL, src, tgt)
    out["sl_on"] = R.sl_on(DEL, src)
    out["dl_on"] = R.dl_on(DEL, src)
    out["sl_off_tree"] = R.sl_off_tree(1, DEL, src, tgt)
    out["dl_off_tree"] = R.dl_off_tree(1, DEL, src, tgt)
    out["sl_on_tree"] = R.sl_on_tree(1, DEL, src)
    out["dl_on_tree"] = R.dl_on_tree(1, DEL, src)
    u = rng.uniform(-1, 1, 3 * (s1["n"] + s2["n"]))
    out["u"] = u
    out["matvec"] = R.problem_matvec(u)
    out["eval_sl"] = R.problem_eval_sl(s1["n"], s2["n"])
    return out


def grid_pressure(G):
    """stokes_pressure_7ord at the band nodes of the grid program (h = 1/16,
    every band node, inside ones too), with f the translating spheroid's
    traction and f0 = f(x0); stokes_pressure_away at the nodes of two box
    faces; p_exact at both."""
    N = 48
    h = 3.0 / N
    DEL = delta(h)
    s = G.generate_surface(N, h)
    s["f"] = G.density(s["x"])
    bt = G.band_targets(N, h, 4.0, 2.0, s)
    tx = (bt["ijk"] * h).astype(np.float64).ravel()
    tgt = dict(n=bt["n"], x=tx, b=bt["b"], n0=bt["n0"], f0=G.density(bt["x0"]))
    p_band = G.pressure_7ord(DEL, s, tgt)
    g = np.arange(N + 1) * h
    face = np.array([(x, y, z) for x in (0.0, N * h) for y in g for z in g])
    p_face = G.pressure_away(s, face.ravel())
    np.savez_compressed(os.path.join(HERE, "grid_pressure_h16.npz"), N=N, h=h, DEL=DEL, ijk=bt["ijk"],
                        b=bt["b"], n0=bt["n0"], f0=tgt["f0"], p_band=p_band, p_band_exact=G.p_exact(tx),
                        face=face.ravel(), p_face=p_face, p_face_exact=G.p_exact(face.ravel()))


def main():
    if sys.argv[1:] == ["pressure"]:
        grid_pressure(RefGrid())
        print("grid_pressure_h16.npz written to", HERE)
        return
    R = Ref2()
    K = R.constants()
    # scalar kernels and coefficients
    lam = np.linspace(-8.0, 8.0, 161)
    r = np.concatenate([[0.0, 5e-15], np.linspace(1e-3, 1.0, 400)])
    d = 0.1
    c = R.reg_coeff(np.array([0.37]))[0]
    np.savez_compressed(os.path.join(HERE, "kernels.npz"), lam=lam, coef=R.reg_coeff(lam), r=r, d=d,
                        c=c, sl=R.kernel("sl", r, d, c), sl_on=R.kernel("sl_on", r, d),
                        dl=R.kernel("dl", r, d, c), dl_on=R.kernel("dl_on", r, d))
    # two-surface problem at h = 1/16 (P, N0, theta are the reference's 6, 1000, 0.6)
    out = two_surface(R, 16, seed=20261203)
    out["gmres_solution"] = R.full_solve(out["h"], len(out["s1_area"]), len(out["s2_area"]))
    out["P"], out["N0"], out["sq_theta"] = K["P"], K["N0"], K["sq_theta"]
    np.savez_compressed(os.path.join(HERE, "two_surface_h16.npz"), **out)
    # extend-to-grid geometry: surface, band targets, exact solution (N = 48, h = 1/16)
    G = RefGrid()
    N = 48
    h = 3.0 / N
    s = G.generate_surface(N, h)
    bt = G.band_targets(N, h, 4.0, 2.0, s)
    tx = (bt["ijk"] * h).astype(np.float64).ravel()
    outside = bt["b"] > 0
    np.savez_compressed(os.path.join(HERE, "grid_h16.npz"), N=N, h=h, ctr=G.constants()["ctr"],
                        s_x=s["x"], s_nrml=s["nrml"], s_area=s["area"], ijk=bt["ijk"], b=bt["b"],
                        x0=bt["x0"], n0=bt["n0"], band2=bt["band2"], f=G.density(s["x"]),
                        exact=G.sl_exact(tx), outside=outside)
    grid_pressure(G)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
