"""Seeded test geometries and densities (built with the package's host geometry)."""
import numpy as np

from hor import geometry as G

SEED = 20261203


def ellipsoid_case(h=1 / 16, seed=SEED, band_factor=3.0, max_band=None, abc=(1.0, 0.8, 0.6)):
    """Config-1 style: ellipsoid centred at 0, SH densities f, g; targets are
    the quadrature points (on-surface) and grid nodes with |b| <= band_factor*delta."""
    body = G.Body(abc)
    s = body.surface(h)
    DEL = G.delta(h)
    cf, cg = G.sh_coefficients(seed), G.sh_coefficients(seed + 1)
    s["f"] = G.sh_density(s["x"], body.ctr, cf)
    s["g"] = G.sh_density(s["x"], body.ctr, cg)
    ngrid = int(round(2 * 1.5 / h))
    t = body.band_targets(s, (-1.5, -1.5, -1.5), h, ngrid, band_factor * DEL)
    if max_band is not None and t["n"] > max_band:
        idx = np.sort(np.random.default_rng(seed).choice(t["n"], max_band, replace=False))
        t = dict(n=len(idx), x=t["x"].reshape(-1, 3)[idx].ravel(), b=t["b"][idx],
                 x0=t["x0"].reshape(-1, 3)[idx].ravel(), n0=t["n0"].reshape(-1, 3)[idx].ravel(),
                 ijk=t["ijk"][idx])
    t["f0"] = G.sh_density(t["x0"], body.ctr, cf)
    t["g0"] = G.sh_density(t["x0"], body.ctr, cg)
    return body, s, t, DEL


def spheroid_pair(N=16, gap=1.0 / 4096.0):
    """The reference's two-surface geometry (highreg_stokes_2surf/utilities.h:15-26,
    main.cpp:67-119): spheroid (1, .5, .5) at (0,0,-0.5-gap) and the same body
    rotated 30 degrees about z at (0,0,0.5)."""
    h = 1.0 / N
    DEL = G.delta(h)
    b1 = G.Body((1.0, 0.5, 0.5), (0.0, 0.0, -0.5 - gap))
    b2 = G.Body((1.0, 0.5, 0.5), (0.0, 0.0, 0.5), G.rot_z(30.0))
    s1, s2 = b1.surface(h), b2.surface(h)
    t2 = b1.nearest(h, DEL, s1, s2["x"])   # points of surface 2 w.r.t. surface 1
    t1 = b2.nearest(h, DEL, s2, s1["x"])
    s1["f"] = b1.force(s1["x"], s1["nrml"])
    s2["f"] = b2.force(s2["x"], s2["nrml"])
    return h, DEL, (b1, b2), (s1, s2), (t1, t2)


def sphere_pair(N=16, gap=0.1, radius=0.5):
    """Two spheres of radius 0.5 with centres (0,0,+-(2r+gap)/2)."""
    h = 1.0 / N
    DEL = G.delta(h)
    zc = radius + gap / 2
    b1 = G.Body((radius,) * 3, (0.0, 0.0, -zc))
    b2 = G.Body((radius,) * 3, (0.0, 0.0, zc))
    s1, s2 = b1.surface(h), b2.surface(h)
    t2 = b1.nearest(h, DEL, s1, s2["x"])
    t1 = b2.nearest(h, DEL, s2, s1["x"])
    return h, DEL, (b1, b2), (s1, s2), (t1, t2)


def problem_dicts(bodies, surfs, tgts):
    """Surface dicts for the oracle / solver: nearest data of each surface's
    points relative to the other surface."""
    out = []
    for k in range(2):
        s, t = surfs[k], tgts[k]
        out.append(dict(n=s["n"], x=s["x"], nrml=s["nrml"], area=s["area"], plane=s["plane"],
                        ctr=bodies[k].ctr, nrst=t["nrst"], x0=t["x0"], n0=t["n0"], b=t["b"],
                        i123=t["i123"], f=s.get("f"), f0=t.get("f0")))
    return out
