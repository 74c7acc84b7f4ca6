"""The grid program end to end (highreg_stokes_extend_to_grid/main.cpp:107-968):
the translating spheroid (1, .5, .5) in the box [0, 3]^3, boundary-integral
pressure and velocity on the box faces and in the band (GPU plans), extended to
the whole grid by the host Poisson solver.

Checks: the extended fields agree with the exact (Chwang & Wu) pressure and
velocity over the nodes outside or on the body, the errors fall with h (orders
measured at h = 1/16 -> 1/32: pressure l2 4.5, velocity l2 5.2, velocity max
3.3), and the reference's mid-plane slice files are written from the result.
"""
import numpy as np
import pytest

from hor import grid as HG
from hor import io as HIO

pytestmark = pytest.mark.gpu


def test_grid_program_converges(ctx, tmp_path):
    res = []
    for N in (48, 96):
        r = HG.grid_program(ctx, N=N, log=print)
        e = r["errors"]
        res.append(e)
        # the grid program's own sanity levels: errors well below the field sizes
        assert e["p_max_outside_on"] < 0.1 * e["p_exact_max"]
        assert e["u_max_outside_on"] < 0.05 * e["u_exact_max"]
    for key, lo in (("p_l2_outside_on", 3.5), ("u_l2_outside_on", 4.0), ("u_max_outside_on", 2.5),
                    ("gradu_l2_grid", 2.0)):
        order = np.log2(res[0][key] / res[1][key])
        print(key, res[0][key], res[1][key], "order", order)
        assert order > lo, key
    paths = HIO.write_grid_files(r, str(tmp_path), tag="h32")
    sl = HIO.read_mid_y_slice(paths["err_u_midY"])
    assert sl.shape == (97, 97)
    assert sl.max() <= r["errors"]["u_max_outside_on"] * (1 + 1e-15)
