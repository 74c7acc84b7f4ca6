"""The sampled-row oracle of the two-surface operators (tests/twosurf.py)
equals the rows of the oracle's full matvec and right-hand side (CPU only).

The GPU tests at benchmark sizes compare sampled rows, so the row helper is
itself checked here against the full operator on a small geometry: treecode
and direct sums, matvec (with the stencil-interpolated g0) and eval_SL."""
import numpy as np
import pytest

from conftest import rel_l2
from hor.solver import two_sphere_problem
from twosurf import operator_rows, sample_rows


@pytest.mark.parametrize("use_tree", [False, True])
def test_rows_equal_full_operator(oracle, use_tree):
    h = 1.0 / 16
    DEL, bodies, pd, b = two_sphere_problem(h, 0.1)
    tree = (4, 200, 0.6)
    pb = oracle.problem(h, DEL, pd, use_tree=use_tree, tree=(tree, tree))
    rows = sample_rows(pd, 40, seed=3)
    assert all(np.any(np.asarray(pd[k]["nrst"])[rows[k]] != 0) for k in range(2))
    u = np.random.default_rng(1).uniform(-1, 1, b.size)
    n1 = pd[0]["n"]
    for full, part in ((oracle.matvec(pb, u), operator_rows(oracle, h, DEL, pd, rows, u=u, tree=tree if use_tree else None)),
                       (oracle.eval_sl(pb), operator_rows(oracle, h, DEL, pd, rows, tree=tree if use_tree else None,
                                                          rhs=True))):
        for k, base in ((0, 0), (1, 3 * n1)):
            want = full[base:base + 3 * pd[k]["n"]].reshape(-1, 3)[rows[k]].ravel()
            err = rel_l2(part[k], want)
            print(f"tree={use_tree} surface {k + 1}: {rows[k].size} rows, rel diff {err:.2e}")
            assert err <= 1e-14
