"""CPU pins of the learned preconditioner's pieces.

* the oracle's hand-derived backward pass equals central finite differences
  of its FP64 forward pass (every parameter block);
* the BF16 rounding helper rounds to nearest even;
* the library's k-nearest-neighbour graph equals a brute-force search.
"""
import numpy as np
import pytest

import mpnn_oracle as MO
from hor import _lib


def _knn_brute(x, k):
    x = x.reshape(-1, 3)
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    n = len(x)
    out = np.zeros((n, k), np.int64)
    for i in range(n):
        order = np.lexsort((np.arange(n), d2[i]))
        out[i] = order[:k]
    return out


def _two_clouds(n1=40, n2=30, seed=3):
    rng = np.random.default_rng(seed)

    def sphere(n, c, r):
        u = rng.standard_normal((n, 3))
        u /= np.linalg.norm(u, axis=1)[:, None]
        return c + r * u, u

    x1, n_1 = sphere(n1, np.array([0.0, 0.0, -0.55]), 0.5)
    x2, n_2 = sphere(n2, np.array([0.0, 0.0, 0.55]), 0.5)
    x = np.concatenate([x1, x2]).ravel()
    nr = np.concatenate([n_1, n_2]).ravel()
    return n1, n2, x, nr, np.array([0, 0, -0.55, 0, 0, 0.55]), np.array([0.5, 0.5])


def test_knn_graph_matches_brute_force():
    n1, n2, x, nr, ctr, size = _two_clouds(200, 150)
    k = 16
    nbr = np.zeros((n1 + n2) * k, np.int32)
    _lib.check(_lib.lib().hor_knn_graph(n1 + n2, _lib.ptr(x), k, _lib.ptr(nbr)), "hor_knn_graph")
    assert np.array_equal(nbr.reshape(-1, k), _knn_brute(x, k))


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e-5, 1.0 + 2.0 ** -9])
    got = MO.bf16(vals)
    # 1 + 2^-9 lies halfway between 1 and 1 + 2^-7 rounded to 8 bits: ties go to even (1.0)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, -2.5, np.float32(3.0e-5), 1.0])
    assert abs(got[4] - 3.0e-5) <= 3.0e-5 * 2 ** -8
    assert np.array_equal(got[[0, 1, 2, 3, 5]], want[[0, 1, 2, 3, 5]])


@pytest.mark.parametrize("batch", [1, 2])
def test_oracle_backward_matches_finite_differences(batch):
    n1, n2, x, nr, ctr, size = _two_clouds()
    k = 6
    nbr = _knn_brute(x, k)
    Gr = MO.Graph(n1, n2, x, nr, ctr, size, nbr)
    p = MO.init_params(5, scale_out=0.5).astype(np.float64)
    rng = np.random.default_rng(11)
    p += 0.05 * rng.standard_normal(p.size) * (p != 0)      # nonzero biases, keep padding at zero
    for off, n in ((MO.OFF_BIN, MO.W), (MO.OFF_BOUT, 3)):
        p[off:off + n] = 0.1 * rng.standard_normal(n)
    for l in range(MO.LAYERS):
        b = MO.OFF_L0 + l * MO.L_SIZE
        p[b + MO.L_CC:b + MO.L_CC + MO.W] = 0.1 * rng.standard_normal(MO.W)
        p[b + MO.L_D:b + MO.L_D + MO.W] = 0.1 * rng.standard_normal(MO.W)
    npts = n1 + n2
    r = rng.standard_normal((batch, 3 * npts))
    v = rng.standard_normal((batch, 3 * npts))
    net = MO.Net(p, bf=False)
    net.p = p.copy()
    loss, g, _ = net.loss_grad(Gr, r, v)

    def f(q):
        m = MO.Net(q, bf=False)
        m.p = q
        return m.loss_grad(Gr, r, v)[0]

    # a few entries of every parameter block
    blocks = [(MO.OFF_WIN, MO.W * MO.INPAD), (MO.OFF_BIN, MO.W), (MO.OFF_WOUT, 3 * MO.W), (MO.OFF_BOUT, 3)]
    for l in range(MO.LAYERS):
        b = MO.OFF_L0 + l * MO.L_SIZE
        blocks += [(b + MO.L_WA, MO.W * MO.W), (b + MO.L_BD, MO.W * 4), (b + MO.L_CC, MO.W),
                   (b + MO.L_WU, MO.W * MO.UW), (b + MO.L_D, MO.W)]
    idx = []
    for off, n in blocks:
        cand = [off + j for j in rng.choice(n, min(n, 4), replace=False)]
        if off != MO.OFF_WIN:
            idx += cand
        else:
            idx += [off + (j // MO.INPAD) * MO.INPAD + (j % 9) for j in rng.choice(MO.W * 9, 4, replace=False)]
    for (off, n) in blocks:
        if n == MO.W * 4:
            idx = [i for i in idx if not (off <= i < off + n and (i - off) % 4 == 3)]
    eps = 1e-6
    for i in idx:
        q1, q2 = p.copy(), p.copy()
        q1[i] += eps
        q2[i] -= eps
        fd = (f(q1) - f(q2)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    assert loss > 0
