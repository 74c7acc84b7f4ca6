"""Grid extension (highreg_stokes_extend_to_grid/main.cpp:969-1830) on the host:
the oracle's restatement pinned by the mathematics it must satisfy, and the
library's FFT-based version against it.  (The reference's main.cpp includes
fftw3.h, which this image lacks, so it is not built; these checks pin the
restatement instead.)

Oracle pins:
* DST-I Poisson: the returned solution satisfies the 4th-order stencil
  equations with zero Dirichlet data (residual ~1e-13), for random data;
* transfinite extension: reproduces the face values, and is exact for
  trilinear data;
* the 4th-order Laplacian stencil times h^2 is exact on cubic polynomials,
  and the 4th-order first differences (central and one-sided) on quartics;
* the extension of a harmonic function from the faces converges as h^4.
Product vs oracle: 1e-13 (DST) and 1e-12 (whole solve) relative.
"""
import numpy as np
import pytest

from hor import grid as HG


def _lap4_apply(v, n):
    """(2/3)(lap7 + (corners - 8 v)/8) of an n^3 array with zero padding."""
    u = np.zeros((n + 2,) * 3)
    u[1:-1, 1:-1, 1:-1] = v.reshape(n, n, n)
    c = u[1:-1, 1:-1, 1:-1]
    lap7 = (u[2:, 1:-1, 1:-1] + u[:-2, 1:-1, 1:-1] + u[1:-1, 2:, 1:-1] + u[1:-1, :-2, 1:-1] +
            u[1:-1, 1:-1, 2:] + u[1:-1, 1:-1, :-2] - 6 * c)
    corners = sum(u[1 + a:n + 1 + a, 1 + b:n + 1 + b, 1 + d:n + 1 + d]
                  for a in (-1, 1) for b in (-1, 1) for d in (-1, 1))
    return (2.0 / 3.0 * (lap7 + (corners - 8 * c) / 8.0)).ravel()


@pytest.mark.parametrize("n", [7, 12, 13])
def test_oracle_dst_solves_the_stencil_equations(oracle, n):
    f = np.random.default_rng(n).standard_normal(n ** 3)
    v = oracle.dst_poisson3d(f)
    r = _lap4_apply(v, n) - f
    assert np.abs(r).max() <= 1e-12 * np.abs(f).max()


@pytest.mark.parametrize("n", [11, 13, 23, 47])
def test_product_dst_matches_oracle(oracle, n):
    """n = 11, 23, 47: FFT lengths 24, 48, 96; n = 13 (length 28, factor 7):
    the dense-matrix path."""
    f = np.random.default_rng(n + 1).standard_normal(n ** 3)
    a = HG.dst_poisson3d(f.reshape(n, n, n)).ravel()
    b = oracle.dst_poisson3d(f)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def _grid(N, L):
    g = np.arange(N + 1) * (L / N)
    return np.meshgrid(g, g, g, indexing="ij")


def test_oracle_transfinite_extension(oracle):
    N, L = 10, 3.0
    rng = np.random.default_rng(2)
    g = rng.standard_normal((N + 1,) * 3)
    G = oracle.extend_into_cube(N, L, g.ravel()).reshape((N + 1,) * 3)
    for sl in (np.s_[0], np.s_[N], np.s_[:, 0], np.s_[:, N], np.s_[:, :, 0], np.s_[:, :, N]):
        assert np.abs(G[sl] - g[sl]).max() <= 1e-14
    X, Y, Z = _grid(N, L)
    t = 1.0 + 2 * X - Y + 0.5 * Z + 0.3 * X * Y - 0.7 * Y * Z + 0.2 * X * Z + 0.1 * X * Y * Z
    T = oracle.extend_into_cube(N, L, t.ravel()).reshape(t.shape)
    assert np.abs(T - t).max() <= 1e-13


def test_oracle_stencils_exact_on_polynomials(oracle):
    N, L = 12, 3.0
    h = L / N
    X, Y, Z = _grid(N, L)
    u = X ** 3 - 2 * X * Y * Z + Y ** 2 * Z + 0.5 * Z ** 3 + X * Y
    lap = 6 * X + 2 * Z + 3 * Z            # Delta u
    rhs = oracle.lapl_in_cube(N, u.ravel())
    assert np.abs(rhs + h * h * lap[1:-1, 1:-1, 1:-1].ravel()).max() <= 1e-12
    # the near-surface Laplacian adds +h^2 Delta u where band == 2
    band = np.zeros((N + 1) ** 3, np.int8)
    band[::3] = 2
    r2 = oracle.lapl_near_surface(N, band, u.ravel(), np.zeros((N - 1) ** 3))
    want = np.where(band.reshape(u.shape)[1:-1, 1:-1, 1:-1] == 2, h * h * lap[1:-1, 1:-1, 1:-1], 0.0).ravel()
    assert np.abs(r2 - want).max() <= 1e-12
    # 4th-order first differences, central and one-sided, exact on quartics
    p = X ** 4 - 3 * X ** 2 * Y + Y ** 4 * 0.5 + Z ** 4 - X * Y * Z ** 2
    grads = (4 * X ** 3 - 6 * X * Y - Y * Z ** 2, -3 * X ** 2 + 2 * Y ** 3 - X * Z ** 2, 4 * Z ** 3 - 2 * X * Y * Z)
    band0 = np.zeros((N + 1) ** 3, np.int8)
    outside = np.ones((N + 1) ** 3, np.uint8)
    for ax in range(3):
        d = oracle.dp_4ord(N, L, ax, band0, outside, p.ravel(), np.zeros((N - 1) ** 3))
        assert np.abs(d - h * h * grads[ax][1:-1, 1:-1, 1:-1].ravel()).max() <= 1e-10


def test_oracle_harmonic_extension_converges_fourth_order(oracle):
    """u = 1/|x - c| with c outside the box, given on the faces only: the
    discrete harmonic extension converges to u as h^4."""
    L, c = 3.0, np.array([4.0, 1.3, -0.7])
    errs = []
    for N in (12, 24):
        X, Y, Z = _grid(N, L)
        u = 1.0 / np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
        g = u.copy()
        g[1:-1, 1:-1, 1:-1] = 0.0
        band = np.zeros((N + 1) ** 3, np.int8)
        outside = np.zeros((N + 1) ** 3, np.uint8)
        e = oracle.solve_poisson(N, L, band, outside, None, g.ravel(), 0).reshape(u.shape)
        errs.append(np.abs(e - u).max())
    order = np.log2(errs[0] / errs[1])
    print("harmonic extension errors", errs, "order", order)
    assert order > 3.7


def test_product_solve_poisson_matches_oracle(oracle):
    N, L = 24, 3.0
    rng = np.random.default_rng(9)
    n3 = (N + 1) ** 3
    band = rng.choice(np.array([0, 1, 2], np.int8), n3, p=[0.7, 0.15, 0.15])
    outside = rng.integers(0, 2, n3).astype(np.uint8)
    p = rng.standard_normal(n3)
    for variable in (0, 1, 2, 3):
        u = rng.standard_normal(n3)
        a = HG.solve_poisson(L, band.reshape((N + 1,) * 3), outside.reshape((N + 1,) * 3),
                             p.reshape((N + 1,) * 3), u.reshape((N + 1,) * 3), variable).ravel()
        b = oracle.solve_poisson(N, L, band, outside, p, u, variable)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), variable
