"""GPU two-surface solve against the CPU oracle: stencil interpolation, matvec,
RHS (eval_SL) and GMRES, direct and treecode.

Matvec / RHS / interpolation: 1e-12 relative (||.||_2 and max norm).
GMRES: the oracle and the GPU run the same restarted GMRES; both stop at
relative residual 1e-10, and the solutions agree to 1e-9.
"""
import numpy as np
import pytest

import hor
from cases import problem_dicts, spheroid_pair, sphere_pair
from conftest import rel_l2, rel_max
from hor import solver as S

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _check(a, b, tol=TOL):
    assert np.all(np.isfinite(a))
    assert rel_l2(a, b) <= tol, rel_l2(a, b)
    assert rel_max(a, b) <= tol, rel_max(a, b)


@pytest.fixture(scope="module")
def pair16():
    return spheroid_pair(N=16)


@pytest.fixture(scope="module")
def pair32():
    return spheroid_pair(N=32)


def _make(ctx, oracle, case, use_tree, tree=(6, 1000, 0.6)):
    h, DEL, bodies, surfs, tgts = case
    pd = problem_dicts(bodies, surfs, tgts)
    tp = hor.TreeParams(*tree)
    sol = S.Solver(ctx, h, DEL, pd[0], pd[1], method=hor.TREE if use_tree else hor.DIRECT, tree1=tp, tree2=tp)
    pb = oracle.problem(h, DEL, pd, use_tree=use_tree, tree=(tree, tree))
    return sol, pb, pd


@pytest.mark.parametrize("use_tree", [False, True])
def test_interp_and_matvec(ctx, oracle, pair32, use_tree):
    sol, pb, pd = _make(ctx, oracle, pair32, use_tree)
    rng = np.random.default_rng(3)
    u = rng.uniform(-1, 1, sol.n)
    n1 = pd[0]["n"]
    # g0 at surface-2 points from the surface-1 density
    g_gpu = sol.interp_g0(2, u[: 3 * n1])
    g_orc = oracle.interp_g0(pd[0]["ctr"], pb.h, pd[0]["x"], u[: 3 * n1],
                             dict(n=pd[1]["n"], nrst=pd[1]["nrst"], x0=pd[1]["x0"], i123=pd[1]["i123"]))
    _check(g_gpu, g_orc)
    print("matvec", use_tree, sol.n, flush=True)
    _check(sol.matvec(u), oracle.matvec(pb, u))


@pytest.mark.parametrize("use_tree", [False, True])
def test_rhs(ctx, oracle, pair32, use_tree):
    sol, pb, pd = _make(ctx, oracle, pair32, use_tree)
    sl = sol.eval_sl(pd[0]["f"], pd[1]["f"], pd[0]["f0"], pd[1]["f0"])
    _check(sl, oracle.eval_sl(pb))


def test_gmres_matches_oracle(ctx, oracle, pair16):
    sol, pb, pd = _make(ctx, oracle, pair16, use_tree=True)
    b = -(2.0 / 3.0) * oracle.eval_sl(pb)                  # main.cpp:199-207
    x_o, it_o, hist_o, res_o = oracle.gmres(pb, b, m=15, tol=1e-10, iter_max=100, fix_restart=True)
    x_g, it_g, hist_g, res_g = sol.gmres(b, m=15, tol=1e-10, maxit=100)
    print("gmres iters", it_o, it_g, res_o, res_g, flush=True)
    assert it_g == it_o
    assert res_g <= 1e-10
    assert rel_l2(x_g, x_o) <= 1e-9
    # the final residual checked with an independent FP64 matvec
    r = b - oracle.matvec(pb, x_g)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 2e-10


def test_gmres_restart_path(ctx, oracle):
    """m = 3 forces restarts; the GPU restarts from the true residual."""
    case = sphere_pair(N=8, gap=0.1)
    sol, pb, pd = _make(ctx, oracle, case, use_tree=False)
    rng = np.random.default_rng(5)
    b = rng.uniform(-1, 1, sol.n)
    x_o, it_o, _, _ = oracle.gmres(pb, b, m=3, tol=1e-10, iter_max=200, fix_restart=True)
    x_g, it_g, _, res = sol.gmres(b, m=3, tol=1e-10, maxit=200)
    print("restart gmres", it_o, it_g, res, flush=True)
    assert abs(it_g - it_o) <= 1
    assert rel_l2(x_g, x_o) <= 1e-8


def test_mixed_gmres_ends_on_fp64_residual(ctx):
    """Mixed-precision treecode GMRES: Arnoldi with the FP32-far-field
    operator, restarts and the final check with the FP64 operator, so the
    reported residual is the FP64 one and meets the tolerance."""
    h, DEL, bodies, surfs, tgts = sphere_pair(32, gap=0.1)
    pd = problem_dicts(bodies, surfs, tgts)
    tp = hor.TreeParams(6, 500, 0.6)
    s64 = S.Solver(ctx, h, DEL, pd[0], pd[1], method=hor.TREE, precision=hor.FP64, tree1=tp, tree2=tp)
    smx = S.Solver(ctx, h, DEL, pd[0], pd[1], method=hor.TREE, precision=hor.MIXED, tree1=tp, tree2=tp)
    b = np.random.default_rng(4).uniform(-1, 1, s64.n)
    for tol in (1e-8, 1e-11):
        x, it, hist, res = smx.gmres(b, m=15, tol=tol, maxit=200)
        true = np.linalg.norm(b - s64.matvec(x)) / np.linalg.norm(b)
        x64, it64, _, res64 = s64.gmres(b, m=15, tol=tol, maxit=200)
        print(f"tol {tol:g}: mixed {it} iterations, FP64 residual {res:.3e} (recomputed {true:.3e}); "
              f"FP64 solve {it64} iterations")
        assert res <= tol and abs(true - res) <= 1e-3 * res + 1e-16
        assert np.linalg.norm(x - x64) <= 10 * tol * np.linalg.norm(x64) / 0.3
    s64.close()
    smx.close()


def test_gmres_nonfinite_rhs_is_an_error(ctx):
    """A NaN or Inf in b ends the solve with HOR_E_NAN (no iteration on
    non-finite data)."""
    h, DEL, bodies, surfs, tgts = sphere_pair(N=8, gap=0.1)
    pd = problem_dicts(bodies, surfs, tgts)
    sol = S.Solver(ctx, h, DEL, pd[0], pd[1])
    for bad in (np.nan, np.inf):
        b = np.random.default_rng(6).uniform(-1, 1, sol.n)
        b[17] = bad
        for call in (lambda: sol.gmres(b, m=15, tol=1e-8, maxit=50),
                     lambda: sol.fgmres(b, m=15, tol=1e-8, maxit=50)):
            with pytest.raises(hor.HorError) as e:
                call()
            assert e.value.code == -4  # HOR_E_NAN
    # the solver stays usable
    b = np.random.default_rng(6).uniform(-1, 1, sol.n)
    _, it, _, res = sol.gmres(b, m=15, tol=1e-8, maxit=50)
    assert res <= 1e-8
    sol.close()


def test_maxit_returns_the_true_residual(ctx):
    """A mixed-precision GMRES or a flexible GMRES stopped by maxit reports
    the FP64 true residual |b - A x|/|b|, not the Arnoldi estimate."""
    h, DEL, bodies, surfs, tgts = sphere_pair(32, gap=0.1)
    pd = problem_dicts(bodies, surfs, tgts)
    tp = hor.TreeParams(6, 500, 0.6)
    s64 = S.Solver(ctx, h, DEL, pd[0], pd[1], method=hor.TREE, precision=hor.FP64, tree1=tp, tree2=tp)
    smx = S.Solver(ctx, h, DEL, pd[0], pd[1], method=hor.TREE, precision=hor.MIXED, tree1=tp, tree2=tp)
    b = np.random.default_rng(8).uniform(-1, 1, s64.n)
    for name, call in (("mixed gmres", lambda mi: smx.gmres(b, m=15, tol=1e-14, maxit=mi)[::3]),
                       ("fgmres", lambda mi: s64.fgmres(b, m=15, tol=1e-14, maxit=mi)[::3]),
                       ("mixed fgmres m=2", lambda mi: smx.fgmres(b, m=2, tol=1e-14, maxit=mi)[::3])):
        for maxit in (3, 4):
            x, res = call(maxit)
            true = np.linalg.norm(b - s64.matvec(x)) / np.linalg.norm(b)
            print(f"{name}, maxit {maxit}: reported {res:.6e}, true {true:.6e}")
            assert abs(res - true) <= 1e-9 * true
    s64.close()
    smx.close()
