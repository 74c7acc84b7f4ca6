"""Two-sphere second-kind solve at the benchmark sizes, GPU against the oracle.

Workloads (spheres of radius 0.5, centres (0, 0, -+(1 + gap)/2), GMRES(15)):
  * h = 1/128, treecode P = 8, N0 = 4000, theta = 0.6, gaps 0.1, 0.05, 0.025,
    FP64: the full matvec and right-hand side against the oracle's, and the
    GMRES solve to 1e-8 against the oracle's GMRES (same iteration count);
  * h = 1/256, treecode P = 10, N0 = 8000, theta = 0.6: sampled rows of the
    matvec and the right-hand side against the oracle's treecode (FP64) and
    its direct sums, and the mixed-precision operator and solve with their
    error against both stated;
  * h = 1/256 with the learned preconditioner shipped with the package, in
    gated flexible GMRES: reaches 1e-8 on the FP64 residual in no more
    iterations than plain GMRES.

Parity bar: FP64 results within 1e-12 of the oracle (2-norm and max-norm,
relative); GMRES iteration counts equal and solutions within 1e-9.  The mixed
operator (FP32 far field and FP32 pairs beyond 4 delta, FP64 accumulation) is
held to 1e-6 of the FP64 operator.
"""
import numpy as np
import pytest

import hor
from conftest import rel_l2, rel_max
from hor import mpnn as MP
from hor.solver import Solver, two_sphere_problem
from twosurf import gpu_rows, operator_rows, sample_rows

pytestmark = pytest.mark.gpu
TOL = 1e-12
MIXED_TOL = 1e-6


def _check(a, b, tol=TOL, what=""):
    assert np.all(np.isfinite(a))
    e2, em = rel_l2(a, b), rel_max(a, b)
    print(f"{what}: rel l2 {e2:.2e}, rel max {em:.2e}", flush=True)
    assert e2 <= tol and em <= tol, (what, e2, em)
    return e2


def _rows_err(a, b):
    return rel_l2(np.concatenate(a), np.concatenate(b)), rel_max(np.concatenate(a), np.concatenate(b))


@pytest.mark.parametrize("gap", [0.1, 0.05, 0.025])
def test_h128_treecode_matvec_rhs_gmres(ctx, oracle, gap):
    h, tree = 1.0 / 128, (8, 4000, 0.6)
    print(f"two spheres h=1/128 gap {gap}, P=8 N0=4000", flush=True)
    DEL, bodies, pd, b_rigid = two_sphere_problem(h, gap, ctx=ctx)
    tp = hor.TreeParams(*tree)
    S = Solver(ctx, h, DEL, pd[0], pd[1], method=hor.TREE, precision=hor.FP64, tree1=tp, tree2=tp)
    pb = oracle.problem(h, DEL, pd, use_tree=True, tree=(tree, tree))
    u = np.random.default_rng(11).uniform(-1, 1, S.n)
    _check(S.matvec(u), oracle.matvec(pb, u), what="matvec")
    sl_o = oracle.eval_sl(pb)
    _check(S.eval_sl(pd[0]["f"], pd[1]["f"], pd[0]["f0"], pd[1]["f0"]), sl_o, what="rhs")
    b = -(2.0 / 3.0) * sl_o                                 # main.cpp:199-207
    x_o, it_o, _, res_o = oracle.gmres(pb, b, m=15, tol=1e-8, iter_max=100, fix_restart=True)
    x_g, it_g, _, res_g = S.gmres(b, m=15, tol=1e-8, maxit=100)
    print(f"gmres: oracle {it_o} iterations ({res_o:.2e}), GPU {it_g} iterations ({res_g:.2e}), "
          f"solution rel diff {rel_l2(x_g, x_o):.2e}", flush=True)
    assert it_g == it_o
    assert res_g <= 1e-8
    assert rel_l2(x_g, x_o) <= 1e-9
    S.close()


@pytest.fixture(scope="module")
def h256(ctx):
    h, tree = 1.0 / 256, (10, 8000, 0.6)
    print("two spheres h=1/256 gap 0.1, P=10 N0=8000: geometry and solvers", flush=True)
    DEL, bodies, pd, b_rigid = two_sphere_problem(h, 0.1, ctx=ctx)
    tp = hor.TreeParams(*tree)
    s64 = Solver(ctx, h, DEL, pd[0], pd[1], method=hor.TREE, precision=hor.FP64, tree1=tp, tree2=tp)
    smx = Solver(ctx, h, DEL, pd[0], pd[1], method=hor.TREE, precision=hor.MIXED, tree1=tp, tree2=tp)
    b_force = -(2.0 / 3.0) * s64.eval_sl(pd[0]["f"], pd[1]["f"], pd[0]["f0"], pd[1]["f0"])
    rows = sample_rows(pd, 400, seed=7)
    d = dict(h=h, DEL=DEL, tree=tree, bodies=bodies, pd=pd, s64=s64, smx=smx, b_rigid=b_rigid,
             b_force=b_force, rows=rows)
    yield d
    s64.close()
    smx.close()


def test_h256_matvec_rows_fp64_and_mixed(oracle, h256):
    d = h256
    h, DEL, pd, rows, tree = d["h"], d["DEL"], d["pd"], d["rows"], d["tree"]
    u = np.random.default_rng(12).uniform(-1, 1, d["s64"].n)
    y64, ymx = d["s64"].matvec(u), d["smx"].matvec(u)
    print(f"h=1/256: {d['s64'].n} unknowns, {rows[0].size}+{rows[1].size} sampled rows", flush=True)
    o_tree = operator_rows(oracle, h, DEL, pd, rows, u=u, tree=tree)
    o_dir = operator_rows(oracle, h, DEL, pd, rows, u=u, tree=None)
    _check(np.concatenate(gpu_rows(y64, pd, rows)), np.concatenate(o_tree), what="FP64 treecode matvec rows")
    mt = _rows_err(gpu_rows(ymx, pd, rows), o_tree)
    md = _rows_err(gpu_rows(ymx, pd, rows), o_dir)
    td = _rows_err(o_tree, o_dir)
    print(f"mixed matvec rows vs oracle treecode: rel l2 {mt[0]:.2e}, rel max {mt[1]:.2e}", flush=True)
    print(f"mixed matvec rows vs oracle direct sums: rel l2 {md[0]:.2e}, rel max {md[1]:.2e}", flush=True)
    print(f"(FP64 treecode vs direct sums: rel l2 {td[0]:.2e}, rel max {td[1]:.2e})", flush=True)
    full = rel_l2(ymx, y64)
    print(f"mixed vs FP64 GPU matvec, all rows: rel l2 {full:.2e}", flush=True)
    assert mt[0] <= MIXED_TOL and mt[1] <= MIXED_TOL and full <= MIXED_TOL
    # the mixed-precision error stays below the treecode's own truncation error
    assert md[0] <= 2.0 * td[0] + MIXED_TOL


def test_h256_rhs_rows(oracle, h256):
    d = h256
    pd = d["pd"]
    sl = -1.5 * d["b_force"]
    o = operator_rows(oracle, d["h"], d["DEL"], pd, d["rows"], tree=d["tree"], rhs=True)
    _check(np.concatenate(gpu_rows(sl, pd, d["rows"])), np.concatenate(o), what="FP64 treecode SL rows")


def test_h256_gmres_fp64_and_mixed(oracle, h256):
    d = h256
    h, DEL, pd, rows, tree = d["h"], d["DEL"], d["pd"], d["rows"], d["tree"]
    b = d["b_force"]
    x64, it64, _, est64 = d["s64"].gmres(b, m=15, tol=1e-8, maxit=100)
    xmx, itmx, _, resmx = d["smx"].gmres(b, m=15, tol=1e-8, maxit=100)
    r64 = np.linalg.norm(b - d["s64"].matvec(x64)) / np.linalg.norm(b)
    rmx = np.linalg.norm(b - d["s64"].matvec(xmx)) / np.linalg.norm(b)
    print(f"h=1/256 forcing RHS: FP64 {it64} iterations (estimate {est64:.2e}, FP64 residual {r64:.2e}); "
          f"mixed {itmx} iterations (FP64 residual {resmx:.2e}, recomputed {rmx:.2e}); "
          f"solutions rel diff {rel_l2(xmx, x64):.2e}", flush=True)
    assert est64 <= 1e-8 and r64 <= 1.01e-8
    assert resmx <= 1e-8 and abs(rmx - resmx) <= 1e-3 * resmx + 1e-16
    assert rel_l2(xmx, x64) <= 1e-6
    # residuals of both solutions under the oracle's treecode and direct operators, sampled rows
    bs = np.concatenate(gpu_rows(b, pd, rows))
    for name, x in (("FP64", x64), ("mixed", xmx)):
        rt = np.linalg.norm(bs - np.concatenate(operator_rows(oracle, h, DEL, pd, rows, u=x, tree=tree)))
        rd = np.linalg.norm(bs - np.concatenate(operator_rows(oracle, h, DEL, pd, rows, u=x, tree=None)))
        print(f"{name} solution, sampled rows: residual under the oracle treecode {rt / np.linalg.norm(bs):.2e}, "
              f"under the oracle direct sums {rd / np.linalg.norm(bs):.2e}", flush=True)
        assert rt <= 1e-7 * np.linalg.norm(bs)
        assert rd <= 1e-6 * np.linalg.norm(bs)


def test_h256_learned_preconditioner_gated(ctx, h256):
    """The shipped network in gated flexible GMRES on the mixed operator: the
    FP64 residual meets 1e-8 and the solve takes no more iterations than
    plain GMRES on the same operator (the gain or loss is printed)."""
    params = MP.load_weights()
    assert params is not None, "precond_weights.npz missing"
    d = h256
    S = d["smx"]
    net = MP.Network(ctx, params=params)
    net.set_geometry(d["pd"], d["bodies"])
    for name, b in (("forcing", d["b_force"]), ("rigid", d["b_rigid"])):
        _, st0 = MP.solve(S, b, net=None, tol=1e-8, maxit=100)
        x, st1 = MP.solve(S, b, net=net, tol=1e-8, maxit=100)
        r = np.linalg.norm(b - d["s64"].matvec(x)) / np.linalg.norm(b)
        print(f"{name}: plain {st0['iterations']} iterations, learned+gate {st1['iterations']} "
              f"({st1['learned_steps']} learned steps, gate on/off {st1['cycles_gated_on']}/"
              f"{st1['cycles_gated_off']}, monitor stops {st1['monitor_stops']}, misfit I {st1['gate_misfit_identity']:.3f} "
              f"M {st1['gate_misfit_network']:.3f}); gain {st0['iterations'] - st1['iterations']} iterations; "
              f"FP64 residual {st1['relres']:.2e} (recomputed {r:.2e})", flush=True)
        assert st1["relres"] <= 1e-8 and r <= 1.01e-8
        assert st1["iterations"] <= st0["iterations"]
        assert st1["nonfinite_outputs"] == 0
    net.close()
