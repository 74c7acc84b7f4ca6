"""Benchmark: regularized Stokeslet + stresslet velocity on the band of grid
nodes around a translating ellipsoid (the extend-to-grid band evaluation).

Workload: ellipsoid semi-axes (1, 0.8, 0.6) at the origin, grid h = 1/64 in
[-1.5, 1.5]^3, quadrature = grid-line/surface intersections (70 degree
partition of unity), delta = 4 (1/64)^(2/7) h^(5/7); targets = every grid node
with |b| <= 0.3.  Densities: f = exact traction of the body translating with
U = (1,0,0) (SL), g = seeded degree-3 spherical harmonics (DL).  One step =
the FP64 direct SL_OFF and DL_OFF sums over all band nodes with inputs
resident in HBM, as one fused pass (HOR_SLDL_OFF: both kernels for every pair,
the two velocity fields of the separate operators); the separate SL_OFF +
DL_OFF evaluations are timed beside it.

Prints one JSON line (metric = pair interactions per second).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "high-order-regularization_amd"), os.path.join(ROOT, "oracle")]

# OpenMP placement of the CPU baseline legs (read by libgomp when the first
# OpenMP library loads, so before any import of the package): threads packed
# onto cores.  OMP_NUM_THREADS is left to the environment (the process's CPU
# share); host_cpu() records what the threads ran on.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
try:   # the process's CPUs, read before OpenMP binds this thread to one of them
    CPUS_AVAILABLE = len(os.sched_getaffinity(0))
except AttributeError:
    CPUS_AVAILABLE = os.cpu_count() or 1

from hor import solver as TS  # noqa: E402

H = 1.0 / 64.0
BAND = 0.3
ABC = (1.0, 0.8, 0.6)
SEED = 20261203
PEAK_FP64 = 78.6e12          # MI355X FP64 (vector and matrix) FLOP/s
VALU_ISSUE = 256 * 4 * 16 * 2.4e9   # VALU lane-instructions/s: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
# Algorithmic operation counts per pair (reference expressions, per-target
# invariants hoisted, FMA = 2): plain SL 30, regularized off-surface SL 61,
# plain DL 30, regularized off-surface DL 127.  The pressure pair is counted in
# its reduced form dx.w (w = -a f + (f0.n0) a n + (n0 x f0) x a n), fewer
# operations than the reference's normal + tangential expression: plain 36,
# regularized 63.
# The treecode's double-layer plans apply the target-side subtraction g - g0
# through per-target coefficients computed once per plan, so their matvec
# kernels evaluate the pair and proxy expressions on g alone: the reference's
# DL proxy interaction (KITC.cpp:748-771) without its 9 products and 9
# differences g0_l W_n,m is 33 operations, and its on-surface and plain DL
# pairs without g - g0 are 49 and 27.
OPS = {"sl_plain": 30, "sl_reg": 61, "dl_plain": 30, "dl_reg": 127, "p_plain": 36, "p_reg": 63,
       "dl_on_reg": 52, "dl_proxy": 51, "dl_moment": 37,
       "dl_proxy_nosub": 33, "dl_on_reg_nosub": 49, "dl_plain_nosub": 27}
PEAK_FP32 = 157.3e12         # MI355X FP32 (vector and matrix) FLOP/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """The host CPUs of the CPU baseline legs: lscpu's model, sockets, cores
    per socket and threads per core, the CPUs this process may run on, and the
    OpenMP settings (threads = OMP_NUM_THREADS, else the CPUs available)."""
    info = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k] = v
    except (OSError, ValueError):
        pass
    avail = CPUS_AVAILABLE
    try:
        physical = int(info["Socket(s)"]) * int(info["Core(s) per socket"])
    except (KeyError, ValueError):
        physical = None
    threads = int(os.environ.get("OMP_NUM_THREADS", avail))
    return {"threads": threads, "cpus_available": avail, "physical_cores_machine": physical,
            "lscpu": info, "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"), "OMP_PLACES": os.environ.get("OMP_PLACES")}


def build_workload(h=H, band=BAND):
    from hor import geometry as G
    body = G.Body(ABC)
    s = body.surface(h)
    DEL = G.delta(h)
    ngrid = int(round(3.0 / h))
    t0 = time.time()
    t = body.band_targets(s, (-1.5, -1.5, -1.5), h, ngrid, band)
    log(f"band targets: {t['n']} nodes in {time.time() - t0:.1f}s; sources {s['n']}; delta {DEL:.5f}")
    s["f"] = body.traction(s["x"])
    cg = G.sh_coefficients(SEED + 1)
    s["g"] = G.sh_density(s["x"], body.ctr, cg)
    t["f0"] = body.traction(t["x0"])
    t["g0"] = G.sh_density(t["x0"], body.ctr, cg)
    return body, s, t, DEL


def near_pair_fraction(s, t, DEL, sample=2000, seed=1):
    """Fraction of (target, source) pairs with r/delta <= 8 (regularized
    pairs), estimated on a random target sample."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(t["n"], min(sample, t["n"]), replace=False)
    tx = t["x"].reshape(-1, 3)[idx]
    sx = s["x"].reshape(-1, 3)
    near = 0
    for a in range(0, len(tx), 200):
        d = np.linalg.norm(tx[a:a + 200, None, :] - sx[None, :, :], axis=2)
        near += int((d / DEL <= 8.0).sum())
    return near / (len(idx) * s["n"])


def cpu_baseline(s, t, DEL, target_seconds=15.0):
    """The reference's own eval_Stokes_SL_offSurf / eval_Stokes_DL_offSurf
    (compiled from its sources into oracle/_ref) on a bounded target sample,
    all host cores; falls back to the oracle restatement when the reference
    build is absent."""
    hc = host_cpu()
    cores = hc["threads"]
    kind = "reference"
    try:
        from pyoracle import Ref2
        impl = Ref2()
    except (FileNotFoundError, OSError):
        from pyoracle import Oracle
        impl, kind = Oracle(), "port"
    rng = np.random.default_rng(2)

    def sub(n):
        idx = np.sort(rng.choice(t["n"], n, replace=False))
        return dict(n=n, x=t["x"].reshape(-1, 3)[idx].ravel(), b=t["b"][idx],
                    x0=t["x0"].reshape(-1, 3)[idx].ravel(), n0=t["n0"].reshape(-1, 3)[idx].ravel(),
                    f0=t["f0"].reshape(-1, 3)[idx].ravel(), g0=t["g0"].reshape(-1, 3)[idx].ravel())

    def run(tt):
        t0 = time.perf_counter()
        impl.sl_off(DEL, s, tt)
        impl.dl_off(DEL, s, tt)
        return time.perf_counter() - t0

    n0 = 1000
    dt0 = run(sub(n0))
    # three timed runs of a third of the time budget each; the median is reported
    n = int(min(t["n"], max(n0, n0 * target_seconds / 3.0 / max(dt0, 1e-3))))
    tt = sub(n)
    runs = [run(tt) for _ in range(3)]
    dt = float(np.median(runs))
    pairs = 2.0 * n * s["n"]
    return {"value": pairs / dt, "unit": "pair interactions/s", "cores": cores, "kind": kind,
            "sample": f"{n} of {t['n']} band targets x {s['n']} sources, SL_OFF + DL_OFF, median of 3 runs "
                      f"({', '.join(f'{x:.2f}' for x in runs)} s)",
            "runs_s": runs, "host": hc}


def pressure_cpu_baseline(s, t, DEL, target_seconds=6.0):
    """The reference's stokes_pressure_7ord (compiled into oracle/_ref) on a
    bounded target sample, all host cores; the oracle restatement when the
    reference build is absent."""
    hc = host_cpu()
    cores = hc["threads"]
    try:
        from pyoracle import RefGrid
        impl, kind = RefGrid().pressure_7ord, "reference"
    except (FileNotFoundError, OSError):
        from pyoracle import Oracle
        impl, kind = Oracle().pressure_off, "port"
    rng = np.random.default_rng(3)

    def sub(n):
        idx = np.sort(rng.choice(t["n"], n, replace=False))
        return dict(n=n, x=t["x"].reshape(-1, 3)[idx].ravel(), b=t["b"][idx],
                    n0=t["n0"].reshape(-1, 3)[idx].ravel(), f0=t["f0"].reshape(-1, 3)[idx].ravel())

    def run(tt):
        t0 = time.perf_counter()
        impl(DEL, s, tt)
        return time.perf_counter() - t0

    n0 = 500
    dt0 = run(sub(n0))
    n = int(min(t["n"], max(n0, n0 * target_seconds / 3.0 / max(dt0, 1e-3))))
    tt = sub(n)
    runs = [run(tt) for _ in range(3)]
    dt = float(np.median(runs))
    return {"value": n * s["n"] / dt, "unit": "pair interactions/s", "cores": cores, "kind": kind,
            "sample": f"{n} of {t['n']} band targets x {s['n']} sources, stokes_pressure_7ord, median of 3 runs "
                      f"({', '.join(f'{x:.2f}' for x in runs)} s)", "runs_s": runs}


def geometry_setup(ctx, h, gap, pd_gpu, gpu_s, sample=256):
    """Geometry set-up on the GPU against the host: the same quadrature and
    nearest data built by the host functions (time, and whether every
    integer and position is identical), the GPU stencil build, and the
    reference's stencil algorithm (the oracle's scan of every quadrature
    point per target) timed on a sample of near targets and scaled to all."""
    from hor import geometry as G
    t0 = time.perf_counter()
    _, _, pd_host, _ = TS.two_sphere_problem(h, gap)
    host_s = time.perf_counter() - t0
    same = all(np.array_equal(pd_gpu[k][key], pd_host[k][key]) for k in range(2)
               for key in ("x", "nrml", "plane", "nrst", "i123", "x0", "n0", "b", "f0"))
    area = max(float(np.abs(pd_gpu[k]["area"] - pd_host[k]["area"]).max()) / h ** 2 for k in range(2))
    srcs = [dict(pd_gpu[1], ctr=pd_gpu[1]["ctr"]), dict(pd_gpu[0], ctr=pd_gpu[0]["ctr"])]
    G.stencils(ctx, srcs[0], pd_gpu[0], h)  # warm-up
    t0 = time.perf_counter()
    for k in range(2):
        G.stencils(ctx, srcs[k], pd_gpu[k], h)
    st_s = time.perf_counter() - t0
    out = {"gpu_s": gpu_s, "host_s": host_s, "speedup": host_s / gpu_s, "identical_to_host": bool(same),
           "area_max_abs_diff_over_h2": area, "stencil_gpu_s": st_s,
           "note": "GPU: quadrature of both spheres + nearest data both ways (host arrays in and out); "
                   "host: the same functions on the host cores (k-d tree nearest search)"}
    from pyoracle import Oracle
    O = Oracle()
    near = [np.where(pd_gpu[k]["nrst"] != 0)[0] for k in range(2)]
    tot = 0.0
    for k in range(2):
        pick = near[k][:sample]
        sub = dict(n=len(pick), nrst=pd_gpu[k]["nrst"][pick], x0=pd_gpu[k]["x0"].reshape(-1, 3)[pick].ravel(),
                   i123=pd_gpu[k]["i123"].reshape(-1, 3)[pick].ravel())
        t0 = time.perf_counter()
        O.stencil(srcs[k]["ctr"], h, srcs[k]["x"], sub)
        tot += (time.perf_counter() - t0) * len(near[k]) / max(len(pick), 1)
    out["stencil_cpu_baseline"] = {"seconds": tot, "kind": "port", "cores": 1,
                                   "sample": f"{sample} near targets per surface, scaled to "
                                             f"{len(near[0]) + len(near[1])} near targets (per matvec in the "
                                             "reference, once per geometry here)"}
    return out


def cpu_matvec(h, DEL, pd, P, N0, theta, u, reps=1):
    """One treecode matvec of the two-surface operator by the oracle
    restatement on all host cores (kind "port": the reference's KITC fixes
    P = 6, N0 = 1000 at compile time, KITC.h:14-16, so it cannot run these
    parameters; at its own parameters the restatement and the compiled
    reference take the same time and give identical results)."""
    from pyoracle import Oracle
    O = Oracle()
    pb = O.problem(h, DEL, pd, use_tree=True, tree=((P, N0, theta), (P, N0, theta)))
    t0 = time.perf_counter()
    for k in range(reps):
        y = O.matvec(pb, u)
        log(f"  CPU matvec {k + 1}/{reps}: {time.perf_counter() - t0:.1f}s")
    t_mv = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    O.eval_sl(pb)
    return t_mv, time.perf_counter() - t0, y


def gmres_workload(h=1.0 / 128.0, gap=0.1, P=8, N0=4000, theta=0.6, tol=1e-8, seed=SEED,
                   direct_check=False, cpu=False, cpu_reps=1, geometry=False):
    """Second-kind double-layer equation A u = u - 2/3 DL u = b on two spheres
    (hor.solver.two_sphere_problem), treecode matvec, GMRES(15) to relative residual tol,
    FP64 and mixed-precision treecodes, both right-hand sides.  Every solution's
    residual is rechecked with the FP64 treecode operator (and, with
    direct_check, with the FP64 direct-sum operator)."""
    import hor
    from hor.solver import Solver
    ctx = hor.Context(0)
    TS.two_sphere_problem(1.0 / 16.0, gap, seed, ctx=ctx)  # first GPU geometry call (code object load)
    t0 = time.perf_counter()
    DEL, bodies, pd, b_rigid = TS.two_sphere_problem(h, gap, seed, ctx=ctx)
    geo_s = time.perf_counter() - t0
    tp = hor.TreeParams(P, N0, theta)
    n = b_rigid.size
    out = {"workload": f"two spheres r=0.5, gap {gap}, h=1/{round(1 / h)}, treecode P={P} N0={N0} "
                       f"theta={theta}, GMRES(15) to {tol:g}",
           "unknowns": int(n), "delta": DEL, "geometry_s": geo_s}
    if geometry:
        out["geometry"] = geometry_setup(ctx, h, gap, pd, geo_s)
    solvers = {}
    for name, prec in (("fp64", hor.FP64), ("mixed", hor.MIXED)):
        t0 = time.time()
        solvers[name] = Solver(ctx, h, DEL, pd[0], pd[1], method=hor.TREE, precision=prec, tree1=tp, tree2=tp)
        out.setdefault("setup_s", {})[name] = time.time() - t0
    S64 = solvers["fp64"]
    b_force = -(2.0 / 3.0) * S64.eval_sl(pd[0]["f"], pd[1]["f"], pd[0]["f0"], pd[1]["f0"])
    for name, S in solvers.items():
        S.matvec(b_rigid)  # warm-up
        t0 = time.perf_counter()
        for _ in range(3):
            S.matvec(b_rigid)
        out.setdefault("matvec_ms", {})[name] = 1e3 * (time.perf_counter() - t0) / 3
    if abs(h - 1.0 / 256.0) < 1e-12:
        ud, yd = ctx.alloc(b_rigid.nbytes), ctx.alloc(b_rigid.nbytes)
        ctx.upload(ud, b_rigid)
        freg = onsurf_reg_fraction(S64, pd, DEL)
        counted = profile_tree()
        out["roofline"] = {name: tree_roofline(S, phase_mean(S, ud, yd, 3), freg, name == "mixed", counted)
                           for name, S in solvers.items()}
        out["roofline"]["note"] = ("phase times from HIP events on the solver's stream around each plan's "
                                   "k_moments, k_leaf and k_far launches, summed over the four DL plans of a "
                                   "matvec and averaged over 3 matvecs; counted figures from "
                                   "profiles/tree_matvec_h256_pmc.json")
        ctx.free(ud)
        ctx.free(yd)
        out["sl_far_field"] = sl_far_kernels(ctx, pd, DEL, tp)
    st = S64.stats()
    mv_s = out["matvec_ms"]["fp64"] * 1e-3
    out["matvec_work"] = dict(st, direct_equivalent_pairs_per_s=st["direct_pairs"] / mv_s,
                              evaluated_interactions_per_s=(st["near_pairs"] + st["proxy_pairs"] +
                                                            st["moment_terms"]) / mv_s)
    if cpu:
        cores = host_cpu()["threads"]
        log(f"CPU treecode matvec (oracle restatement, {cores} threads)")
        t_cpu, t_rhs, y_cpu = cpu_matvec(h, DEL, pd, P, N0, theta, b_force, reps=cpu_reps)
        y_gpu = S64.matvec(b_force)
        out["cpu_baseline"] = {"matvec_s": t_cpu, "matvecs_timed": cpu_reps, "rhs_s": t_rhs, "cores": cores,
                               "kind": "port",
                               "gpu_fp64_matvec_rel_diff": float(np.linalg.norm(y_gpu - y_cpu) / np.linalg.norm(y_cpu)),
                               "note": f"mean of {cpu_reps} timed matvec(s); CPU GMRES time = iterations x matvec "
                                       "(extrapolated, the RHS excluded as on the GPU side)"}
    S_dir = Solver(ctx, h, DEL, pd[0], pd[1], method=hor.DIRECT, precision=hor.FP64) if direct_check else None
    # learned right preconditioner (trained weights shipped with the package)
    from hor import mpnn as MP
    params = MP.load_weights()
    net = None
    if params is not None:
        t0 = time.time()
        net = MP.Network(ctx, params=params)
        net.set_geometry(pd, bodies)
        out["learned_setup_s"] = time.time() - t0
    for rhs_name, b in (("rigid", b_rigid), ("forcing", b_force)):
        res = {}
        xs = {}
        runs = [(name, S, None) for name, S in solvers.items()]
        if net is not None:
            runs.append(("mixed_learned", solvers["mixed"], net))
        for name, S, M in runs:
            t0 = time.perf_counter()
            if M is None:
                x, it, hist, est = S.gmres(b, m=15, tol=tol, maxit=300)
                info = {"iterations": it, "relres_estimate": est}
                if S is S64:
                    # the FP64 solve stops on its estimate; its FP64 residual check is part of the
                    # timed solve, as the mixed and learned solves end on one inside theirs
                    S64.matvec(x)
                    info["final_fp64_check_in_time"] = True
            else:
                x, st = MP.solve(S, b, net=M, m=15, tol=tol, maxit=300)
                info = {"iterations": st["iterations"], "learned_steps": st["learned_steps"],
                        "plain_steps": st["plain_steps"], "gate_cycles_on": st["cycles_gated_on"],
                        "gate_cycles_off": st["cycles_gated_off"], "monitor_stops": st["monitor_stops"],
                        "gate_misfit_identity": st["gate_misfit_identity"],
                        "gate_misfit_network": st["gate_misfit_network"],
                        "nonfinite_outputs": st["nonfinite_outputs"], "relres_fp64": st["relres"]}
            dt = time.perf_counter() - t0
            xs[name] = x
            r = b - S64.matvec(x)
            res[name] = dict(info, seconds=dt, ms_per_iteration=1e3 * dt / max(info["iterations"], 1),
                             relres_fp64_recheck=float(np.linalg.norm(r) / np.linalg.norm(b)))
            if S_dir is not None:
                rd = b - S_dir.matvec(x)
                res[name]["relres_fp64_direct"] = float(np.linalg.norm(rd) / np.linalg.norm(b))
        for name in xs:
            if name != "fp64":
                res[name]["rel_diff_vs_fp64_solution"] = float(np.linalg.norm(xs[name] - xs["fp64"]) /
                                                               np.linalg.norm(xs["fp64"]))
                res[f"speedup_{name}_vs_fp64"] = res["fp64"]["seconds"] / res[name]["seconds"]
        out[rhs_name] = res
        if "cpu_baseline" in out:
            it = res["fp64"]["iterations"]
            out["cpu_baseline"][f"gmres_s_{rhs_name}"] = it * out["cpu_baseline"]["matvec_s"]
            res["speedup_fp64_gpu_vs_cpu"] = it * out["cpu_baseline"]["matvec_s"] / res["fp64"]["seconds"]
    if net is not None:
        net.close()
    for S in solvers.values():
        S.close()
    if S_dir is not None:
        S_dir.close()
    ctx.close()
    return out


def onsurf_reg_fraction(S, pd, DEL, sample=2000, seed=3):
    """Fraction of the on-surface near-leaf pairs that are regularized
    (r/delta <= 8), from the mean number of same-surface sources within
    8 delta of a seeded sample of targets (k-d tree)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    nreg = 0.0
    for k in range(2):
        x = pd[k]["x"].reshape(-1, 3)
        idx = rng.choice(x.shape[0], min(sample, x.shape[0]), replace=False)
        cnt = cKDTree(x).query_ball_point(x[idx], 8.0 * DEL, return_length=True)
        nreg += float(np.mean(cnt)) * x.shape[0]
    near_on = S.plan_stats(0)["near_pairs"] + S.plan_stats(1)["near_pairs"]
    return min(1.0, nreg / max(near_on, 1.0))


def tree_roofline(S, phase, freg, mixed, counted=None):
    """Roofline figures of one treecode matvec's kernels from its phase times
    (HIP events, ms) and the plans' work, counted as the kernels evaluate it
    (the subtraction g - g0 is applied through the plans' coefficients): the DL
    far field k_far (33 flops per proxy interaction), the near sums k_leaf (49
    flops per regularized on-surface pair, 27 per plain pair; the fraction
    regularized sampled) and the moments (37 flops per moment term, FP64
    matrix cores)."""
    st = [S.plan_stats(k) for k in range(4)]
    proxy = sum(x["proxy_pairs"] for x in st)
    near = sum(x["near_pairs"] for x in st)
    mom = sum(x["moment_terms"] for x in st)
    peak_far = PEAK_FP32 if mixed else PEAK_FP64
    far_flops = proxy * OPS["dl_proxy_nosub"]
    near_flops = near * (freg * OPS["dl_on_reg_nosub"] + (1 - freg) * OPS["dl_plain_nosub"])
    mom_flops = mom * OPS["dl_moment"]
    out = {
        "far": {"kernel": "k_far<true, true, 2, true>" if mixed else "k_far<true, false, 1, true>", "bound": "valu",
                "compute": "FP32 vector ALU, FP64 row sums" if mixed else "FP64 vector ALU",
                "ms": phase["far_ms"], "flop": far_flops, "achieved": far_flops / (phase["far_ms"] * 1e-3) / 1e12,
                "peak": peak_far / 1e12, "unit": "TFLOP/s",
                "frac": far_flops / (phase["far_ms"] * 1e-3) / peak_far},
        "near": {"kernel": "k_leaf (DL_ON and DL_OFF)", "bound": "valu", "ms": phase["near_ms"],
                 "flop": near_flops, "regularized_fraction": freg,
                 "achieved": near_flops / (phase["near_ms"] * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12,
                 "unit": "TFLOP/s", "frac": near_flops / (phase["near_ms"] * 1e-3) / PEAK_FP64},
        "moments": {"kernel": "k_moments (FP64 matrix cores) + reduction", "bound": "mfma", "ms": phase["moments_ms"],
                    "flop": mom_flops, "achieved": mom_flops / max(phase["moments_ms"] * 1e-3, 1e-12) / 1e12,
                    "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                    "frac": mom_flops / max(phase["moments_ms"] * 1e-3, 1e-12) / PEAK_FP64},
        "work": {"proxy_interactions": proxy, "near_pairs": near, "moment_terms": mom},
    }
    if counted:
        for key, name in (("far", out["far"]["kernel"]), ("near", "k_leaf<2, %s, true>" % ("true" if mixed else "false"))):
            c = counted.get(name)
            if c:
                out[key]["counted"] = {k: c[k] for k in ("fp64_tflops", "fp64_frac_of_peak", "fp32_tflops",
                                                         "valu_issue_frac", "hbm_bytes_per_launch") if k in c}
    return out


def profile_tree():
    """Per-kernel counters of the h=1/256 treecode matvec from the committed
    rocprofv3 summary (VALU FLOPs, issue rate, HBM bytes)."""
    path = os.path.join(ROOT, "profiles", "tree_matvec_h256_pmc.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("kernels")
    return None


def phase_mean(S, u_dev, y_dev, reps):
    """Mean phase times (ms) over reps device matvecs."""
    acc = {"moments_ms": 0.0, "near_ms": 0.0, "far_ms": 0.0}
    for _ in range(reps):
        S.matvec_dev(u_dev, y_dev)
        for k, v in S.phase_timing().items():
            acc[k] += v / reps
    return acc


def sl_far_kernels(ctx, pd, DEL, tp, reps=3):
    """Single-layer far field of the forcing right-hand side -2/3 SL[f]
    (main.cpp:199-207): sphere 1's traction at its own points (SL_ON) and at
    sphere 2's (SL_OFF), mixed precision, the proxy interactions on the FP32
    vector units (k_far) and as a GEMM on FP32 matrix cores (k_far_sl_mfma).
    Far-phase kernel time (HIP events on the plan's stream, mean of reps
    evaluations after one warm-up) and the relative l2 difference of each
    result from the FP64 plan's."""
    import hor
    s = pd[0]
    out = {}
    for name, kind, tgt, d0 in (("on_surface", hor.SL_ON, None, None),
                                ("other_sphere", hor.SL_OFF, pd[1], pd[1]["f0"])):
        ref_plan = hor.Plan(ctx, kind, s, tgt, DEL, hor.TREE, hor.FP64, s["ctr"], tp)
        ref = ref_plan(s["f"], d0)
        ref_plan.close()
        row = {}
        for kname, kern in (("valu", hor.FAR_VALU), ("mfma", hor.FAR_MFMA)):
            plan = hor.Plan(ctx, kind, s, tgt, DEL, hor.TREE, hor.MIXED, s["ctr"], tp).set_far_kernel(kern)
            y = plan(s["f"], d0)
            far = 0.0
            for _ in range(reps):
                y = plan(s["f"], d0)
                far += plan.phase_timing()["far_ms"]
            row[kname] = {"far_ms": far / reps,
                          "rel_l2_vs_fp64": float(np.linalg.norm(y - ref) / np.linalg.norm(ref))}
            row["proxy_interactions"] = plan.stats()["proxy_pairs"]
            plan.close()
        row["mfma_time_over_valu"] = row["mfma"]["far_ms"] / row["valu"]["far_ms"]
        out[name] = row
        log(f"SL far field {name}: FP32 VALU {row['valu']['far_ms']:.2f} ms, "
            f"FP32 MFMA {row['mfma']['far_ms']:.2f} ms")
    return out


def matvec_workload(h=1.0 / 256.0, gap=0.1, P=10, N0=8000, theta=0.6, reps=3, seed=SEED):
    """Only the two-sphere treecode matvecs (the GMRES operator): reps FP64
    and reps mixed-precision matvecs of the seeded rigid-body vector, after one
    warm-up each.  The short run that the profiler passes measure."""
    import hor
    from hor.solver import Solver
    ctx = hor.Context(0)
    DEL, bodies, pd, b = TS.two_sphere_problem(h, gap, seed, ctx=ctx)
    tp = hor.TreeParams(P, N0, theta)
    out = {"workload": f"two spheres r=0.5, gap {gap}, h=1/{round(1 / h)}, treecode P={P} N0={N0} theta={theta}: "
                       f"{reps} FP64 and {reps} mixed matvecs", "unknowns": int(b.size)}
    for name, prec in (("fp64", hor.FP64), ("mixed", hor.MIXED)):
        S = Solver(ctx, h, DEL, pd[0], pd[1], method=hor.TREE, precision=prec, tree1=tp, tree2=tp)
        ud, yd = ctx.alloc(b.nbytes), ctx.alloc(b.nbytes)
        ctx.upload(ud, b)
        S.matvec_dev(ud, yd)
        t0 = time.perf_counter()
        for _ in range(reps):
            S.matvec_dev(ud, yd)
        out[f"{name}_matvec_ms"] = 1e3 * (time.perf_counter() - t0) / reps
        out[f"{name}_work"] = S.stats()
        out[f"{name}_phase_ms"] = S.phase_timing()
        out[f"{name}_plan_stats"] = [S.plan_stats(k) for k in range(4)]
        log(f"{name}: {out[f'{name}_matvec_ms']:.1f} ms per matvec")
        ctx.free(ud)
        ctx.free(yd)
        S.close()
    out["sl_far_field"] = sl_far_kernels(ctx, pd, DEL, tp)
    ctx.close()
    return out


def profile_traffic():
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)."""
    path = os.path.join(ROOT, "profiles", "direct_band_h64_pmc.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("traffic_bytes_per_launch")
    return None


def profile_valu():
    """Per-launch FP64/FP32 FLOPs and VALU instructions counted by rocprofv3
    (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64/F32, calibrated on FMA chains of
    known length) from the committed summary."""
    path = os.path.join(ROOT, "profiles", "direct_band_h64_valu_pmc.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("kernels")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-gmres", action="store_true", help="skip the two-sphere GMRES workload")
    ap.add_argument("--no-grid", action="store_true", help="skip the grid-program leg")
    ap.add_argument("--no-h256", action="store_true", help="skip the h=1/256 two-sphere GMRES workload")
    ap.add_argument("--no-cpu-h256", action="store_true",
                    help="skip the CPU treecode matvecs and RHS at h=1/256 (about 80 s of CPU time)")
    ap.add_argument("--gmres-only", type=int, default=0, metavar="N",
                    help="run only the two-sphere GMRES workload at h=1/N (128 or 256) and print its JSON")
    ap.add_argument("--matvec-only", type=int, default=0, metavar="N",
                    help="run only FP64 and mixed two-sphere treecode matvecs at h=1/N (256: P=10, N0=8000; "
                         "128: P=8, N0=4000) and print their JSON (the profiler runs)")
    ap.add_argument("--h", type=float, default=H)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (the path shards by "
              "targets across GPUs by design, but this build measures one GPU)", file=sys.stderr)
        sys.exit(2)
    import hor

    if args.matvec_only:
        P, N0 = (10, 8000) if args.matvec_only >= 256 else (8, 4000)
        print(json.dumps(matvec_workload(h=1.0 / args.matvec_only, P=P, N0=N0)), flush=True)
        return
    if args.gmres_only:
        if args.gmres_only == 128:
            print(json.dumps(gmres_workload()), flush=True)
        else:
            print(json.dumps(gmres_workload(h=1.0 / args.gmres_only, P=10, N0=8000, cpu=not args.no_cpu_h256,
                                            cpu_reps=2, geometry=True)), flush=True)
        return
    body, s, t, DEL = build_workload(h=args.h)
    ctx = hor.Context(0)
    t0 = time.time()
    p_v = hor.Plan(ctx, hor.SLDL_OFF, s, t, DEL)
    p_sl = hor.Plan(ctx, hor.SL_OFF, s, t, DEL)
    p_dl = hor.Plan(ctx, hor.DL_OFF, s, t, DEL)
    log(f"plans built in {time.time() - t0:.1f}s")
    ns, nt = s["n"], t["n"]
    dev = {}
    for name, arr in (("f", s["f"]), ("g", s["g"]), ("f0", t["f0"]), ("g0", t["g0"]),
                      ("fg", np.concatenate([s["f"], s["g"]])), ("fg0", np.concatenate([t["f0"], t["g0"]]))):
        dev[name] = ctx.alloc(arr.nbytes)
        ctx.upload(dev[name], np.ascontiguousarray(arr))
    out_sl, out_dl, out_v = ctx.alloc(24 * nt), ctx.alloc(24 * nt), ctx.alloc(48 * nt)

    def step():
        p_v.eval_dev(dev["fg"], dev["fg0"], out_v)
        return ctx.last_timing()[1]

    for _ in range(args.warmup):
        step()
    kv = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kv.append(step())
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    pairs_per_step = 2.0 * ns * nt
    value = pairs_per_step / (wall / args.steps)
    ref_v = ctx.download(out_v, np.zeros(6 * nt))

    # the same sums as two separate operators (SL_OFF, then DL_OFF)
    ksl, kdl = [], []
    p_sl.eval_dev(dev["f"], dev["f0"], out_sl)
    t0 = time.perf_counter()
    for _ in range(max(2, args.steps // 2)):
        p_sl.eval_dev(dev["f"], dev["f0"], out_sl)
        ksl.append(ctx.last_timing()[1])
        p_dl.eval_dev(dev["g"], dev["g0"], out_dl)
        kdl.append(ctx.last_timing()[1])
    sep_ms = 1e3 * (time.perf_counter() - t0) / len(ksl)
    ref_sl = ctx.download(out_sl, np.zeros(3 * nt))
    ref_dl = ctx.download(out_dl, np.zeros(3 * nt))
    fused_vs_separate = max(float(np.abs(ref_v[:3 * nt] - ref_sl).max() / np.abs(ref_sl).max()),
                            float(np.abs(ref_v[3 * nt:] - ref_dl).max() / np.abs(ref_dl).max()))

    # the same step through the host-buffer entry point (densities in, velocities
    # out over PCIe): reported beside the value, never as it
    fg, fg0 = np.concatenate([s["f"], s["g"]]), np.concatenate([t["f0"], t["g0"]])
    p_v(fg, fg0)
    t0 = time.perf_counter()
    for _ in range(2):
        p_v(fg, fg0)
    host_ms = 1e3 * (time.perf_counter() - t0) / 2

    # mixed-precision variant (FP32 far pairs, FP64 accumulation) of the fused
    # step: time and error against the FP64 step
    m_v = hor.Plan(ctx, hor.SLDL_OFF, s, t, DEL, precision=hor.MIXED)
    m_v.eval_dev(dev["fg"], dev["fg0"], out_v)
    tm = time.perf_counter()
    for _ in range(args.steps):
        m_v.eval_dev(dev["fg"], dev["fg0"], out_v)
    mixed_ms = 1e3 * (time.perf_counter() - tm) / args.steps
    mix_v = ctx.download(out_v, np.zeros(6 * nt))
    err_sl = float(np.linalg.norm(mix_v[:3 * nt] - ref_v[:3 * nt]) / np.linalg.norm(ref_v[:3 * nt]))
    err_dl = float(np.linalg.norm(mix_v[3 * nt:] - ref_v[3 * nt:]) / np.linalg.norm(ref_v[3 * nt:]))
    m_v.close()

    # error against the exact translating-body velocity: U = (1,0,0) inside, Oberbeck's flow outside
    inside, outside = t["b"] < 0, t["b"] > 0
    u_sl = ref_v[:3 * nt].reshape(-1, 3)
    exact = body.exact_velocity(t["x"]).reshape(-1, 3)
    e_node = np.abs(u_sl - exact).max(axis=1)
    accuracy = {"sl_max_err_inside_vs_exact_U": float(e_node[inside].max()) if inside.any() else None,
                "sl_max_err_outside_vs_exact": float(e_node[outside].max()) if outside.any() else None,
                "sl_rms_err_outside_vs_exact": float(np.sqrt(np.mean(((u_sl - exact)[outside]) ** 2)))
                if outside.any() else None,
                "nodes_inside": int(inside.sum()), "nodes_outside": int(outside.sum()),
                "note": "FP64 single layer of the exact traction at every band node against the exact velocity of "
                        "the ellipsoid translating with U = (1,0,0) (Oberbeck's exterior flow, |U| = 1)"}

    # roofline of the dominant kernel (the fused SL + DL direct sum):
    # algorithmic FP64 flops of both kernels per launch / average launch time
    # (HIP events on the plan's stream)
    fnear = near_pair_fraction(s, t, DEL)
    dl_flops = ns * nt * (fnear * OPS["dl_reg"] + (1 - fnear) * OPS["dl_plain"])
    sl_flops = ns * nt * (fnear * OPS["sl_reg"] + (1 - fnear) * OPS["sl_plain"])
    kv_ms = float(np.mean(kv))
    kdl_ms = float(np.mean(kdl))
    ksl_ms = float(np.mean(ksl))
    achieved = (sl_flops + dl_flops) / (kv_ms * 1e-3) / 1e12
    roof = {"bound": "valu", "compute": "FP64 vector ALU issue (peak 78.6 TFLOP/s, equal to the FP64 matrix peak)",
            "kernel": "k_direct_sldl", "achieved": achieved, "peak": PEAK_FP64 / 1e12,
            "unit": "TFLOP/s", "frac": achieved * 1e12 / PEAK_FP64, "traffic": profile_traffic(),
            "avg_kernel_ms": kv_ms, "flop_per_launch": sl_flops + dl_flops,
            "near_pair_fraction": fnear,
            "separate": {"sl_kernel_ms": ksl_ms, "dl_kernel_ms": kdl_ms,
                         "sl_achieved_tflops": sl_flops / (ksl_ms * 1e-3) / 1e12,
                         "dl_achieved_tflops": dl_flops / (kdl_ms * 1e-3) / 1e12}}
    counted = profile_valu()
    if counted and "k_direct_sldl" in counted:
        c = counted["k_direct_sldl"]
        t_v = kv_ms * 1e-3
        roof["counted"] = {
            "fp64_flop_per_launch": c["fp64_flop_per_launch"],
            "fp64_tflops": c["fp64_flop_per_launch"] / t_v / 1e12,
            "fp64_frac": c["fp64_flop_per_launch"] / t_v / PEAK_FP64,
            "valu_issue_frac": c["counters"]["SQ_INSTS_VALU"] * 64 / t_v / VALU_ISSUE,
            "note": "hardware-counted FP64 FLOPs (rocprofv3 VALU counters, profiles/direct_band_h64_valu_pmc.json) "
                    "over this run's launch time; the FLOPs the kernel executes, including the 1/r refinement"}

    # regularized pressure (stokes_pressure_7ord) on the same band workload
    pp = hor.Plan(ctx, hor.P_OFF, s, t, DEL)
    out_p = ctx.alloc(8 * nt)
    pp.eval_dev(dev["f"], dev["f0"], out_p)
    kp = []
    tp0 = time.perf_counter()
    for _ in range(args.steps):
        pp.eval_dev(dev["f"], dev["f0"], out_p)
        kp.append(ctx.last_timing()[1])
    p_ms = 1e3 * (time.perf_counter() - tp0) / args.steps
    kp_ms = float(np.mean(kp))
    p_flops = ns * nt * (fnear * OPS["p_reg"] + (1 - fnear) * OPS["p_plain"])
    pressure = {"ms_per_eval": p_ms, "value": ns * nt / (p_ms * 1e-3), "unit": "pair interactions/s",
                "kernel": "k_pressure", "avg_kernel_ms": kp_ms,
                "achieved_tflops": p_flops / (kp_ms * 1e-3) / 1e12,
                "frac": p_flops / (kp_ms * 1e-3) / PEAK_FP64}
    pp.close()
    ctx.free(out_p)
    if not args.no_cpu:
        pressure["cpu_baseline"] = pressure_cpu_baseline(s, t, DEL)

    # the grid program end to end: boundary-integral pressure and velocity on
    # the GPU, extension to the whole grid on the host
    grid = None
    if not args.no_grid:
        from hor import grid as HG
        log("grid program (translating spheroid, h=1/64)")
        t0 = time.perf_counter()
        r = HG.grid_program(ctx, N=192)
        grid = {"workload": "grid program: spheroid (1,0.5,0.5) translating in [0,3]^3, h=1/64 (193^3 nodes)",
                "sources": r["sources"], "band_nodes": r["band_nodes"], "seconds": time.perf_counter() - t0,
                "times_s": r["times"], "errors": r["errors"]}

    cpu = None if args.no_cpu else cpu_baseline(s, t, DEL)
    gm = None
    if not args.no_gmres:
        log("two-sphere GMRES workload (h=1/128, treecode P=8 N0=4000)")
        gm = {"h128": gmres_workload(cpu=not args.no_cpu)}
        for gap in (0.05, 0.025):
            log(f"two-sphere GMRES workload (h=1/128, gap {gap})")
            gm[f"h128_gap{gap}"] = gmres_workload(gap=gap)
        if not args.no_h256:
            log("two-sphere GMRES workload (h=1/256, treecode P=10 N0=8000)")
            gm["h256"] = gmres_workload(h=1.0 / 256.0, P=10, N0=8000, direct_check=True,
                                        cpu=not (args.no_cpu or args.no_cpu_h256), cpu_reps=2,
                                        geometry=not args.no_cpu)
    line = {
        "metric": "pair interactions/s",
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "extend-to-grid band: ellipsoid (1,0.8,0.6), h=1/64, box [-1.5,1.5]^3, "
                               "|b|<=0.3, FP64 direct SL_OFF + DL_OFF (one fused pass)",
                   "sources": ns, "targets": nt, "delta": DEL, "pairs_per_step": pairs_per_step},
        "roofline": roof,
        "gmres_roofline": (gm or {}).get("h256", {}).get("roofline"),
        "cpu_baseline": cpu,
        "separate_ops": {"ms_per_step": sep_ms, "value": pairs_per_step / (sep_ms * 1e-3),
                         "max_rel_diff_vs_fused": fused_vs_separate,
                         "note": "the same step as an SL_OFF and a DL_OFF evaluation"},
        "host_buffers": {"ms_per_step": host_ms, "value": pairs_per_step / (host_ms * 1e-3),
                         "note": "hor_plan_eval with host arrays: density upload, sums, velocity download"},
        "mixed": {"ms_per_step": mixed_ms, "value": pairs_per_step / (mixed_ms * 1e-3),
                  "rel_err_vs_fp64_sl": err_sl, "rel_err_vs_fp64_dl": err_dl},
        "accuracy": accuracy,
        "pressure": pressure,
        "grid_program": grid,
        "gmres": gm,
    }
    for p in dev.values():
        ctx.free(p)
    ctx.free(out_sl)
    ctx.free(out_dl)
    ctx.free(out_v)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
