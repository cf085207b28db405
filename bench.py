#!/usr/bin/env python3
"""MG-PCG benchmark: DOF-iterations/s and time-to-solution of the equilibrium
solve K(rho) u = f on the 192x96x96 cantilever (5.4M DOFs, 5-level GMG),
FP64 preconditioner vs mixed (FP32 V-cycle inside FP64 PCG), on one MI355X.

One step = one full MG-PCG solve from x0 = 0 to ||r||/||f|| <= 1e-8 with the
hierarchy, densities and load already resident in HBM (setup is timed and
reported separately, as the reference's setupK / setupPrec / system split).
Prints one JSON line.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "topopt-slp-3d_amd"))

HBM_PEAK_GBS = 8000.0
# Serving target, stated ahead of the measurement: p99 latency (arrival to
# result) of the 256-request stream at 0.8 x single-stream capacity, about six
# times the stream's mean serial service time (~16 ms per request)
SERVE_P99_TARGET_MS = 100.0
FP64_VEC_PEAK_TFLOPS = 78.6
FP32_VEC_PEAK_TFLOPS = 78.6   # packed FP32 is disabled in this build; unpacked FP32 = FP64 vector rate


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", default="192x96x96")
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-slp", action="store_true", help="skip the bounded sample of the multiresolution SLP loop")
    ap.add_argument("--slp-sample-iters", type=int, default=2,
                    help="SLP iterations of the multiresolution sample in the default run")
    ap.add_argument("--solve-only", default=None, choices=["fp64", "mixed"],
                    help="run only this solve (profiling aid; prints no JSON line)")
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--workload", default="solve", choices=["solve", "slp"],
                    help="solve: MG-PCG solves + serving stream (default); slp: the multiresolution SLP loop")
    ap.add_argument("--slp-iters", type=int, default=30)
    ap.add_argument("--slp-grid", default="240x120x120", help="displacement mesh of the SLP workload")
    ap.add_argument("--slp-cycle", type=int, default=1, choices=[0, 1, 2],
                    help="mg.cycle of the SLP workload (1 = W, the reference's default, Implementation/TopOpt3D.m:172)")
    ap.add_argument("--probe-kernel", type=int, default=None,
                    help="profiling aid: only launch this solver kernel (tmg_time_kernel id) --kernel-reps times "
                         "on the workload and print its average; no JSON bench line")
    ap.add_argument("--probe-precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--probe-mr", type=int, default=1, choices=[1, 2],
                    help="probe on the multiresolution problem (mr.n = 2, W-cycle, uniform density 0.3) of --grid")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="profiling aid: only run three elementwise kernels (scale, copy, add) on vectors of this "
                         "many bytes (calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def profile_traffic(name):
    """HBM bytes per launch recorded in a committed rocprofv3 counter summary
    (line 'traffic_bytes_per_launch: N'), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prob, x, L, gpu_iters):
    """FP64 CPU restatement of the reference's assembled GMG-PCG path (the
    oracle) timed on this host, on the benchmark's own workload: setup
    (assembly + Galerkin + coarse Cholesky) and one full solve on every
    allotted core, then a bounded one-core sample (2 PCG iterations from
    the same start)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    prev = oracle.set_threads(cores)
    o = oracle.Oracle(prob)
    t0 = time.perf_counter()
    o.assemble(x)
    o.setup_gmg(L, 0.6)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    u, it, rr = o.solve(prob.f)
    t_all = time.perf_counter() - t0
    oracle.set_threads(1)
    k1 = 2
    t0 = time.perf_counter()
    o.solve(prob.f, maxiter=k1)
    t_one = time.perf_counter() - t0
    oracle.set_threads(prev)
    v_all = prob.nfree * (it - 1) / t_all
    v_one = prob.nfree * k1 / t_one
    return {
        "value": v_all,
        "unit": "DOF-iters/s",
        "cores": cores,
        "kind": "port",
        "one_core_value": v_one,
        "sample": (f"FP64 C/OpenMP restatement of the reference's assembled SolveMG path (the oracle; baseline only, "
                   f"not MATLAB) on the benchmark's workload ({prob.nelx}x{prob.nely}x{prob.nelz}, {prob.nfree} "
                   f"free DOFs, {L}-level V(1,1) Jacobi, tol 1e-8) on {cores} threads of a {cpu_model()}: one full "
                   f"solve, {it - 1} PCG iterations (GPU FP64: {gpu_iters - 1}), {t_all:.2f} s; setup (assembly + "
                   f"Galerkin + coarse Cholesky) {t_setup:.2f} s; one-core sample of {k1} PCG iterations "
                   f"{t_one:.2f} s = {v_one:.3e} DOF-iters/s"),
    }


def iteration_bytes(prob, L, prec):
    """Algorithmic HBM bytes of one PCG iteration (each array the kernels read
    or write counted once per pass; no halo or cache effects), for the
    solver's schedule: FP64 recurrences and q = K d (x updated one iteration
    late inside the direction update, the cycle's first Jacobi sweep inside
    the residual update); the fine level and level 1 matrix-free (moduli and
    free-DOF mask only); assembled symmetric stencils (126 coefficients per
    node) below; the dense coarse inverse (its FP32 copy in the mixed cycle)."""
    s = 4 if prec == "mixed" else 8
    n0, ne0 = prob.ndof, prob.nelem
    b = 8 * (2 * n0 + ne0) + n0       # q = K d (read d, E, mask; write q)
    b += (24 + 2 * s) * n0            # r -= alpha q, ||r||^2, x0 = invD r (read r, q, invD; write r, x0)
    b += (s + 32) * n0                # x += alpha d_old, d = s + beta d_old (read s, d, x; write d, x)
    b += s * (2 * n0 + ne0) + 9 * n0  # fine residual (read x, E, r, mask; write)
    b += s * (3 * n0 + ne0) + 8 * n0  # fine post-sweep (read x, invD, E, r; write)
    ex, ey, ez = prob.nelx, prob.nely, prob.nelz
    nprev = n0
    for l in range(1, L):
        ex, ey, ez = ex // 2, ey // 2, ez // 2
        n = 3 * (ex + 1) * (ey + 1) * (ez + 1)
        b += s * (nprev + n)          # restriction into level l
        b += s * (2 * nprev + n)      # prolongation back onto level l - 1
        if l == L - 1:
            b += s * n * n + 2 * s * n     # dense coarse inverse times the coarse residual
        elif l == 1:
            b += s * (3 * n + 8 * ex * ey * ez) * 2 + s * 3 * n   # two matrix-free sweeps + sweep from zero
        else:
            b += 2 * s * (42 * n + 3 * n)  # two stencil sweeps (126 coefficients per node = 42 per DOF)
        nprev = n
    return b


def probe(args):
    """Profiling aid: launch one solver kernel (or a copy of known size)
    --kernel-reps times; prints its HIP-event average. Kernel ids are those of
    tmg_time_kernel (0 FP64 fine apply q = K d, 1 fine Jacobi sweep, 2 second-level
    sweep, 3 V-cycle, 4/5 fused fine residual / post-sweep, 6 fused direction
    update + apply, 7/8 fine residual / post-sweep on the FP64 PCG residual)."""
    import torch
    torch.cuda.init()
    if args.copy_bytes:
        # three kernels of known HBM traffic on FP32 vectors of copy_bytes
        # each: scale (read 1, write 1), copy (read 1, write 1), add (read 2,
        # write 1); their counter readings calibrate FETCH_SIZE / WRITE_SIZE
        n = args.copy_bytes // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        src2 = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(args.kernel_reps):
            torch.mul(src, 2.0, out=dst)
        for _ in range(args.kernel_reps):
            dst.copy_(src2)
        for _ in range(args.kernel_reps):
            torch.add(src, src2, out=dst)
        torch.cuda.synchronize()
        log(json.dumps({"copy_bytes": args.copy_bytes, "reps": args.kernel_reps,
                        "kernels": ["mul (r1 w1)", "copy (r1 w1)", "add (r2 w1)"]}))
        return
    import tmg
    from tmg import problem as P
    nelx, nely, nelz = (int(v) for v in args.grid.split("x"))
    prob = P.cantilever(nelx, nely, nelz)
    if args.probe_mr == 2:
        import numpy as np
        x = np.full(8 * prob.nelem, 0.3)
        s = tmg.Solver(prob, args.levels, precision=args.probe_precision, mr=2, cycle=1)
    else:
        x = P.random_density(nelx, nely, nelz, args.seed)
        s = tmg.Solver(prob, args.levels, precision=args.probe_precision)
    s.update_density(x)
    ms = s.time_kernel(args.probe_kernel, args.kernel_reps)
    log(json.dumps({"kernel": args.probe_kernel, "precision": args.probe_precision, "mr": args.probe_mr,
                    "avg_ms": ms, "reps": args.kernel_reps, "n_dof": prob.ndof, "n_el": prob.nelem}))


def slp_workload(args):
    """The multiresolution SLP loop: 240x120x120 displacement elements
    (10.5M DOFs), 2x2x2 density sub-elements (27.6M design variables),
    volfrac 0.3, filter radius 1.5, 4-level GMG, mixed precision; solve,
    Galerkin refresh, filter and gradient on the GPU, the LP on the host."""
    import torch

    from tmg import problem as P
    from tmg import slp
    torch.cuda.init()
    nelx, nely, nelz = (int(v) for v in args.slp_grid.split("x"))
    prob = P.cantilever(nelx, nely, nelz)
    log(f"SLP: {nelx}x{nely}x{nelz} MR 2, {prob.nfree} free DOFs, {8 * prob.nelem} design variables")
    t0 = time.perf_counter()
    res = slp.StrSLP(prob, 0.3, p=3.0, emin=1e-9, opfilter=2, rmin=1.5, ngrids=4, precision="mixed", mr=2,
                     slp=slp.SLPParams(maxiter=args.slp_iters), log=log, cycle=args.slp_cycle)
    wall = time.perf_counter() - t0
    nsolve = res.ns
    line = {
        "metric": f"SLP iterations/s (MR 2x2x2, {nelx}x{nely}x{nelz}, mixed precision GMG-PCG, host LP)",
        "value": round(res.iter / wall, 4), "unit": "accepted SLP iterations/s", "n_gpus": 1,
        "steps": res.iter, "warmup": 0, "ms_per_step": round(1e3 * wall / max(res.iter, 1), 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (PCG, LP, gradient) + f32 (V-cycle)", "data": "synthetic",
        "config": {
            "workload": f"cantilever {nelx}x{nely}x{nelz} Q1 displacement mesh, multiresolution 2x2x2 density/design "
                        f"sub-elements ({2 * nelx}x{2 * nely}x{2 * nelz}), volfrac 0.3, linear filter rmin 1.5, "
                        f"4-level GMG {'VWF'[args.slp_cycle]}(1,1) cycle, tol 1e-8, maxiterMG 5000",
            "n_free_dofs": prob.nfree, "design_variables": 8 * prob.nelem,
            "mg_cycle": "VWF"[args.slp_cycle], "iterations": res.iter, "rejected": res.itrej, "solves": nsolve,
            "pcg_iterations": res.pcgit, "unconverged_solves": int(sum(1 for it in res.pcgit if it > 5000)),
            "F": res.F, "vol": res.vol, "opstop": res.opstop, "wall_s": round(wall, 3),
            "time_s": {kk: round(v, 3) for kk, v in res.time.items()},
            "F_history": [float(f"{h['F']:.8g}") for h in res.history],
        },
    }
    print(json.dumps(line), flush=True)


def slp_sample_leg(args):
    """A bounded sample of the multiresolution SLP workload inside the default
    run: the first args.slp_sample_iters accepted iterations of the 240x120x120
    loop (2x2x2 density sub-elements, W-cycle, mixed precision) that
    bench.py --workload slp runs for 30 iterations."""
    from tmg import problem as P
    from tmg import slp
    nelx, nely, nelz = (int(v) for v in args.slp_grid.split("x"))
    prob = P.cantilever(nelx, nely, nelz)
    log(f"SLP sample: {nelx}x{nely}x{nelz} MR 2, {args.slp_sample_iters} iterations ...")
    t0 = time.perf_counter()
    res = slp.StrSLP(prob, 0.3, p=3.0, emin=1e-9, opfilter=2, rmin=1.5, ngrids=4, precision="mixed", mr=2,
                     slp=slp.SLPParams(maxiter=args.slp_sample_iters), log=lambda *a, **kw: None,
                     cycle=args.slp_cycle)
    wall = time.perf_counter() - t0
    out = {
        "workload": f"cantilever {nelx}x{nely}x{nelz} Q1, multiresolution 2x2x2 density sub-elements "
                    f"({8 * prob.nelem} design variables, {prob.nfree} free DOFs), volfrac 0.3, rmin 1.5, "
                    f"4-level GMG {'VWF'[args.slp_cycle]}(1,1), tol 1e-8, mixed precision; the first "
                    f"{args.slp_sample_iters} iterations of the loop bench.py --workload slp runs",
        "iterations": res.iter, "rejected": res.itrej, "pcg_iterations": res.pcgit,
        "unconverged_solves": int(sum(1 for it in res.pcgit if it > 5000)),
        "F": res.F, "vol": res.vol, "wall_s": round(wall, 3),
        "time_s": {kk: round(v, 3) for kk, v in res.time.items()},
    }
    log("SLP sample:", json.dumps(out))
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "slp":
        slp_workload(args)
        return
    if args.probe_kernel is not None or args.copy_bytes:
        probe(args)
        return
    import numpy as np
    import torch

    import tmg
    from tmg import problem as P

    nelx, nely, nelz = (int(v) for v in args.grid.split("x"))
    L = args.levels
    torch.cuda.init()
    dev = torch.device("cuda:0")
    log(f"problem {nelx}x{nely}x{nelz}, {L} levels, seed {args.seed}")
    prob = P.cantilever(nelx, nely, nelz)
    x = P.random_density(nelx, nely, nelz, args.seed)
    f_full = torch.from_numpy(prob.full(prob.f)).to(dev)
    # the FP64 oracle's solve of this workload (minted in the build container
    # by tests/golden/make_golden.py); every bench line states each variant's
    # error against it
    gpath = os.path.join(ROOT, "tests", "golden", f"oracle_{nelx}x{nely}x{nelz}_seed{args.seed}.json")
    if not os.path.exists(gpath):
        raise SystemExit(f"bench.py: no oracle fixture {gpath}; mint it with tests/golden/make_golden.py --only "
                         f"{nelx}x{nely}x{nelz}")
    golden = json.load(open(gpath))
    gidx = np.asarray(golden["u_sample_index"])
    gu = np.asarray(golden["u_sample"])

    results = {}
    solvers = {}
    for prec in ((args.solve_only,) if args.solve_only else ("fp64", "mixed")):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s = tmg.Solver(prob, L, precision=prec)
        t1 = time.perf_counter()
        s.update_density(x)
        t2 = time.perf_counter()
        solvers[prec] = s
        xd = torch.zeros(prob.ndof, dtype=torch.float64, device=dev)
        for _ in range(args.warmup):
            xd.zero_()
            s.solve_device(xd, f_full)
        torch.cuda.synchronize()
        times, its = [], []
        for _ in range(args.steps):
            xd.zero_()
            torch.cuda.synchronize()
            ts = time.perf_counter()
            it, rr, rc = s.solve_device(xd, f_full)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - ts)
            its.append(it)
        if args.solve_only:
            log(f"{prec}: it={its[-1]} t_solve={1e3 * sum(times) / len(times):.3f} ms")
            return
        u = xd.cpu().numpy()[prob.freedofs - 1]
        # true residual with the FP64 operator of the FP64 hierarchy
        Ku = solvers["fp64"].apply(1, u)
        relres_true = float(np.linalg.norm(prob.f - Ku) / np.linalg.norm(prob.f))
        C = float(prob.f @ u)
        t_solve = sum(times) / len(times)
        bytes_it = iteration_bytes(prob, L, prec)
        results[prec] = dict(it=its[-1], relres=rr, relres_true=relres_true, compliance=C, t_solve=t_solve,
                             t_setup_structure=t1 - t0, t_setup_operator=t2 - t1,
                             dof_iters=prob.nfree * (its[-1] - 1) / t_solve,
                             it_oracle=golden["it"],
                             compliance_rel_err_vs_oracle=abs(C - golden["compliance"]) / abs(golden["compliance"]),
                             u_sample_rel_err_vs_oracle=float(np.linalg.norm(u[gidx] - gu) / np.linalg.norm(gu)),
                             algorithmic_bytes_per_iteration=bytes_it,
                             solve_algorithmic_gbs=bytes_it * (its[-1] - 1) / t_solve / 1e9)
        log(f"{prec}: it={its[-1]} t_solve={t_solve * 1e3:.2f} ms relres={rr:.2e} true={relres_true:.2e} C={C:.10e} "
            f"setup={t1 - t0:.3f}+{t2 - t1:.3f} s DOF-iters/s={results[prec]['dof_iters']:.3e}")

    results["mixed"]["compliance_rel_err_vs_fp64"] = abs(results["mixed"]["compliance"] - results["fp64"]["compliance"]) / abs(results["fp64"]["compliance"])

    # dominant kernels, HIP events on the solver's stream
    n_dof, n_el = prob.ndof, prob.nelem
    sm = solvers["mixed"]
    sf = solvers["fp64"]
    reps = args.kernel_reps
    k = {
        "fine_apply_fp64_ms": sf.time_kernel(0, reps),
        "fine_resid_fp32_ms": sm.time_kernel(7, reps),
        "fine_post_fp32_ms": sm.time_kernel(8, reps),
        "fine_resid_fp64_ms": sf.time_kernel(7, reps),
        "fine_post_fp64_ms": sf.time_kernel(8, reps),
        "fine_jacobi_fp32_ms": sm.time_kernel(1, reps),
        "fine_jacobi_fp64_ms": sf.time_kernel(1, reps),
        "l2_jacobi_fp32_ms": sm.time_kernel(2, reps) if L >= 3 else None,
        "l2_jacobi_fp64_ms": sf.time_kernel(2, reps) if L >= 3 else None,
        "vcycle_mixed_ms": sm.time_kernel(3, max(reps // 5, 3)),
        "vcycle_fp64_ms": sf.time_kernel(3, max(reps // 5, 3)),
    }
    # the same two kernels timed inside a solve (HIP events around each of
    # their launches in one eagerly launched mixed solve): the back-to-back
    # loops above keep their ~100 MB working set in the 256 MB Infinity
    # Cache, the solve does not, so these are the figures of the solve
    t_apply_is, n_apply_is = sm.time_in_solve(0)
    t_post_is, n_post_is = sm.time_in_solve(8)
    k["fine_apply_fp64_in_solve_ms"] = t_apply_is
    k["fine_post_fp32_in_solve_ms"] = t_post_is
    k["in_solve_launches"] = n_apply_is
    log("kernels:", json.dumps(k))
    # Dominant solve kernel: the FP64 fine-level operator apply q = K(rho) d
    # with the PCG dot d'q folded in (once per iteration in both modes).
    # Algorithmic bytes: read d, write q (8 B x n_dof each), read E (8 B x n_el).
    # Arithmetic of the spectral element form per element: plane transform 24,
    # z butterflies 21, 45 coefficients x 2, modulus 21, inverse z 36, inverse
    # plane transform 24, corner sums 9 = 225 flop (a dense 24x24 Ke product
    # is 1152).
    flops_apply = 225 * n_el
    bytes_apply64 = 8 * 2 * n_dof + 8 * n_el
    t = k["fine_apply_fp64_in_solve_ms"] * 1e-3
    achieved = bytes_apply64 / t / 1e9
    roofline = {
        "bound": "hbm", "kernel": "k_ebe_lds<double,DOTU> (fine-level FP64 operator apply q = K d with d'q, spectral element form, separable node exchange)",
        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": profile_traffic("k_ebe_lds_apply_fp64_cantilever192x96x96_hbm_traffic.txt"),
        "bytes_per_launch": bytes_apply64, "avg_launch_ms": round(k["fine_apply_fp64_in_solve_ms"], 5),
        "avg_launch_ms_measured": f"HIP events on the solver's stream around each of the {n_apply_is} launches "
                                  f"in one mixed solve",
        "avg_launch_ms_back_to_back": round(k["fine_apply_fp64_ms"], 5),
        "valu_tflops": round(flops_apply / t / 1e12, 2), "valu_peak_tflops": FP64_VEC_PEAK_TFLOPS,
        "valu_frac": round(flops_apply / t / 1e12 / FP64_VEC_PEAK_TFLOPS, 4),
    }
    # the preconditioner's fine-level post-sweep (FP32), the largest fine
    # V-cycle kernel: x += omega invD (r - K x) with r the FP64 PCG residual;
    # read x, invD (4 B), r (8 B), E (4 B per element); write x
    bytes_post32 = (4 + 4 + 8 + 4) * n_dof + 4 * n_el
    t32 = k["fine_post_fp32_in_solve_ms"] * 1e-3
    secondary = {
        "kernel": "k_ebe_lds<float,POSTR> (fine-level fused apply + damped Jacobi post-sweep on the FP64 residual, FP32)",
        "achieved_gbs": round(bytes_post32 / t32 / 1e9, 1), "frac": round(bytes_post32 / t32 / 1e9 / HBM_PEAK_GBS, 4),
        "bytes_per_launch": bytes_post32, "avg_launch_ms": round(k["fine_post_fp32_in_solve_ms"], 5),
        "avg_launch_ms_back_to_back": round(k["fine_post_fp32_ms"], 5),
        "traffic": profile_traffic("k_ebe_lds_post_fp32_cantilever192x96x96_hbm_traffic.txt"),
        "valu_tflops": round(flops_apply / t32 / 1e12, 2), "valu_peak_tflops": FP32_VEC_PEAK_TFLOPS,
        "valu_frac": round(flops_apply / t32 / 1e12 / FP32_VEC_PEAK_TFLOPS, 4),
    }
    serving = None
    if not args.no_serve:
        from tmg import serve
        log(f"serving: building {args.serve_requests} requests ...")
        reqs = serve.stream_workload(args.serve_requests, seed=4)
        serving = serve.benchmark(reqs, budget_bytes=200e9, max_active=16, cap_sample=32, load=0.8, log=log)
        serving = {kk: (float(f"{v:.6g}") if isinstance(v, float) else v) for kk, v in serving.items()}
        serving["p99_target_ms"] = SERVE_P99_TARGET_MS
        serving["p99_target_met"] = bool(serving["p99_ms"] <= SERVE_P99_TARGET_MS)
        log("serving:", json.dumps(serving))
    slp_sample = None
    if not args.no_slp:
        slp_sample = slp_sample_leg(args)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(prob, x, L, results["fp64"]["it"])
        log("cpu:", json.dumps(cpu))
    mixed, fp64 = results["mixed"], results["fp64"]
    line = {
        "metric": "MG-PCG DOF-iters/s + solve time @5.4M DOFs (FP64 vs mixed); served p99",
        "value": round(mixed["dof_iters"], 1),
        "unit": "DOF-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["t_solve"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (PCG, outer apply, coarse factorization) + f32 (V-cycle)",
        "data": "synthetic",
        "config": {
            "workload": f"cantilever {nelx}x{nely}x{nelz} Q1, {prob.nfree} free DOFs, {L}-level GMG V(1,1) "
                        f"Jacobi omega 0.6, filtered random densities seed {args.seed}, tol 1e-8",
            "n_free_dofs": prob.nfree,
            "fp64": {kk: (float(f"{v:.6g}") if isinstance(v, float) else v) for kk, v in fp64.items()},
            "mixed": {kk: (float(f"{v:.6g}") if isinstance(v, float) else v) for kk, v in mixed.items()},
            "speedup_mixed_vs_fp64_time_to_solution": round(fp64["t_solve"] / mixed["t_solve"], 3),
            "kernels_ms": k,
            "preconditioner_sweep": secondary,
            "serving": serving,
            "served_p99_ms": serving["p99_ms"] if serving else None,
            "served_p99_target_ms": SERVE_P99_TARGET_MS,
            "slp_sample": slp_sample,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    if serving and serving["failed"]:
        log(f"serving: {serving['failed']} request(s) failed: {serving.get('failed_requests')}")
        sys.exit(3)


if __name__ == "__main__":
    main()
