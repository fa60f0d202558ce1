"""Benchmark of the elasto-plastic load-stepping loop on one MI355X.

Default workload (config 3): quarter plate [0,1]^2 with a hole of radius
0.1, N_theta x N_r = 2000 x 1000 bilinear quads (2,000,000 elements,
2,003,001 nodes, 4,006,002 DOF), plane stress, E = 2.1e5 MPa, nu = 0.3, von
Mises secant (deformation theory) update with Swift hardening
sigma = sigma_Y (1 + eps_p/eps0)^0.2, sigma_Y = 240 MPa, eps0 = sigma_Y/E,
40 equal load increments of the far-edge traction up to 300 MPa,
Aitken-accelerated secant iterations to a 1e-6 stress tolerance, linear
solves to ||r||/||b|| = 1e-9.

A "step" is one load increment: warm-started from the previous one,
assemble -> CG -> return map until the secant iteration converges.  The
timed region runs the 40-increment ramp on a fresh device problem with all
inputs already in HBM; with --steps K < 40 the first 40 - K increments run
untimed first and the last K are timed.  W warm-up increments (--warmup)
run before, on a separate problem.  `value` is the mixed-precision path:
FP32 CG preconditioned by a geometric multigrid V-cycle, FP64 iterative
refinement, FP64 return map and Aitken.  The same line reports the FP64
path with the same solver settings, the plain FP64 GPU path (the
reference's algorithm: unpreconditioned CG to the full tolerance at every
secant iteration) over the first increments, the roofline of the dominant
kernel, the 632x316 plate (config 2) and the 1M-quad graded ring (config 4)
as sub-records, a served stream of independent problems at three arrival
rates (p50/p99 latency) and the CPU oracle on a capped sample.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "load-steps/s + CG GB/s at 1e-6 rel err; p50/p99 latency (mixed+served vs FP64)"
HBM_PEAK_GBS = 8000.0
NSTEPS = 20
# Fabric bytes per launch of the roofline kernel, measured with rocprofv3
# --pmc FETCH_SIZE and --pmc WRITE_SIZE on `bench.py --roofline-only`
# (632x316 quarter plate), each scaled by what the counter reads for a plain
# 256 MiB device copy (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000):
# the fused block-Jacobi CG iteration (--precond 1) 48.19 MB read + 8.05 MB
# written; the finest level's segment smoothing step (multigrid, the
# default) 44.93 MB read + 3.21 MB written; on the 2000x1000 plate
# (HBM-resident) the same step 475.8 MB read + 33.0 MB written in the
# sliced-ELL rows, 290.3 MB read + 32.1 MB written in the symmetric lattice
# layout (non-temporal matrix loads); the step that reads the combined
# iterate x + d (the one the V-cycle runs), with ordinary loads, 245.2 MB
# read + 32.7 MB written in that layout; the multigrid CG's operator
# product on that plate (FP32, symmetric lattice layout) 196.8 MB read +
# 16.2 MB written.
MEASURED_TRAFFIC = {(632, 316, "cg"): 56.24e6, (632, 316, "mg_smooth_seg"): 48.14e6,
                    (2000, 1000, "mg_smooth_seg"): 508.8e6, (2000, 1000, "mg_smooth_seg_sym"): 322.4e6,
                    (2000, 1000, "mg_cheb_x_seg_sym"): 277.9e6, (2000, 1000, "mg_cg_dots_sym"): 213.0e6}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_case(A, Nt, Nr, S=300.0):
    mesh = A.gen_quarter_plate(1.0, 0.1, Nt, Nr)
    loads = A.Loads(mesh).force(1, S, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    E = 2.1e5
    mat = A.material(E, 0.3, 240.0, A.LAW_LINEAR, K_T=E / 101)
    return mesh, loads, mat


def solver_options(A, args, precision, reference=False, precond=None, tol=1e-6):
    if reference:
        # the reference's algorithm: unpreconditioned CG to cg_tol at every
        # secant iteration, previous level's state as the initial guess
        return A.options(cg_tol=args.cg_tol, tol=tol, aitken=True, max_outer=500, precision=precision)
    return A.options(cg_tol=args.cg_tol, tol=tol, aitken=True, max_outer=500, precision=precision,
                     precond=args.precond if precond is None else precond, predictor=1, cg_forcing=args.forcing,
                     inner_tol=args.inner_tol)


GRAPHS = True   # --no-graphs: multigrid CG launches its kernels one by one (for kernel traces)


def run_ramp(A, ctx, mesh, loads, mat, ksteps, opt, tag, grid=None):
    """First `ksteps` increments of the 20-increment ramp; (seconds, stats, problem)."""
    t0 = time.perf_counter()
    P = A.Problem(ctx, mesh, loads, mat)
    if opt.precond == 2:
        P.set_grid(*grid)
        if not GRAPHS:
            P.set_graphs(False)
    P.setup_seconds = time.perf_counter() - t0
    per = []
    t0 = time.perf_counter()
    for n in range(1, ksteps + 1):
        o = A.Options.from_buffer_copy(opt)
        o.warm = 1 if n > 1 else 0
        if n == 1:
            P.set_load(n / NSTEPS)
        else:
            P.advance_load(n / NSTEPS, int(opt.predictor))
        st = P.solve(o)
        per.append(st)
        log(f"  {tag} increment {n}/{NSTEPS}: secant {st.outer_iters} cg {st.cg_iters} refinements "
            f"{st.refinements} rel_err {st.rel_err:.2e} {st.seconds:.3f} s")
    return time.perf_counter() - t0, per, P


def cg_iteration_bytes(nnz, N, s):
    """Algorithmic bytes of one CG iteration over a CSR matrix with nnz scalar
    entries and N unknowns, s-byte values: SpMV nnz(s+4) + 4(N+1) + 2Ns (values,
    columns, row pointers, p read once, Ap written once) plus the vector
    updates 9Ns and the one-byte Dirichlet mask."""
    return nnz * (s + 4) + 4 * (N + 1) + 2 * N * s + 9 * N * s + N


def scalar_nnz(A, P):
    import ctypes as C
    n = C.c_longlong()
    A.check(A.lib().cp_problem_csr_nnz(P.ptr, C.byref(n)))
    return n.value


def cheb_step_bytes(nnz, N, s):
    """Algorithmic bytes of one Chebyshev smoothing step of the finest level
    in the reference's CSR structure: the SpMV nnz(s+4) + 4(N+1) + 2Ns, the
    second input vector (the direction) and the right-hand side Ns each, the
    2x2 block-diagonal inverse 1.5Ns, the second output Ns and the mask N."""
    return nnz * (s + 4) + 4 * (N + 1) + 2 * N * s + N * s + N * s + 1.5 * N * s + N * s + N


def smooth_kernel(P):
    """Name and traffic key of the finest level's FP32 Chebyshev step kernel
    (row template -2: the symmetric lattice layout, 9: sliced-ELL rows; step
    mode 9 on a segment level: the step that reads the combined iterate
    x + d, 1 on a point-smoother level)."""
    w, tag = (-2, "_sym") if P.mg_level0_sym() else (9, "")
    if P.mg_smoothers()[0] == 1:
        return "cpk::k_mg_smooth_seg<%d,9,float>" % w, "mg_cheb_x_seg" + tag
    return "cpk::k_mg_smooth<%d,1,float>" % w, "mg_smooth" + tag


def hbm_copy(A, ctx, nbytes=1 << 30, reps=20):
    """Achievable HBM bandwidth: a plain device copy of 1 GiB (four times the
    Infinity Cache; read once and written once per launch) and a read-only
    pass over 1 GiB."""
    ms, msr = A.bench_copy(ctx, nbytes, reps)
    gbps, rgbps = 2 * nbytes / ms / 1e6, nbytes / msr / 1e6
    return {"bytes": nbytes, "copy_ms": ms, "copy_gbps": gbps, "copy_frac_of_peak": gbps / HBM_PEAK_GBS,
            "read_ms": msr, "read_gbps": rgbps, "read_frac_of_peak": rgbps / HBM_PEAK_GBS}


def bench_mg(A, P, precision, reps):
    import ctypes as C
    v, sm, b = C.c_double(), C.c_double(), C.c_double()
    A.check(A.lib().cp_bench_mg(P.ptr, precision, reps, C.byref(v), C.byref(sm), C.byref(b)))
    return v.value, sm.value, b.value


def bench_mg_spmv(A, P, precision, reps):
    """The multigrid CG's operator product alone (w = mask (A u) and the
    iteration's three dot products): average ms and the bytes one launch
    must move in level 0's layout."""
    import ctypes as C
    ms, b = C.c_double(), C.c_double()
    A.check(A.lib().cp_bench_mg_spmv(P.ptr, precision, reps, C.byref(ms), C.byref(b)))
    return ms.value, b.value


def spmv_kernel(P):
    """Name and traffic key of the multigrid CG's FP32 operator-product kernel."""
    w, tag = (-2, "_sym") if P.mg_level0_sym() else (9, "")
    return "cpk::k_mg_cg_dots<%d,float>" % w, "mg_cg_dots" + tag


def spmv_record(A, P, nt, nr, nnz, N, reps):
    """The SpMV of the north star on the path that runs: the multigrid CG's
    w = mask (A u) kernel (FP32), its layout bytes per launch over the mean
    launch time against the HBM peak, and the rocprofv3 traffic measured for
    it where there is one."""
    ms, fmt = bench_mg_spmv(A, P, A.MIXED, reps)
    kname, tkey = spmv_kernel(P)
    traffic = MEASURED_TRAFFIC.get((nt, nr, tkey))
    return {"kernel": kname + " (multigrid CG: w = mask (A u) with (r, u), (w, u), (r, r); FP32)",
            "avg_launch_ms": ms, "reps": reps, "bytes_per_launch": fmt,
            "bytes_model": ("what one launch must move in level 0's layout: the stored operator once (the "
                            "symmetric lattice layout: five 2x2 blocks per node), u and r read, w written, the "
                            "mask and the per-block partials"),
            "achieved": fmt / ms / 1e6, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": fmt / ms / 1e6 / HBM_PEAK_GBS,
            "traffic": traffic, "traffic_gbps": traffic / ms / 1e6 if traffic else None,
            # back-to-back launches over a working set below the 256 MB
            # Infinity Cache are served partly from it
            "working_set_bytes": fmt, "cache_resident": fmt < 256e6,
            "reference_csr_bytes_per_launch": nnz * 8 + 4 * (N + 1) + 2 * N * 4}


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def bench_cg(A, P, precision, reps, precond=True):
    import ctypes as C
    ms, b = C.c_double(), C.c_double()
    A.check(A.lib().cp_bench_cg(P.ptr, precision, int(precond), reps, C.byref(ms), C.byref(b)))
    return ms.value, b.value


def cpu_baseline(nt, nr, per_step, cores, what="the reference algorithm (FP64, unpreconditioned CG)"):
    """The oracle (CPU FP64 restatement of the reference solver, OpenMP) on
    the same mesh.  Two bounded runs of the first increment with 200 and 60200
    CG iterations give the cost of one CG iteration and of the rest of one
    secant iteration (pattern, B, assembly, return map); these are applied
    to the secant and CG iteration counts `per_step` of the GPU run of the
    reference algorithm (FP64, unpreconditioned CG) on the same increments."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    om = O.gen_quarter_plate(1.0, 0.1, nt, nr)
    ol = O.Problem(om).force(1, 300.0 / NSTEPS, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    E = 2.1e5
    mat = O.material(E, 0.3, 240.0, O.LAW_LINEAR, K_T=E / 101)
    t, n = {}, {}
    for it in (200, 60200):
        t0 = time.perf_counter()
        _, _, st = O.solve(ol, mat, O.options(cg_tol=1e-30, cg_maxit=it, fixed_outer=1))
        t[it] = time.perf_counter() - t0
        n[it] = st.cg_iters
        log(f"cpu baseline: first increment with {n[it]} CG iterations: {t[it]:.2f} s")
    per_it = (t[60200] - t[200]) / (n[60200] - n[200])
    per_outer = max(0.0, t[200] - n[200] * per_it)
    outer = sum(s.outer_iters for s in per_step)
    cg = sum(s.cg_iters for s in per_step)
    secs = outer * per_outer + cg * per_it
    return {"value": len(per_step) / secs, "unit": "load-steps/s", "cores": cores, "kind": "port",
            "sample": (f"oracle/cp_oracle.c (OpenMP, {cores} threads) on the same {nt}x{nr} mesh: the first "
                       f"increment run with 200 and 60200 CG iterations ({t[200] + t[60200]:.1f} s) gives "
                       f"{per_it * 1e3:.3f} ms per CG iteration and {per_outer:.3f} s per secant iteration "
                       f"(pattern, B, assembly, return map); applied to the {outer} secant and {cg} CG "
                       f"iterations of the FP64 GPU run of {what} on increments "
                       f"1..{len(per_step)}: {secs:.1f} s projected")}


def digest(*arrays):
    """sha256 over the arrays' bytes (-0.0 read as 0.0): equal digests are
    bitwise-equal results."""
    h = hashlib.sha256()
    for a in arrays:
        h.update((np.ascontiguousarray(a, dtype=np.float64) + 0.0).tobytes())
    return h.hexdigest()


def solve_alone(A, ctx, stream, opt, nsteps, keep=(), times=None):
    """Every problem of the stream solved alone, back to back: total seconds
    and the digest of each result (u, sigma_i); the results of the problems
    in `keep` and (list `times`) each problem's seconds are kept too."""
    from cudaplasticity_amd import serving
    t0 = time.perf_counter()
    sums = []
    kept = {}
    for k, (mesh, loads, mat, desc) in enumerate(stream):
        P = A.Problem(ctx, mesh, loads, mat)
        if opt.precond == 2:
            P.set_grid(*serving.lattice(desc))
        ts = time.perf_counter()
        per = P.solve_steps(nsteps, 1.0, opt)
        log(f"  alone {k + 1}/{len(stream)}: {desc['kind']} {desc['quads']} quads, secant "
            f"{sum(s.outer_iters for s in per)}, cg {sum(s.cg_iters for s in per)}, {time.perf_counter() - ts:.3f} s")
        if times is not None:
            times.append(time.perf_counter() - ts)
        sums.append(digest(P.field("u"), P.field("sigI")))
        if k in keep:
            kept[k] = (P.field("u"), P.field("sigI"))
        P.close()
    if keep:
        return time.perf_counter() - t0, sums, kept
    return time.perf_counter() - t0, sums


def serve_once(A, ctx, args, stream, opt, rate, sums):
    """The stream served at Poisson rate `rate`; latency percentiles,
    throughput and the bitwise check of every result against `sums`."""
    from cudaplasticity_amd import serving
    budget = int(args.serve_budget_gb * (1 << 30))
    recs = serving.serve(ctx, stream, args.serve_steps, opt, rate, budget, max_slots=args.serve_slots, log=log)
    ok = all(r["code"] == A.CP_OK for r in recs)
    same = ok and all(digest(r["u"], r["sig"]) == d for r, d in zip(recs, sums))
    lat = np.array([r["latency"] for r in recs])
    queue = np.array([r["queue"] for r in recs])
    span = max(r["t_done"] for r in recs)
    return {"arrival_rate_per_s": rate, "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "p99_queue_s": float(np.percentile(queue, 99)),
            "throughput_problems_per_s": len(stream) / span,
            "throughput_load_steps_per_s": len(stream) * args.serve_steps / span,
            "all_ok": ok, "results_equal_alone_bitwise": bool(same)}


def served_stream(A, ctx, args, precond=2, n=None, qmax=None):
    """Reduced config-5 stream: served with Poisson arrivals at `serve_load`
    times the rate the same problems are solved back to back one at a time,
    and every served result compared with the same problem solved alone.
    precond 2: multigrid CG (each request replays its own V-cycle graphs);
    precond 1: block-Jacobi CG batched across requests."""
    from cudaplasticity_amd import serving
    n = args.serve_n if n is None else n
    stream = serving.make_stream(n, qmin=2e3, qmax=args.serve_qmax if qmax is None else qmax)
    opt = solver_options(A, args, A.MIXED, precond=precond, tol=args.serve_tol)
    t_seq, sums = solve_alone(A, ctx, stream, opt, args.serve_steps)
    rate = args.serve_load * len(stream) / t_seq
    log(f"served stream: {len(stream)} problems solved back to back in {t_seq:.2f} s; serving at {rate:.2f}/s")
    rec = serve_once(A, ctx, args, stream, opt, rate, sums)
    rec.pop("p99_queue_s")
    out = {"problems": len(stream), "load_steps_each": args.serve_steps,
           "quads_range": [int(min(s[3]["quads"] for s in stream)), int(max(s[3]["quads"] for s in stream))]}
    out.update(rec)
    out["sequential_alone_load_steps_per_s"] = len(stream) * args.serve_steps / t_seq
    return out


def c5_workload(A, args):
    """Config 5 at full size: a stream of 512 problems (plate or graded
    ring, 2k-500k quads log-uniform, random hole radius, yield and hardening
    law, 5 increments each, mixed precision, multigrid CG), solved alone back
    to back, then served at Poisson arrival rates swept over multiples of the
    back-to-back rate.  `value` is the served throughput at the highest
    rate; every served result of every rate is checked bitwise against the
    problem solved alone."""
    from cudaplasticity_amd import serving
    ctx = A.Context(0)
    n = args.c5_n
    stream = serving.make_stream(n, qmin=2e3, qmax=args.serve_qmax)
    quads = np.array([s[3]["quads"] for s in stream])
    log(f"stream of {n} problems: {int(quads.sum())} quads in all, {quads.min()}-{quads.max()} each")
    opt = solver_options(A, args, A.MIXED, precond=2, tol=args.serve_tol)
    # the sample solved again with the plain FP64 GPU path
    sample = [k for k, s in enumerate(stream) if s[3]["quads"] <= args.c5_ref_qmax][:args.c5_ref_n]
    times = []
    t_seq, sums, kept = solve_alone(A, ctx, stream, opt, args.serve_steps, keep=set(sample), times=times)
    base = n / t_seq
    log(f"back to back: {t_seq:.2f} s, {base:.2f} problems/s")
    ref_opt = solver_options(A, args, A.FP64, reference=True, tol=args.serve_tol)
    f64_opt = solver_options(A, args, A.FP64, precond=2, tol=args.serve_tol)
    t_ref, eu, es, pu, ps = 0.0, [], [], [], []
    for k in sample:
        mesh, loads, mat, desc = stream[k]
        P = A.Problem(ctx, mesh, loads, mat)
        ts = time.perf_counter()
        per = P.solve_steps(args.serve_steps, 1.0, ref_opt)
        t_ref += time.perf_counter() - ts
        eu.append(rel(kept[k][0], P.field("u")))
        es.append(rel(kept[k][1], P.field("sigI")))
        P.close()
        # precision alone: FP64 with the served options
        P = A.Problem(ctx, mesh, loads, mat)
        P.set_grid(*serving.lattice(desc))
        P.solve_steps(args.serve_steps, 1.0, f64_opt)
        pu.append(rel(kept[k][0], P.field("u")))
        ps.append(rel(kept[k][1], P.field("sigI")))
        P.close()
        log(f"  FP64 reference algorithm {k}: {desc['kind']} {desc['quads']} quads, secant "
            f"{sum(s.outer_iters for s in per)}, cg {sum(s.cg_iters for s in per)}; mixed vs it: u {eu[-1]:.1e} "
            f"sigma_i {es[-1]:.1e}; mixed vs FP64 with the same options: u {pu[-1]:.1e} sigma_i {ps[-1]:.1e}")
    t_mix = sum(times[k] for k in sample)
    ref_rate = len(sample) * args.serve_steps / t_ref
    sweep = []
    for load in args.serve_loads:
        log(f"serving at {load:g} x the back-to-back rate ({load * base:.2f}/s)")
        rec = serve_once(A, ctx, args, stream, opt, load * base, sums)
        rec["load"] = load
        log("  " + json.dumps(rec))
        sweep.append(rec)
    top = sweep[-1]
    ok = all(r["all_ok"] and r["results_equal_alone_bitwise"] for r in sweep)
    line = {"metric": METRIC, "value": top["throughput_load_steps_per_s"], "unit": "load-steps/s", "n_gpus": 1,
            "steps": n * args.serve_steps, "warmup": 0, "ms_per_step": 1e3 / top["throughput_load_steps_per_s"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": (f"served stream of {n} problems, plate with hole or graded ring uniformly, "
                                    f"2k-{args.serve_qmax:.0f} quads log-uniform, hole radius U(0.05,0.3), yield "
                                    f"U(200,400) MPa, linear or Swift hardening, {args.serve_steps} increments each, "
                                    f"Poisson arrivals; mixed precision, multigrid CG, secant iterations to "
                                    f"{args.serve_tol:g}"),
                       "problems": n, "quads_total": int(quads.sum()),
                       "quads_range": [int(quads.min()), int(quads.max())], "max_slots": args.serve_slots,
                       "hbm_budget_gb": args.serve_budget_gb},
            "sequential_alone_load_steps_per_s": n * args.serve_steps / t_seq,
            "fp64_reference_algorithm_sample": {
                "what": ("the plain FP64 GPU path (the reference's algorithm: unpreconditioned FP64 CG to cg_tol "
                         "at every secant iteration, no predictor) on the first problems of the stream of at most "
                         f"{args.c5_ref_qmax:.0f} quads"),
                "problems": len(sample), "quads": [int(stream[k][3]["quads"]) for k in sample],
                "secant_tol": args.serve_tol,
                "load_steps_per_s": ref_rate,
                "mixed_alone_load_steps_per_s_same_problems": len(sample) * args.serve_steps / t_mix,
                "mixed_alone_speedup": t_ref / t_mix,
                "served_capacity_vs_fp64_reference": top["throughput_load_steps_per_s"] / ref_rate,
                "max_rel_err_mixed_vs_reference": {"u": max(eu), "sigma_i": max(es)},
                "max_rel_err_mixed_vs_fp64_same_options": {"u": max(pu), "sigma_i": max(ps)}},
            "sweep": sweep, "all_results_equal_alone_bitwise": bool(ok),
            "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def other_workload(A, args):
    """The 2M-quad plate (Swift law, 40 increments to 300 MPa) or the 1M-quad
    graded quarter ring (linear hardening, 30 increments to 1.2 sigma_Y
    ln(b/a) internal pressure): mixed and FP64 multigrid ramps, their
    difference and the finest-level smoothing step; one JSON line."""
    import math
    E = 2.1e5
    ctx = A.Context(0)
    if args.workload == "c3":
        nt, nr, nsteps, tol = 2000, 1000, 40, 1e-6
        mesh = A.gen_quarter_plate(1.0, 0.1, nt, nr)
        loads = A.Loads(mesh).force(1, 300.0, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
        mat = A.material(E, 0.3, 240.0, A.LAW_SWIFT, n=0.2)
        name = (f"plate with hole, quarter model {nt}x{nr} quads ({mesh.ne}), Swift hardening sigma_Y 240 MPa n 0.2, "
                f"{nsteps} increments to 300 MPa, Aitken to {tol:g}")
    else:
        # the whole ring flows on the hardening slope past the limit
        # pressure: a stress difference of the exit tolerance shows up about
        # a hundredfold in u, so the secant loop runs to 1e-8 here
        nt, nr, nsteps, tol = 16700, 60, 30, 1e-8
        mesh = A.gen_quarter_ring(0.5, 1.0, nt, nr, 1.05)
        p = 1.2 * 240.0 * math.log(2.0)
        loads = A.Loads(mesh).force(0, -p, 0.0, norm_orient=True).fix_vert_axis(0.0).fix_hor_axis(0.0)
        mat = A.material(E, 0.3, 240.0, A.LAW_LINEAR, K_T=E / 101)
        name = (f"thick ring a 0.5 b 1.0, quarter model {nt}x{nr} quads ({mesh.ne}) graded 1.05 per element radially "
                f"(overall 1.05^59 = {1.05 ** 59:.1f}), linear hardening H E/100, {nsteps} increments to "
                f"1.2 sigma_Y ln(b/a) = {p:.1f} MPa, Aitken to {tol:g}")
    log(name)
    grid = (nt + 1, nr + 1)
    out = {}
    for tag, prec in (("mixed", A.MIXED), ("fp64", A.FP64)):
        P = A.Problem(ctx, mesh, loads, mat)
        P.set_grid(*grid)
        opt = A.options(cg_tol=args.cg_tol, tol=tol, aitken=True, max_outer=500, precision=prec, precond=2,
                        predictor=1, cg_forcing=args.forcing, inner_tol=args.inner_tol)
        t0 = time.perf_counter()
        per = []
        for n in range(1, nsteps + 1):
            o = A.Options.from_buffer_copy(opt)
            o.warm = 1 if n > 1 else 0
            if n == 1:
                P.set_load(n / nsteps)
            else:
                P.advance_load(n / nsteps, 1)
            per.append(P.solve(o))
            log(f"  {tag} increment {n}/{nsteps}: secant {per[-1].outer_iters} cg {per[-1].cg_iters} "
                f"{per[-1].seconds:.3f} s")
        out[tag] = (time.perf_counter() - t0, per, P)
    (tm, per_m, Pm), (tf, per_f, Pf) = out["mixed"], out["fp64"]
    nnz = scalar_nnz(A, Pm)
    _, ms_s, fmt = bench_mg(A, Pm, A.MIXED, args.reps // 4)
    byt = cheb_step_bytes(nnz, 2 * mesh.nn, 4)
    epsT = mat.sigmaT / mat.E
    line = {"metric": METRIC, "value": nsteps / tm, "unit": "load-steps/s", "n_gpus": 1, "steps": nsteps, "warmup": 0,
            "ms_per_step": tm / nsteps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": name, "nodes": mesh.nn, "elements": mesh.ne, "dof": 2 * mesh.nn,
                       "mg_levels": Pm.mg_levels()},
            "fp64_load_steps_per_s": nsteps / tf, "speedup_vs_fp64": tf / tm,
            "rel_err_vs_fp64": {"u": rel(Pm.field("u"), Pf.field("u")), "sigma_i": rel(Pm.field("sigI"), Pf.field("sigI")),
                                "plastic_flags_equal": bool(np.array_equal(Pm.field("epsI") >= epsT,
                                                                           Pf.field("epsI") >= epsT))},
            "plastic_elements": int((Pm.field("epsI") >= epsT).sum()),
            "iterations": {"mixed": [[s.outer_iters, s.cg_iters] for s in per_m],
                           "fp64": [[s.outer_iters, s.cg_iters] for s in per_f]},
            "roofline": {"bound": "hbm", "achieved": fmt / ms_s / 1e6, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": fmt / ms_s / 1e6 / HBM_PEAK_GBS,
                         "traffic": MEASURED_TRAFFIC.get((nt, nr, smooth_kernel(Pm)[1])),
                         "kernel": smooth_kernel(Pm)[0] + " (finest-level Chebyshev smoothing step)",
                         "bytes_per_launch": fmt, "avg_launch_ms": ms_s, "reference_csr_bytes_per_launch": byt},
            "hbm_copy": hbm_copy(A, ctx),
            "cpu_baseline": None}
    print(json.dumps(line), flush=True)


def q8_workload(A, args):
    """8-node quads: the reference's energyConvTest plate (Main.cpp:105-178),
    genRectWithHole(-3, 3, -2, 2, 0.5, 3k, 2k, order 2), 304L power law,
    traction 170 MPa on x = +-3 in 20 increments, Aitken to 1e-6; multigrid
    CG on the mesh's corner lattice in mixed precision (the line's value) and
    in FP64, and block-Jacobi CG in mixed precision (the path without a
    lattice) beside them; the finest level's smoothing step on the 8-node
    rows, the block-Jacobi CG iteration and the element-block assembly.  One
    JSON line."""
    k, nsteps, tol, S = args.q8_k, NSTEPS, 1e-6, 170.0
    E = 193e3
    mesh = A.gen_rect_with_hole(-3., 3., -2., 2., 0.5, 3 * k, 2 * k, order=2)
    loads = A.Loads(mesh).force(1, S, 0.0).force(3, -S, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    mat = A.material(E, 0.29, 170.0, A.LAW_POWER, A=0.01 * E, m=0.015)
    lat = A.corner_lattice2("hole", -3., 3., -2., 2., 0.5, 3 * k, 2 * k)
    name = (f"plate with hole 6x4, r 0.5, 8-node quads ({mesh.ne}, genRectWithHole {3 * k}x{2 * k} order 2), "
            f"304L power law, {nsteps} increments to {S:g} MPa, Aitken to {tol:g}")
    log(name)
    ctx = A.Context(0)
    out = {}
    for tag, prec, pre in (("mixed", A.MIXED, 2), ("fp64", A.FP64, 2), ("mixed_block_jacobi", A.MIXED, 1)):
        P = A.Problem(ctx, mesh, loads, mat)
        if pre == 2:
            P.set_grid(lat[0], lat[1], lat[2], periodic=True)
        opt = A.options(cg_tol=args.cg_tol, tol=tol, aitken=True, max_outer=500, precision=prec, precond=pre,
                        predictor=1, cg_forcing=args.forcing, inner_tol=args.inner_tol)
        t0 = time.perf_counter()
        per = P.solve_steps(nsteps, 1.0, opt)
        out[tag] = (time.perf_counter() - t0, per, P)
        log(f"  {tag}: {nsteps} increments in {out[tag][0]:.3f} s, secant {sum(s.outer_iters for s in per)} "
            f"cg {sum(s.cg_iters for s in per)}")
    (tm, per_m, Pm), (tf, per_f, Pf), (tb, per_b, Pb) = out["mixed"], out["fp64"], out["mixed_block_jacobi"]
    nnz = scalar_nnz(A, Pm)
    N = 2 * mesh.nn
    mv, ms_s, fmt_s = bench_mg(A, Pm, A.MIXED, args.reps)
    ms_cg, fmt_cg = bench_cg(A, Pb, A.MIXED, args.reps, precond=True)
    import ctypes as C
    asm = {}
    for tag, prec in (("fp64", A.FP64), ("fp32", A.MIXED)):
        t = C.c_double()
        A.check(A.lib().cp_bench_assemble(Pb.ptr, prec, 20, C.byref(t)))
        asm[tag] = {"ms": t.value, "elements_per_s": mesh.ne / (t.value * 1e-3)}
    epsT = mat.sigmaT / mat.E
    line = {"metric": METRIC, "value": nsteps / tm, "unit": "load-steps/s", "n_gpus": 1, "steps": nsteps, "warmup": 0,
            "ms_per_step": tm / nsteps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": name, "nodes": mesh.nn, "elements": mesh.ne, "dof": N, "nnz": nnz,
                       "nodes_per_element": 8,
                       "path": ("mixed: FP32 CG preconditioned by a multigrid V-cycle (level 0 the 8-node mesh, "
                                "level 1 its corner lattice) + FP64 refinement, FP64 return map and Aitken")},
            "fp64_load_steps_per_s": nsteps / tf, "speedup_vs_fp64": tf / tm,
            "block_jacobi_mixed_load_steps_per_s": nsteps / tb, "speedup_vs_block_jacobi": tb / tm,
            "rel_err_vs_fp64": {"u": rel(Pm.field("u"), Pf.field("u")), "sigma_i": rel(Pm.field("sigI"), Pf.field("sigI")),
                                "plastic_flags_equal": bool(np.array_equal(Pm.field("epsI") >= epsT,
                                                                           Pf.field("epsI") >= epsT))},
            "rel_err_block_jacobi_vs_multigrid": {"u": rel(Pb.field("u"), Pm.field("u")),
                                                  "sigma_i": rel(Pb.field("sigI"), Pm.field("sigI"))},
            "plastic_elements": int((Pm.field("epsI") >= epsT).sum()),
            "multigrid": {"levels": Pm.mg_levels(), "vcycle_ms": mv, "smooth_step_ms": ms_s,
                          "vcycle_over_finest_step": mv / ms_s},
            "iterations": {"mixed": [[s.outer_iters, s.cg_iters] for s in per_m],
                           "fp64": [[s.outer_iters, s.cg_iters] for s in per_f],
                           "mixed_block_jacobi": [[s.outer_iters, s.cg_iters] for s in per_b]},
            "assembly": asm,
            "roofline": {"bound": "hbm", "achieved": fmt_s / ms_s / 1e6, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": fmt_s / ms_s / 1e6 / HBM_PEAK_GBS, "traffic": None,
                         "kernel": "cpk::k_mg_smooth<0,1,float> (finest-level Chebyshev smoothing step, 8-node rows)",
                         "bytes_per_launch": fmt_s, "avg_launch_ms": ms_s,
                         "working_set_bytes": fmt_s, "cache_resident": fmt_s < 256e6},
            "block_jacobi_cg_iteration": {"kernel": "cpk::k_cg_iter_fused<0,true,float>", "avg_launch_ms": ms_cg,
                                          "bytes_per_launch": fmt_cg, "gbps": fmt_cg / ms_cg / 1e6,
                                          "reference_csr_bytes_per_launch": cg_iteration_bytes(nnz, N, 4)},
            "cpu_baseline": None}
    print(json.dumps(line), flush=True)


NSTEPS_C3 = 40


def cpu_model():
    """CPU model name of this host (lscpu's "Model name", from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def plate_case(A, nt, nr, law):
    E = 2.1e5
    mesh = A.gen_quarter_plate(1.0, 0.1, nt, nr)
    loads = A.Loads(mesh).force(1, 300.0, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    if law == "swift":
        mat = A.material(E, 0.3, 240.0, A.LAW_SWIFT, n=0.2)
    else:
        mat = A.material(E, 0.3, 240.0, A.LAW_LINEAR, K_T=E / 101)
    return mesh, loads, mat


def ring_case(A, nt=16700, nr=60):
    import math
    E = 2.1e5
    mesh = A.gen_quarter_ring(0.5, 1.0, nt, nr, 1.05)
    p = 1.2 * 240.0 * math.log(2.0)
    loads = A.Loads(mesh).force(0, -p, 0.0, norm_orient=True).fix_vert_axis(0.0).fix_hor_axis(0.0)
    return mesh, loads, A.material(E, 0.3, 240.0, A.LAW_LINEAR, K_T=E / 101), p


def ramp_timed(A, ctx, mesh, loads, mat, grid, nsteps, opt, tag, timed=None, upto=None):
    """Increments 1..upto (default nsteps) of an nsteps ramp on a fresh
    problem; the last `timed` of them (default all) are timed, the ones
    before run untimed.  Every solve ends with its stream synchronised, so
    the clock brackets only the timed increments' device work.
    Returns (timed seconds, per-increment stats, problem)."""
    upto = nsteps if upto is None else upto
    timed = upto if timed is None else timed
    P = A.Problem(ctx, mesh, loads, mat)
    if opt.precond == 2:
        P.set_grid(*grid)
        if not GRAPHS:
            P.set_graphs(False)
    per = []
    t0 = None
    for n in range(1, upto + 1):
        o = A.Options.from_buffer_copy(opt)
        o.warm = 1 if n > 1 else 0
        if n == upto - timed + 1:
            t0 = time.perf_counter()
        if n == 1:
            P.set_load(n / nsteps)
        else:
            P.advance_load(n / nsteps, int(opt.predictor))
        st = P.solve(o)
        per.append(st)
        log(f"  {tag} increment {n}/{nsteps}: secant {st.outer_iters} cg {st.cg_iters} refinements "
            f"{st.refinements} rel_err {st.rel_err:.2e} {st.seconds:.3f} s")
    return time.perf_counter() - t0, per, P


def mg_options(A, args, precision, tol):
    return A.options(cg_tol=args.cg_tol, tol=tol, aitken=True, max_outer=500, precision=precision, precond=2,
                     predictor=1, cg_forcing=args.forcing, inner_tol=args.inner_tol)


def pair_record(A, Pm, Pf, mat):
    epsT = mat.sigmaT / mat.E
    return {"u": rel(Pm.field("u"), Pf.field("u")), "sigma_i": rel(Pm.field("sigI"), Pf.field("sigI")),
            "plastic_flags_equal": bool(np.array_equal(Pm.field("epsI") >= epsT, Pf.field("epsI") >= epsT))}


def sub_config(A, ctx, args, which):
    """Config 2 (632x316 plate, linear hardening, 20 increments) or config 4
    (16700x60 graded ring, 30 increments): the whole ramp in mixed precision
    and in FP64 with the same options, each timed."""
    if which == "c2":
        nt, nr, nsteps, tol = 632, 316, 20, 1e-6
        mesh, loads, mat = plate_case(A, nt, nr, "linear")
        name = (f"plate with hole, quarter model {nt}x{nr} quads ({mesh.ne}), linear hardening sigma_Y 240 MPa "
                f"H E/100, {nsteps} increments to 300 MPa, Aitken to {tol:g}")
    else:
        nt, nr, nsteps, tol = 16700, 60, 30, 1e-8
        mesh, loads, mat, p = ring_case(A, nt, nr)
        name = (f"thick ring a 0.5 b 1.0, quarter model {nt}x{nr} quads ({mesh.ne}) graded 1.05 per element "
                f"radially, linear hardening H E/100, {nsteps} increments to 1.2 sigma_Y ln(b/a) = {p:.1f} MPa, "
                f"Aitken to {tol:g} (past the limit pressure a stress difference of the exit tolerance shows up "
                f"about a hundredfold in u)")
    log(name)
    grid = (nt + 1, nr + 1)
    tm, per_m, Pm = ramp_timed(A, ctx, mesh, loads, mat, grid, nsteps, mg_options(A, args, A.MIXED, tol),
                               which + " mixed")
    tf, per_f, Pf = ramp_timed(A, ctx, mesh, loads, mat, grid, nsteps, mg_options(A, args, A.FP64, tol),
                               which + " fp64")
    rec = {"workload": name, "increments": nsteps, "load_steps_per_s": nsteps / tm,
           "fp64_load_steps_per_s": nsteps / tf, "speedup_vs_fp64": tf / tm,
           "rel_err_vs_fp64": pair_record(A, Pm, Pf, mat),
           "cg_iterations": {"mixed": sum(s.cg_iters for s in per_m), "fp64": sum(s.cg_iters for s in per_f)}}
    mv, ms, _ = bench_mg(A, Pm, A.MIXED, 100)
    rec["vcycle_ms_mixed"] = mv
    rec["finest_smooth_step_ms_mixed"] = ms
    rec["vcycle_over_finest_step"] = mv / ms
    Pm.close()
    Pf.close()
    return rec


def served_sweep(A, ctx, args):
    """A config-5 stream of --serve-n problems (plate or graded ring, 2k-500k
    quads log-uniform, random hole radius, yield and law; --serve-steps
    increments each; mixed precision, multigrid CG) solved alone back to back
    for each result's digest, then served with Poisson arrivals at
    --serve-rates times the back-to-back rate, every served result checked
    bit for bit against its digest.  A sample of the stream is solved again
    with the plain FP64 GPU path (the reference's algorithm) for the error of
    the served results against it."""
    from cudaplasticity_amd import serving
    stream = serving.make_stream(args.serve_n, qmin=2e3, qmax=args.serve_qmax)
    opt = solver_options(A, args, A.MIXED, precond=2, tol=args.serve_tol)
    sample = [k for k, s in enumerate(stream) if s[3]["quads"] <= args.serve_ref_qmax][:args.serve_ref_n]
    t_seq, sums, kept = solve_alone(A, ctx, stream, opt, args.serve_steps, keep=set(sample))
    base = len(stream) / t_seq
    log(f"served stream: {len(stream)} problems back to back in {t_seq:.2f} s ({base:.2f} problems/s)")
    sweep = []
    for load in args.serve_rates:
        rec = serve_once(A, ctx, args, stream, opt, load * base, sums)
        rec["load"] = load
        log("  served at %g x: %s" % (load, json.dumps(rec)))
        sweep.append(rec)
    ref_opt = solver_options(A, args, A.FP64, reference=True, tol=args.serve_ref_tol)
    eu, es = [], []
    for k in sample:
        mesh, loads, mat, desc = stream[k]
        P = A.Problem(ctx, mesh, loads, mat)
        P.solve_steps(args.serve_steps, 1.0, ref_opt)
        eu.append(rel(kept[k][0], P.field("u")))
        es.append(rel(kept[k][1], P.field("sigI")))
        log(f"  plain FP64 path, problem {k} ({desc['kind']}, {desc['quads']} quads): served mixed vs it "
            f"u {eu[-1]:.1e} sigma_i {es[-1]:.1e}")
        P.close()
    quads = [s[3]["quads"] for s in stream]
    return {"problems": len(stream), "load_steps_each": args.serve_steps, "quads_range": [min(quads), max(quads)],
            "secant_tol": args.serve_tol, "max_slots": args.serve_slots,
            "sequential_alone_load_steps_per_s": len(stream) * args.serve_steps / t_seq,
            "sweep": sweep,
            "capacity_vs_back_to_back": max(r["throughput_load_steps_per_s"] for r in sweep) /
            (len(stream) * args.serve_steps / t_seq),
            "all_results_equal_alone_bitwise": all(r["results_equal_alone_bitwise"] for r in sweep),
            "rel_err_vs_plain_fp64_sample": {
                "problems": [int(quads[k]) for k in sample],
                "plain_fp64": (f"unpreconditioned FP64 CG to {args.cg_tol:g} at every secant iteration, secant "
                               f"loop to {args.serve_ref_tol:g}"),
                "max_u": max(eu) if eu else None, "max_sigma_i": max(es) if es else None}}


def cpu_baseline_c3(args, ref_cg_first, cores):
    """The oracle (the CPU restatement of the reference solver, OpenMP) on
    the 2000x1000 plate's first increment (7.5 MPa, elastic: one secant
    iteration), the reference algorithm (unpreconditioned FP64 CG), under a
    wall-time cap: a short run gives the time of the set-up and of one CG
    iteration, then one run fills the cap.  The fraction of the increment
    done is the CG iterations completed over the `ref_cg_first` iterations
    the same algorithm needs on the GPU to finish it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    om = O.gen_quarter_plate(1.0, 0.1, 2000, 1000)
    ol = O.Problem(om).force(1, 300.0 / NSTEPS_C3, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    mat = O.material(2.1e5, 0.3, 240.0, O.LAW_SWIFT, n=0.2)
    t0 = time.perf_counter()
    O.solve(ol, mat, O.options(cg_tol=1e-30, cg_maxit=20, fixed_outer=1))
    t_short = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.solve(ol, mat, O.options(cg_tol=1e-30, cg_maxit=120, fixed_outer=1))
    per_it = max(1e-6, (time.perf_counter() - t0 - t_short) / 100)
    n_it = int(max(200, min(ref_cg_first, (args.cpu_cap - t_short) / per_it)))
    t0 = time.perf_counter()
    _, _, st = O.solve(ol, mat, O.options(cg_tol=args.cg_tol, cg_maxit=n_it, fixed_outer=1))
    secs = time.perf_counter() - t0
    done = min(1.0, st.cg_iters / ref_cg_first)
    nnz = O.lib.or_csr_nnz(om.nn, om.ne, O.ip(om.conn))
    cg_bytes = cg_iteration_bytes(nnz, 2 * om.nn, 8)
    return {"value": done / secs, "unit": "load-steps/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
            "seconds": secs, "increments_completed": int(st.status == 0 and st.cg_iters < n_it),
            "cg_iterations_done": int(st.cg_iters), "cg_iterations_needed": int(ref_cg_first),
            "cg_gbps": st.cg_iters * cg_bytes / max(1e-9, secs - t_short) / 1e9,
            "sample": (f"oracle/cp_oracle.c (OpenMP, {cores} threads) on the 2000x1000 plate, first of the 40 "
                       f"increments (7.5 MPa, elastic), the reference algorithm (unpreconditioned FP64 CG to "
                       f"{args.cg_tol:g}) under a {args.cpu_cap:g} s cap: {st.cg_iters} of the {ref_cg_first} CG "
                       f"iterations the increment needs (the GPU run of the same algorithm finishes it) in "
                       f"{secs:.1f} s, i.e. {done:.4f} of one load step")}


def c3_workload(A, args):
    """The headline line: config 3 (2000x1000 plate, Swift law, 40
    increments)."""
    nt, nr, tol = 2000, 1000, 1e-6
    mesh, loads, mat = plate_case(A, nt, nr, "swift")
    grid = (nt + 1, nr + 1)
    K = max(1, min(args.steps, NSTEPS_C3))
    ctx = A.Context(0)
    log("device", ctx.info())
    log(f"mesh {nt}x{nr}: {mesh.nn} nodes, {mesh.ne} quads")
    opt_mx, opt_64 = mg_options(A, args, A.MIXED, tol), mg_options(A, args, A.FP64, tol)
    if args.warmup > 0:
        _, _, P = ramp_timed(A, ctx, mesh, loads, mat, grid, NSTEPS_C3, opt_mx, "warmup",
                             upto=min(args.warmup, NSTEPS_C3))
        P.close()
    t_mx, per_mx, Pmx = ramp_timed(A, ctx, mesh, loads, mat, grid, NSTEPS_C3, opt_mx, "mixed", timed=K)
    log(f"mixed: last {K} of {NSTEPS_C3} increments in {t_mx:.3f} s")
    t_64, per_64, P64 = ramp_timed(A, ctx, mesh, loads, mat, grid, NSTEPS_C3, opt_64, "fp64", timed=K)
    log(f"fp64: last {K} of {NSTEPS_C3} increments in {t_64:.3f} s")
    err = pair_record(A, Pmx, P64, mat)
    P64.close()

    ref = None
    if args.ref_steps > 0:
        R = min(args.ref_steps, NSTEPS_C3)
        t_ref, per_ref, Pref = ramp_timed(A, ctx, mesh, loads, mat, grid, NSTEPS_C3,
                                          solver_options(A, args, A.FP64, reference=True, tol=tol), "plain-fp64",
                                          upto=R)
        t_mx_R = sum(s.seconds for s in per_mx[:R])
        ref = {"what": ("the plain FP64 GPU path: the reference's algorithm (unpreconditioned FP64 CG to cg_tol at "
                        "every secant iteration, the previous increment's state as initial guess)"),
               "increments": R, "seconds": t_ref, "load_steps_per_s": R / t_ref,
               "mixed_same_increments_seconds": t_mx_R, "speedup_mixed_vs_plain_fp64": t_ref / t_mx_R,
               "iterations": [[s.outer_iters, s.cg_iters] for s in per_ref]}
        Pref.close()

    # dominant kernel of the ramp: the finest level's Chebyshev smoothing step
    reps = args.reps
    mv32, ms_s32, fmt_s32 = bench_mg(A, Pmx, A.MIXED, reps)
    nnz = scalar_nnz(A, Pmx)
    ref_bytes = cheb_step_bytes(nnz, 2 * mesh.nn, 4)
    kname, tkey = smooth_kernel(Pmx)
    traffic = MEASURED_TRAFFIC.get((nt, nr, tkey))
    roof = {"bound": "hbm", "achieved": fmt_s32 / ms_s32 / 1e6, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": fmt_s32 / ms_s32 / 1e6 / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": kname + " (finest-level Chebyshev smoothing step, FP32)",
            "bytes_per_launch": fmt_s32, "avg_launch_ms": ms_s32, "reps": reps,
            "bytes_model": (("what one launch must move in the symmetric lattice layout: the five stored 2x2 blocks "
                             "of every node once (the backward couplings are the neighbours' forward blocks), x, d, "
                             "b read, x, d written, the pivot inverses and the mask once each")
                            if tkey.endswith("_sym") else
                            ("what one launch must move in the 2x2-block layout: every stored block (4 values + "
                             "its column) once, slice descriptors, x, d, b read, x, d written, the pivot inverses "
                             "and the mask once each")),
            "working_set_bytes": fmt_s32,
            "traffic_gbps": traffic / ms_s32 / 1e6 if traffic else None,
            "reference_csr_bytes_per_launch": ref_bytes}
    spmv = spmv_record(A, Pmx, nt, nr, nnz, 2 * mesh.nn, reps)
    log("spmv:", spmv)
    mv64, ms_s64, fmt_s64 = bench_mg(A, Pmx, A.FP64, max(20, reps // 4))
    mg = {"levels": Pmx.mg_levels(),
          "smoothers": ["segments" if k else "block-Jacobi" for k in Pmx.mg_smoothers()[:-1]] + ["exact"],
          "vcycle_ms": {"mixed": mv32, "fp64": mv64}, "smooth_step_ms": {"mixed": ms_s32, "fp64": ms_s64},
          "vcycle_over_finest_step": {"mixed": mv32 / ms_s32, "fp64": mv64 / ms_s64},
          "smooth_step_gbps": {"mixed": fmt_s32 / ms_s32 / 1e6, "fp64": fmt_s64 / ms_s64 / 1e6},
          "cg_iterations": {"mixed": sum(s.cg_iters for s in per_mx[NSTEPS_C3 - K:]),
                            "fp64": sum(s.cg_iters for s in per_64[NSTEPS_C3 - K:])}}
    log("multigrid:", mg)
    times = {"cg_s": sum(s.cg_seconds for s in per_mx[NSTEPS_C3 - K:]),
             "assembly_s": sum(s.asm_seconds for s in per_mx[NSTEPS_C3 - K:]),
             "return_map_s": sum(s.rm_seconds for s in per_mx[NSTEPS_C3 - K:])}
    epsT = mat.sigmaT / mat.E
    plastic = int((Pmx.field("epsI") >= epsT).sum())
    Pmx.close()

    subs = {}
    for which in ("c2", "c4"):
        if which not in args.skip:
            subs[which] = sub_config(A, ctx, args, which)
            log(which, subs[which])
    served = None if "serve" in args.skip or args.serve_n <= 0 else served_sweep(A, ctx, args)
    if served:
        log("served:", served)
    copy = hbm_copy(A, ctx)
    ctx.close()

    prefix = "" if K == NSTEPS_C3 else (f"; timed: increments {NSTEPS_C3 - K + 1}..{NSTEPS_C3}, after "
                                        f"{NSTEPS_C3 - K} untimed ones")
    line = {
        "metric": METRIC, "value": K / t_mx, "unit": "load-steps/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": t_mx / K * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": (f"plate with hole, quarter model {nt}x{nr} quads ({mesh.ne}), Swift hardening "
                                f"sigma_Y 240 MPa n 0.2 eps0 sigma_Y/E, {NSTEPS_C3} increments to 300 MPa, Aitken "
                                f"to {tol:g}, cg_tol {args.cg_tol:g}{prefix}"),
                   "nodes": mesh.nn, "elements": mesh.ne, "dof": 2 * mesh.nn,
                   "path": ("mixed: FP32 CG preconditioned by a geometric multigrid V-cycle + FP64 refinement, "
                            f"FP64 return map and Aitken, load predictor, inexact secant iterations (forcing "
                            f"{args.forcing:g})")},
        "fp64_load_steps_per_s": K / t_64, "speedup_vs_fp64": t_64 / t_mx, "rel_err_vs_fp64": err,
        "plastic_elements": plastic,
        "plain_fp64_path": ref,
        "iterations": {"mixed": [[s.outer_iters, s.cg_iters, s.refinements] for s in per_mx],
                       "fp64": [[s.outer_iters, s.cg_iters] for s in per_64]},
        "device_time_mixed": times,
        "roofline": roof, "spmv": spmv, "multigrid": mg, "hbm_copy": copy,
        "c2": subs.get("c2"), "c4": subs.get("c4"), "served": served,
        "cpu_baseline": None,
    }
    if "cpu" not in args.skip and ref:
        cores = len(os.sched_getaffinity(0))
        os.environ.setdefault("OMP_NUM_THREADS", str(cores))
        line["cpu_baseline"] = cpu_baseline_c3(args, ref["iterations"][0][1], int(os.environ["OMP_NUM_THREADS"]))
        log("cpu baseline:", line["cpu_baseline"])
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=("c2", "c3", "c4", "c5", "q8"),
                    help="c3 (default): the headline line, 2000x1000 plate, Swift law, 40 increments, with the "
                         "632x316 plate and the graded ring as sub-records and a served sweep; c2: the 632x316 "
                         "plate alone with its kernel micro-benchmarks; c4: 16700x60 graded quarter ring, 30 "
                         "increments; c5: served stream of --c5-n problems at swept arrival rates; q8: 8-node "
                         "quads on the reference's energyConvTest plate")
    ap.add_argument("--skip", nargs="*", default=[], choices=("c2", "c4", "serve", "cpu"),
                    help="default workload: parts of the line to leave out")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[1.0, 2.0, 4.0],
                    help="default workload: served arrival rates, multiples of the back-to-back rate")
    ap.add_argument("--serve-ref-n", type=int, default=4,
                    help="default workload: served problems solved again with the plain FP64 GPU path")
    ap.add_argument("--serve-ref-qmax", type=float, default=5e4,
                    help="default workload: largest problem (quads) of that sample")
    ap.add_argument("--serve-ref-tol", type=float, default=1e-9,
                    help="default workload: secant tolerance of the plain FP64 runs of the sample")
    ap.add_argument("--cpu-cap", type=float, default=20.0,
                    help="default workload: wall-time cap (s) of the CPU oracle's sample")
    ap.add_argument("--q8-k", type=int, default=64,
                    help="q8 mesh: genRectWithHole(-3, 3, -2, 2, 0.5, 3k, 2k, order 2)")
    ap.add_argument("--steps", type=int, default=NSTEPS_C3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nt", type=int, default=632)
    ap.add_argument("--nr", type=int, default=316)
    ap.add_argument("--cg-tol", type=float, default=1e-9)
    # inexact secant iterations: each solve to max(cg_tol, forcing x the last
    # secant residual); 0.3 is the fastest of 0.05-0.7 for the multigrid path
    # on the 632x316 and 2000x1000 plates (the last solve of a level is exact)
    ap.add_argument("--forcing", type=float, default=0.3)
    ap.add_argument("--precond", type=int, default=2, choices=(1, 2),
                    help="1 block-Jacobi CG, 2 multigrid-preconditioned CG (headline)")
    ap.add_argument("--inner-tol", type=float, default=1e-5)
    ap.add_argument("--ref-steps", type=int, default=2,
                    help="increments of the plain FP64 GPU path (the reference's algorithm) to time")
    ap.add_argument("--large-nt", type=int, default=2000)
    ap.add_argument("--large-nr", type=int, default=1000)
    ap.add_argument("--serve-n", type=int, default=64)
    ap.add_argument("--serve-qmax", type=float, default=5e5)
    ap.add_argument("--serve-steps", type=int, default=5)
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--serve-load", type=float, default=1.0)
    ap.add_argument("--serve-tol", type=float, default=1e-7,
                    help="secant exit tolerance of the served streams: problems loaded near or past their limit "
                         "load need 1e-7 for u within 1e-6 of FP64")
    ap.add_argument("--c5-n", type=int, default=512, help="--workload c5: problems in the stream")
    ap.add_argument("--c5-ref-n", type=int, default=8,
                    help="--workload c5: problems solved again with the plain FP64 GPU path")
    ap.add_argument("--c5-ref-qmax", type=float, default=1e5,
                    help="--workload c5: largest problem (quads) of that sample")
    ap.add_argument("--serve-loads", type=float, nargs="+", default=[0.5, 1.0, 2.0, 4.0],
                    help="--workload c5: arrival rates swept, as multiples of the back-to-back rate")
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--reps", type=int, default=400, help="launches of the CG iteration kernel for the roofline")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="MiB: only run a plain device copy of this size 10 times (HBM counter calibration)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="multigrid CG without HIP graph replay (same kernels; for rocprofv3 kernel traces)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only time the CG iteration kernel on the workload's matrix (for profiler runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bj", action="store_true", help="skip the block-Jacobi comparison ramp")
    ap.add_argument("--no-ref", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    args = ap.parse_args()
    global GRAPHS
    GRAPHS = not args.no_graphs
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: each problem runs on a single MI355X here; how a problem or a "
              "stream would shard over several GPUs is described in DESIGN.md, not run.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c3" and args.copy_calibration <= 0 and not args.roofline_only:
        from cudaplasticity_amd import abi as A
        c3_workload(A, args)
        return
    K = max(1, min(args.steps, NSTEPS))
    if args.workload == "q8":
        import cudaplasticity_amd.abi as A
        q8_workload(A, args)
        return
    if args.workload == "c5":
        from cudaplasticity_amd import abi as A
        c5_workload(A, args)
        return
    if args.workload == "c4":
        from cudaplasticity_amd import abi as A
        other_workload(A, args)
        return
    if args.workload == "c3":   # --roofline-only / --copy-calibration on the 2000x1000 plate
        args.nt, args.nr = 2000, 1000
    if args.copy_calibration > 0:
        import torch
        n = int(args.copy_calibration * (1 << 20)) // 4
        src = torch.rand(n, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}), flush=True)
        return
    from cudaplasticity_amd import abi as A

    mesh, loads, mat = build_case(A, args.nt, args.nr)
    log(f"mesh {args.nt}x{args.nr}: {mesh.nn} nodes, {mesh.ne} quads")
    ctx = A.Context(0)
    log("device", ctx.info())
    grid = (args.nt + 1, args.nr + 1)
    if args.roofline_only:
        # the dominant kernel alone, for profiling, on the workload's matrix
        # (initial secant moduli, full load): the finest level's Chebyshev
        # smoothing step (multigrid) or the block-Jacobi CG iteration
        P = A.Problem(ctx, mesh, loads, mat)
        P.set_load(1.0)
        N, nnz = 2 * mesh.nn, scalar_nnz(A, P)
        spmv = None
        if args.precond == 2:
            P.set_grid(*grid)
            _, ms, fmt = bench_mg(A, P, A.MIXED, args.reps)
            name, byt = smooth_kernel(P)[0], cheb_step_bytes(nnz, N, 4)
            # the CG's operator product after it, in the same run (counter runs cover both)
            spmv = spmv_record(A, P, args.nt, args.nr, nnz, N, args.reps)
        else:
            ms, fmt = bench_cg(A, P, A.MIXED, args.reps)
            name, byt = "cpk::k_cg_iter_fused<9,true,float>", cg_iteration_bytes(nnz, N, 4)
        print(json.dumps({"kernel": name, "launches": args.reps, "avg_launch_ms": ms, "bytes_per_launch": byt,
                          "achieved_gbps": byt / ms / 1e6, "format_bytes_per_launch": fmt,
                          "achieved_format_gbps": fmt / ms / 1e6, "spmv": spmv}), flush=True)
        P.close()
        return
    opt_mx = solver_options(A, args, A.MIXED)
    opt_64 = solver_options(A, args, A.FP64)

    if args.warmup > 0:
        t, _, P = run_ramp(A, ctx, mesh, loads, mat, min(args.warmup, NSTEPS), opt_mx, "warmup", grid)
        P.close()

    t_mx, per_mx, Pmx = run_ramp(A, ctx, mesh, loads, mat, K, opt_mx, "mixed", grid)
    log(f"mixed: {K} increments in {t_mx:.3f} s")
    t_64, per_64, P64 = run_ramp(A, ctx, mesh, loads, mat, K, opt_64, "fp64", grid)
    log(f"fp64: {K} increments in {t_64:.3f} s")
    bj = None
    if args.precond == 2 and not args.no_bj:
        t_bj, per_bj, Pbj = run_ramp(A, ctx, mesh, loads, mat, K, solver_options(A, args, A.MIXED, precond=1),
                                     "mixed-block-jacobi")
        bj = {"load_steps_per_s": K / t_bj, "cg_iterations": sum(s.cg_iters for s in per_bj),
              "rel_err_vs_multigrid": {"u": rel(Pbj.field("u"), Pmx.field("u")),
                                       "sigma_i": rel(Pbj.field("sigI"), Pmx.field("sigI"))}}
        Pbj.close()
    epsT = mat.sigmaT / mat.E
    err = {"u": rel(Pmx.field("u"), P64.field("u")), "sigma_i": rel(Pmx.field("sigI"), P64.field("sigI")),
           "plastic_flags_equal": bool(np.array_equal(Pmx.field("epsI") >= epsT, P64.field("epsI") >= epsT))}

    ref = None
    per_ref_steps = None
    if not args.no_ref:
        R = max(1, min(args.ref_steps, K))
        t_ref, per_ref, Pref = run_ramp(A, ctx, mesh, loads, mat, R, solver_options(A, args, A.FP64, True), "fp64-ref")
        Pchk = A.Problem(ctx, mesh, loads, mat)
        if opt_mx.precond == 2:
            Pchk.set_grid(*grid)
        Pchk.solve_steps(R, R / NSTEPS, opt_mx)
        t_mx_R = sum(s.seconds for s in per_mx[:R])
        per_ref_steps = per_ref
        ref = {"increments": R, "seconds": t_ref, "load_steps_per_s": R / t_ref,
               "mixed_same_increments_seconds": t_mx_R, "speedup_mixed_vs_reference_fp64": t_ref / t_mx_R,
               "rel_err_mixed_vs_reference": {"u": rel(Pchk.field("u"), Pref.field("u")),
                                              "sigma_i": rel(Pchk.field("sigI"), Pref.field("sigI"))},
               "iterations": [[s.outer_iters, s.cg_iters] for s in per_ref]}
        Pref.close()
        Pchk.close()

    # dominant kernel: the fused CG iteration of the mixed path (FP32 block-Jacobi PCG)
    reps = args.reps
    fused = (mesh.nn + 255) // 256 <= 2048   # the library folds the scalar step in up to 2048 blocks
    ms32, fmt32 = bench_cg(A, Pmx, A.MIXED, reps)
    ms64, fmt64 = bench_cg(A, P64, A.FP64, reps)
    sp_ms, sp_stored, sp_algo = P64.bench_spmv(A.FP64, reps)
    nnz = scalar_nnz(A, Pmx)
    bytes32, bytes64 = cg_iteration_bytes(nnz, 2 * mesh.nn, 4), cg_iteration_bytes(nnz, 2 * mesh.nn, 8)
    ach32 = bytes32 / (ms32 * 1e-3) / 1e9
    ach64 = bytes64 / (ms64 * 1e-3) / 1e9
    cg_it = sum(s.cg_iters for s in per_mx)
    cg_s = sum(s.cg_seconds for s in per_mx)
    mg = None
    if args.precond == 2:
        # multigrid path: the finest level's Chebyshev smoothing step (three
        # per V-cycle each way) and the whole V-cycle
        mv32, ms_s32, fmt_s32 = bench_mg(A, Pmx, A.MIXED, reps)
        mv64, ms_s64, fmt_s64 = bench_mg(A, P64, A.FP64, reps)
        bs32, bs64 = cheb_step_bytes(nnz, 2 * mesh.nn, 4), cheb_step_bytes(nnz, 2 * mesh.nn, 8)
        mg = {"levels": Pmx.mg_levels(),
              "smoothers": ["segments" if k else "block-Jacobi" for k in Pmx.mg_smoothers()[:-1]] + ["exact"],
              "vcycle_ms": {"mixed": mv32, "fp64": mv64},
              "smooth_step_ms": {"mixed": ms_s32, "fp64": ms_s64},
              "smooth_step_gbps": {"mixed": bs32 / ms_s32 / 1e6, "fp64": bs64 / ms_s64 / 1e6},
              "smooth_step_format_gbps": {"mixed": fmt_s32 / ms_s32 / 1e6, "fp64": fmt_s64 / ms_s64 / 1e6},
              "cg_iterations": {"mixed": sum(s.cg_iters for s in per_mx), "fp64": sum(s.cg_iters for s in per_64)}}
        log("multigrid:", mg)
    log(f"CG iteration kernel: FP32 {ms32 * 1e3:.1f} us {bytes32 / 1e6:.1f} MB ({fmt32 / 1e6:.1f} MB as stored) "
        f"-> {ach32:.0f} GB/s; "
        f"FP64 {ms64 * 1e3:.1f} us {bytes64 / 1e6:.1f} MB -> {ach64:.0f} GB/s; SpMV FP64 {sp_ms * 1e3:.1f} us "
        f"-> {sp_algo / sp_ms / 1e6:.0f} GB/s; layout {Pmx.layout()}")

    large = None
    if args.large_nt > 0:
        # the same kernels on the 2000x1000 mesh, whose matrix and vectors
        # (about 0.8 GB in FP32) do not fit the 256 MB last-level cache
        lm, ll, lmat = build_case(A, args.large_nt, args.large_nr)
        PL = A.Problem(ctx, lm, ll, lmat)
        PL.set_load(1.0)
        lnnz = scalar_nnz(A, PL)
        l32, lf32 = bench_cg(A, PL, A.MIXED, max(20, reps // 4))
        l64, lf64 = bench_cg(A, PL, A.FP64, max(20, reps // 4))
        lsp_ms, _, lsp_algo = PL.bench_spmv(A.FP64, max(20, reps // 4))
        PL.set_grid(args.large_nt + 1, args.large_nr + 1)
        lv32, lsm32, lfs32 = bench_mg(A, PL, A.MIXED, max(20, reps // 4))
        lbs32 = cheb_step_bytes(lnnz, 2 * lm.nn, 4)
        b32, b64 = cg_iteration_bytes(lnnz, 2 * lm.nn, 4), cg_iteration_bytes(lnnz, 2 * lm.nn, 8)
        large = {"mesh": f"{args.large_nt}x{args.large_nr}", "elements": lm.ne, "nodes": lm.nn,
                 "mixed_iteration_kernel": {"ms": l32, "bytes": b32, "gbps": b32 / l32 / 1e6,
                                            "format_gbps": lf32 / l32 / 1e6},
                 "fp64_iteration_kernel": {"ms": l64, "bytes": b64, "gbps": b64 / l64 / 1e6,
                                           "format_gbps": lf64 / l64 / 1e6},
                 "fp64_spmv": {"ms": lsp_ms, "bytes": lsp_algo, "gbps": lsp_algo / lsp_ms / 1e6},
                 "mixed_mg_smooth_step": {"ms": lsm32, "bytes": lbs32, "gbps": lbs32 / lsm32 / 1e6,
                                          "format_gbps": lfs32 / lsm32 / 1e6},
                 "mixed_mg_vcycle_ms": lv32, "mg_levels": PL.mg_levels()}
        log("large mesh kernels:", large)
        PL.close()

    serve = serve_bj = None
    if not args.no_serve and args.serve_n > 0:
        serve = served_stream(A, ctx, args, precond=2)
        log("served:", serve)
        if not args.no_bj:
            # the block-Jacobi serving path, batched across requests, on a
            # smaller stream (2k-50k quads)
            serve_bj = served_stream(A, ctx, args, precond=1, n=24, qmax=5e4)
            log("served, block-Jacobi:", serve_bj)

    line = {
        "metric": METRIC,
        "value": K / t_mx,
        "unit": "load-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": t_mx / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"plate with hole, quarter model {args.nt}x{args.nr} quads ({mesh.ne}), linear "
                                f"hardening sigma_Y 240 MPa H E/100, increments 1..{K} of {NSTEPS} to 300 MPa, "
                                f"Aitken to 1e-6, cg_tol {args.cg_tol:g}"),
                   "nodes": mesh.nn, "elements": mesh.ne, "dof": 2 * mesh.nn,
                   "path": (("mixed: FP32 CG preconditioned by a geometric multigrid V-cycle" if args.precond == 2
                             else "mixed: FP32 block-Jacobi CG") + " + FP64 refinement, FP64 return map, load "
                            f"predictor, inexact secant iterations (forcing {args.forcing:g})")},
        "problem_setup_seconds": Pmx.setup_seconds,
        "load_steps_per_s_incl_setup_and_upload": K / (t_mx + Pmx.setup_seconds),
        "fp64_load_steps_per_s": K / t_64,
        "speedup_vs_fp64": t_64 / t_mx,
        "rel_err_vs_fp64": err,
        "fp64_reference_algorithm": ref,
        # block-Jacobi CG iteration kernels (the multigrid path's kernels are
        # reported under "multigrid"); the in-solve figure only for block-Jacobi
        "cg_gbps": {"mixed_iteration_kernel": ach32, "fp64_iteration_kernel": ach64,
                    "fp64_spmv": sp_algo / (sp_ms * 1e-3) / 1e9,
                    "mixed_in_solve_incl_scalar_step": (cg_it * bytes32 / cg_s / 1e9)
                    if cg_s > 0 and args.precond != 2 else None},
        "iterations": {"mixed": [[s.outer_iters, s.cg_iters, s.refinements] for s in per_mx],
                       "fp64": [[s.outer_iters, s.cg_iters] for s in per_64]},
        # bytes each launch must move in this layout; config 2's working set
        # (about 48 MB) sits inside the 256 MB Infinity Cache, so these rates
        # are cache-resident ones
        "roofline": ({"bound": "hbm", "achieved": fmt_s32 / ms_s32 / 1e6, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": fmt_s32 / ms_s32 / 1e6 / HBM_PEAK_GBS,
                      "traffic": MEASURED_TRAFFIC.get((args.nt, args.nr, smooth_kernel(Pmx)[1])),
                      "kernel": smooth_kernel(Pmx)[0] + " (finest-level Chebyshev smoothing step)",
                      "bytes_per_launch": fmt_s32, "avg_launch_ms": ms_s32, "reps": reps,
                      "working_set_bytes": fmt_s32, "cache_resident": fmt_s32 < 256e6,
                      "reference_csr_bytes_per_launch": bs32}
                     if mg else
                     {"bound": "hbm", "achieved": fmt32 / (ms32 * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": fmt32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "traffic": MEASURED_TRAFFIC.get((args.nt, args.nr, "cg")),
                      "kernel": ("cpk::k_cg_iter_fused<9,true,float> (block-Jacobi Chronopoulos-Gear iteration, "
                                 "scalar step included)" if fused else "cpk::k_cg_iter<9,true,float>"),
                      "bytes_per_launch": fmt32, "avg_launch_ms": ms32, "reps": reps,
                      "working_set_bytes": fmt32, "cache_resident": fmt32 < 256e6,
                      "reference_csr_bytes_per_launch": bytes32}),
        "multigrid": mg,
        "mixed_block_jacobi": bj,
        "kernels_2000x1000": large,
        "hbm_copy": hbm_copy(A, ctx),
        "served": serve,
        "served_block_jacobi": serve_bj,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))
        os.environ.setdefault("OMP_NUM_THREADS", str(cores))
        threads = int(os.environ["OMP_NUM_THREADS"])
        if per_ref_steps:
            line["cpu_baseline"] = cpu_baseline(args.nt, args.nr, per_ref_steps, threads)
        else:
            line["cpu_baseline"] = cpu_baseline(args.nt, args.nr, per_64, threads, "the FP64 path")
        line["cpu_baseline"]["gpu_same_increments"] = {
            "fp64_reference_algorithm_load_steps_per_s": ref["load_steps_per_s"] if ref else None,
            "mixed_load_steps_per_s": (ref["increments"] / ref["mixed_same_increments_seconds"]) if ref else None}
    Pmx.close()
    P64.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
