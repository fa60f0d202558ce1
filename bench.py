"""Benchmark of the elasto-plastic load-stepping loop on one MI355X.

Workload (config 2): quarter plate [0,1]^2 with a hole of radius 0.1,
N_theta x N_r = 632 x 316 bilinear quads (199 712 elements, 200 661 nodes,
401 322 DOF), plane stress, E = 2.1e5 MPa, nu = 0.3, von Mises secant
(deformation theory) update with linear hardening sigma_Y = 240 MPa,
K_T = E/101 (i.e. H = E/100), 20 equal load increments of the far-edge
traction up to 300 MPa, Aitken-accelerated secant iterations to a 1e-6
stress tolerance, CG to ||r||/||b|| = cg_tol.

A "step" is one load increment: warm-started from the previous increment,
assemble -> CG -> return map until converged.  The timed region runs the
first K increments of the 20-increment ramp on a fresh device problem with
all inputs already in HBM; W warm-up increments run first on a separate
instance.  Both the plain FP64 path and the mixed path (FP32 CG + FP64
refinement, FP64 return map) are timed; `value` is the mixed path and the
line carries the FP64 rate and the error between the two.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "load-steps/s + CG GB/s at 1e-6 rel err; p50/p99 latency (mixed+served vs FP64)"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_case(A, Nt, Nr, S=300.0):
    mesh = A.gen_quarter_plate(1.0, 0.1, Nt, Nr)
    loads = A.Loads(mesh).force(1, S, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    E = 2.1e5
    mat = A.material(E, 0.3, 240.0, A.LAW_LINEAR, K_T=E / 101)
    return mesh, loads, mat


def run_ramp(A, ctx, mesh, loads, mat, nsteps_total, ksteps, opt):
    """First `ksteps` increments of an `nsteps_total` ramp; returns (seconds, stats list, problem)."""
    P = A.Problem(ctx, mesh, loads, mat)
    per = []
    t0 = time.perf_counter()
    for n in range(1, ksteps + 1):
        o = A.Options.from_buffer_copy(opt)
        o.warm = 1 if n > 1 else 0
        P.set_load(n / nsteps_total)
        st = P.solve(o)
        per.append(st)
        log(f"  increment {n}/{nsteps_total}: outer {st.outer_iters} cg {st.cg_iters} refinements {st.refinements} "
            f"rel_err {st.rel_err:.2e} {st.seconds:.3f} s (cg {st.cg_seconds:.3f} asm {st.asm_seconds:.3f} "
            f"rm {st.rm_seconds:.3f})")
    dt = time.perf_counter() - t0
    return dt, per, P


def cpu_baseline(mesh_args, gpu_per_step, cores):
    """Oracle (CPU FP64 restatement of the reference solver) timed on the
    host: per-outer-iteration cost (pattern, B, assembly, return map) and
    per-CG-iteration cost on the same mesh, from two bounded runs of the
    first increment with 60 and 160 CG iterations; the load-steps/s figure
    applies them to the iteration counts of the FP64 GPU run."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    Nt, Nr = mesh_args
    om = O.gen_quarter_plate(1.0, 0.1, Nt, Nr)
    ol = O.Problem(om).force(1, 300.0 / 20, 0.0).fix_vert_axis(0.0).fix_hor_axis(0.0)
    E = 2.1e5
    mat = O.material(E, 0.3, 240.0, O.LAW_LINEAR, K_T=E / 101)
    t = {}
    for it in (60, 160):
        t0 = time.perf_counter()
        O.solve(ol, mat, O.options(cg_tol=1e-30, cg_maxit=it, fixed_outer=1))
        t[it] = time.perf_counter() - t0
    per_it = (t[160] - t[60]) / 100.0
    per_outer = t[60] - 60 * per_it
    outer = sum(s.outer_iters for s in gpu_per_step)
    cg = sum(s.cg_iters for s in gpu_per_step)
    secs = outer * per_outer + cg * per_it
    return {"value": len(gpu_per_step) / secs, "unit": "load-steps/s", "cores": cores, "kind": "port",
            "sample": (f"oracle/cp_oracle.c on the same {Nt}x{Nr} mesh: first increment with 60 and 160 CG "
                       f"iterations timed ({t[60] + t[160]:.1f} s total); {per_it * 1e3:.2f} ms per CG iteration, "
                       f"{per_outer:.2f} s per secant iteration; applied to the FP64 GPU run's {outer} secant "
                       f"and {cg} CG iterations")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nt", type=int, default=632)
    ap.add_argument("--nr", type=int, default=316)
    ap.add_argument("--cg-tol", type=float, default=1e-9)
    ap.add_argument("--inner-tol", type=float, default=1e-3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build runs each problem on a single MI355X "
              "(multi-GPU sharding is described in DESIGN.md, not run).", file=sys.stderr)
        sys.exit(2)
    nsteps_total = 20
    K = max(1, min(args.steps, nsteps_total))
    from cudaplasticity_amd import abi as A

    mesh, loads, mat = build_case(A, args.nt, args.nr)
    log(f"mesh {args.nt}x{args.nr}: {mesh.nn} nodes, {mesh.ne} quads")
    ctx = A.Context(0)
    info = ctx.info()
    log("device", info)
    common = dict(cg_tol=args.cg_tol, tol=1e-6, aitken=True, max_outer=500)
    opt64 = A.options(precision=A.FP64, **common)
    optmx = A.options(precision=A.MIXED, inner_tol=args.inner_tol, **common)

    if args.warmup > 0:
        t, per, P = run_ramp(A, ctx, mesh, loads, mat, nsteps_total, min(args.warmup, nsteps_total), optmx)
        P.close()
        log(f"warmup {args.warmup} increments: {t:.2f} s")

    t_mx, per_mx, Pmx = run_ramp(A, ctx, mesh, loads, mat, nsteps_total, K, optmx)
    log(f"mixed: {K} increments in {t_mx:.3f} s; per step (outer, cg, refinements):",
        [(s.outer_iters, s.cg_iters, s.refinements) for s in per_mx])
    res = {"mixed_seconds": t_mx}
    if not args.no_fp64:
        t_64, per_64, P64 = run_ramp(A, ctx, mesh, loads, mat, nsteps_total, K, opt64)
        log(f"fp64: {K} increments in {t_64:.3f} s; per step (outer, cg):",
            [(s.outer_iters, s.cg_iters) for s in per_64])
        u64, umx = P64.field("u"), Pmx.field("u")
        s64, smx = P64.field("sigI"), Pmx.field("sigI")
        err_u = float(np.linalg.norm(umx - u64) / np.linalg.norm(u64))
        err_s = float(np.linalg.norm(smx - s64) / np.linalg.norm(s64))
        epsT = mat.sigmaT / mat.E
        flags_equal = bool(np.array_equal(P64.field("epsI") >= epsT, Pmx.field("epsI") >= epsT))
    else:
        t_64, per_64, P64, err_u, err_s, flags_equal = None, None, None, None, None, None

    # roofline of the dominant kernel: the fused CG iteration (FP32 in the mixed path)
    reps = 200
    ms32, bytes32 = (lambda r: (r[0], r[1]))(bench_cg(A, Pmx, A.MIXED, reps))
    ms64, bytes64 = bench_cg(A, P64 or Pmx, A.FP64, reps)
    ach32 = bytes32 / (ms32 * 1e-3) / 1e9
    ach64 = bytes64 / (ms64 * 1e-3) / 1e9
    cg_it_mx = sum(s.cg_iters for s in per_mx)
    cg_s_mx = sum(s.cg_seconds for s in per_mx)
    lay = Pmx.layout()
    log(f"CG iteration kernel: FP32 {ms32 * 1e3:.1f} us, {bytes32 / 1e6:.1f} MB -> {ach32:.0f} GB/s; "
        f"FP64 {ms64 * 1e3:.1f} us, {bytes64 / 1e6:.1f} MB -> {ach64:.0f} GB/s; layout {lay}")

    value = K / t_mx
    line = {
        "metric": METRIC,
        "value": value,
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
        "config": {"workload": f"plate-with-hole quarter model {args.nt}x{args.nr} quads ({mesh.ne}), linear "
                               f"hardening, {K} of {nsteps_total} increments to 300 MPa, Aitken, tol 1e-6, "
                               f"cg_tol {args.cg_tol:g}",
                   "nodes": mesh.nn, "elements": mesh.ne, "dof": 2 * mesh.nn, "path": "mixed (FP32 CG + FP64 refinement)"},
        "fp64_load_steps_per_s": (K / t_64) if t_64 else None,
        "speedup_vs_fp64": (t_64 / t_mx) if t_64 else None,
        "rel_err_vs_fp64": {"u": err_u, "sigma_i": err_s, "plastic_flags_equal": flags_equal},
        "cg_gbps": {"mixed_iteration_kernel": ach32, "fp64_iteration_kernel": ach64,
                    "mixed_in_solve": (cg_it_mx * bytes32 / cg_s_mx / 1e9) if cg_s_mx > 0 else None},
        "iterations": {"mixed": [[s.outer_iters, s.cg_iters, s.refinements] for s in per_mx],
                       "fp64": [[s.outer_iters, s.cg_iters] for s in per_64] if per_64 else None},
        "roofline": {"bound": "hbm", "achieved": ach32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach32 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_cg_iter<float> (fused Chronopoulos-Gear CG iteration)",
                     "bytes_per_launch": bytes32, "avg_launch_ms": ms32},
        "cpu_baseline": None,
    }
    if not args.no_cpu and per_64:
        cores = len(os.sched_getaffinity(0))
        os.environ.setdefault("OMP_NUM_THREADS", str(cores))
        line["cpu_baseline"] = cpu_baseline((args.nt, args.nr), per_64, int(os.environ["OMP_NUM_THREADS"]))
    print(json.dumps(line), flush=True)


def bench_cg(A, P, precision, reps):
    import ctypes as C
    ms, b = C.c_double(), C.c_double()
    A.check(A.lib().cp_bench_cg(P.ptr, precision, reps, C.byref(ms), C.byref(b)))
    return ms.value, b.value


if __name__ == "__main__":
    main()
