#!/usr/bin/env python3
"""Benchmark: AMG-preconditioned CG to a relative residual of 1e-8 on a 256^3
lognormal (sigma = 1) variable-coefficient diffusion stencil, mixed precision
(FP32 hierarchy / smoothers / preconditioner SpMV inside an FP64 CG) and plain
FP64, on one MI355X; plus the HBM roofline of the hot kernels.

A step is one complete solve from x = 0 (the hierarchy is set up once,
outside the timed region; the setup time is reported separately).  value is
the mixed-precision time-to-solution in seconds (lower is better).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the finest-level f32 SOR sweeps at 256^3: 2 x
# FETCH_SIZE + WRITE_SIZE of rocprofv3 --pmc runs (the factor 2 is what a plain
# copy kernel of known size gives), profiles/stencil_kernels_256_hbm_counters.md
TRAFFIC_256 = {"stencil_sor_fwd": 826_077_768, "stencil_sor_bwd": 1_000_488_656}


def sor_cache_bytes(n: int, esize: int) -> int:
    """Bytes of the stencil SOR sweeps' coefficient cache for an n^3 box (each
    sweep reads it once): per 8 x 8 column and chunk of 4 steps, a 4-value
    record per lane and step plus 16 halo entries of 4 values."""
    cx = (n + 7) // 8
    nsteps = n + 14
    nc = (nsteps + (4 - nsteps % 4) % 4) // 4
    return cx * cx * nc * (4 * 64 * 4 + 16 * 4) * esize


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_ms(ctx, fn, launches):
    """Average duration of one launch in steady state, HIP events on the
    library's stream.  An untimed pass of the same launches comes first: the
    first launches after other work run up to 20% slower (measured on the
    256^3 SpMV: 0.114 ms in the first 60 launches, 0.094 ms from the third
    pass of 60 on)."""
    for k in range(launches):
        fn(k)
    ctx.synchronize()
    ctx.timer_start()
    for k in range(launches):
        fn(k)
    return ctx.timer_stop() / launches


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """The host threads this process may use (the box's CPU share)."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(n, sigma, runs, x_gpu, x_gpu_fp64):
    """The oracle's FP64 PCG-AMG at the benchmark's size on the host's
    threads (oracle/baseline.c: OpenMP, thread-segment SOR with one axis-0
    slab per thread on the finest level and one row block per thread on the
    coarse levels, built -O3 -march=native on this host): the median of
    `runs` complete solves from x = 0 to 1e-8; the hierarchy is built once,
    untimed.  Also the relative differences of the GPU mixed-precision and
    FP64 solutions to this FP64 CPU solution (its thread segments take
    another iteration path, so the difference measures the solve tolerance,
    not the precision)."""
    import oracle as O
    from lineal_amd import problems

    native = True
    try:
        O.baseline_lib(native=True)
    except Exception as e:  # no compiler for this host: the portable build
        log(f"  native baseline build unavailable ({e}); using the x86-64-v3 build")
        native = False
    T = cpu_threads()
    dims = (n,) * 3
    c = problems.lognormal(dims, sigma, 1)
    S = O.Stencil(dims, coeff=c)
    b = S.rhs()
    t0 = time.perf_counter()
    H = O.BaselineHierarchy(S, native)
    t_setup = time.perf_counter() - t0
    log(f"  cpu hierarchy {t_setup:.1f} s (one thread, untimed)")
    times, its, x = [], [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        x, it = O.baseline_pcg(S, b, H, rel_tol=1e-8, max_it=1000, threads=T, native=native)
        times.append(time.perf_counter() - t0)
        its.append(it)
        log(f"  cpu solve {times[-1]:.2f} s, {it} iterations, {T} threads")
    err = float(np.abs(x_gpu - x).max() / np.abs(x).max())
    err64 = float(np.abs(x_gpu_fp64 - x).max() / np.abs(x).max())
    return {
        "value": float(np.median(times)),
        "unit": "s",
        "cores": T,
        "kind": "port",
        "sample": (f"the oracle's FP64 PCG-AMG to rel. residual 1e-8 at {n}^3 lognormal sigma={sigma}, full "
                   f"solve from x = 0, median of {runs} runs ({', '.join(f'{t:.2f}' for t in times)} s, "
                   f"{its[-1]} iterations) on {T} threads of {cpu_model()} (OpenMP, thread-segment SOR: "
                   f"one axis-0 slab per thread on the finest level, one row block per thread on coarse "
                   f"levels; -O3 -march={'native' if native else 'x86-64-v3'}); hierarchy built once "
                   f"({t_setup:.1f} s, untimed)"),
        "mixed_gpu_vs_cpu_fp64_rel_error": err,
        "fp64_gpu_vs_cpu_fp64_rel_error": err64,
        "cpu_iterations": its[-1],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=3, help="CPU baseline solves (median)")
    ap.add_argument("--workload", default="lognormal", choices=["lognormal", "two-phase", "mixed512"],
                    help="lognormal: the headline AMG-CG metric (default); two-phase and mixed512: "
                         "the other solve configurations, one JSON line each")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this path has no multi-GPU implementation "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import lineal_amd as la
    from lineal_amd import problems

    if not os.path.exists(la.LIB_PATH):
        from lineal_amd import build as hip_build

        hip_build.build()

    if args.workload == "two-phase":
        return two_phase(args)
    if args.workload == "mixed512":
        return mixed512(args)

    n = args.n
    dims = (n, n, n)
    N = n ** 3
    ctx = la.default_context()
    log(f"generating {n}^3 lognormal field sigma={args.sigma}")
    c = problems.lognormal(dims, args.sigma, 1)
    A = la.StencilMatrix.values(la.BoxInfo(dims), c, ctx)
    b = A.rhs()
    # ---- single-kernel rooflines (HIP events on the library stream) ----
    rot = 3  # distinct buffer sets so the 256 MB infinity cache cannot serve repeats
    xs = [la.DenseVector(problems.uniform(N, 10 + k), ctx=ctx) for k in range(rot)]
    ys = [la.DenseVector(N, ctx=ctx) for _ in range(rot)]
    launches = 120
    t_spmv = kernel_ms(ctx, lambda k=0: A.spmv(xs[k % rot], ys[k % rot]), launches)
    spmv_bytes = 24 * N  # coefficient + x read, y written (f64)
    xs32 = [la.DenseVector(problems.uniform(N, 20 + k, np.float32), ctx=ctx) for k in range(rot)]
    ys32 = [la.DenseVector(N, np.float32, ctx=ctx) for _ in range(rot)]
    t_spmv32 = kernel_ms(ctx, lambda k=0: A.spmv(xs32[k % rot], ys32[k % rot]), launches)
    bs = [la.DenseVector(problems.uniform(N, 30 + k), ctx=ctx) for k in range(rot)]
    t_sor = kernel_ms(ctx, lambda k=0: A.sor(xs[k % rot], ys[k % rot], bs[k % rot], 1.0, True, False), 10)
    sor_bytes = 32 * N  # coefficient, x_in, b read, x_out written
    bs32 = [la.DenseVector(problems.uniform(N, 40 + k, np.float32), ctx=ctx) for k in range(rot)]
    t_sor32 = kernel_ms(ctx, lambda k=0: A.sor(None, ys32[k % rot], bs32[k % rot], 1.0, True, False), 10)
    sor32_bytes = 12 * N  # coefficient, b read, x_out written (pre-smoother from zero)
    kern = {
        "spmv_values_f64": {"ms": t_spmv, "GB/s": spmv_bytes / t_spmv / 1e6},
        "spmv_values_f32": {"ms": t_spmv32, "GB/s": 12 * N / t_spmv32 / 1e6},
        "sor_fwd_values_f64": {"ms": t_sor, "GB/s": sor_bytes / t_sor / 1e6},
        "sor_fwd_zero_values_f32": {"ms": t_sor32, "GB/s": sor32_bytes / t_sor32 / 1e6},
    }
    for k, v in kern.items():
        log(f"{k}: {v['ms']:.3f} ms, {v['GB/s']:.0f} GB/s")
    del xs, ys, xs32, ys32, bs, bs32

    log("AMG setup (FP32 hierarchy)")
    M32 = la.MultiGridSolver(precision=np.float32).instantiate(A)
    log(f"  {M32.setup_seconds:.2f} s, levels {[M32.level_size(l) for l in range(M32.level_num)]}")
    M64 = la.MultiGridSolver(precision=np.float64).instantiate(A)
    log(f"AMG setup (FP64 hierarchy) {M64.setup_seconds:.2f} s, by phase "
        + ", ".join(f"{k} {v:.2f}" for k, v in M64.setup_split.items()))

    x = la.DenseVector(N, ctx=ctx)

    def solve(M):
        x.fill(0.0)
        return la.ConjugateGradientsSolver(8, M).solve(A, x, b, rel_tol=1e-8, max_it=500)

    for _ in range(args.warmup):
        r = solve(M32)
        log(f"warmup mixed: {r.iterations} it, status {r.status}")
    ctx.synchronize()
    all_classes = (("stencil_sor_fwd", la.KT_STENCIL_SOR_FWD), ("stencil_sor_bwd", la.KT_STENCIL_SOR_BWD),
                   ("stencil_pass", la.KT_STENCIL_PASS), ("csr_sor", la.KT_CSR_SOR),
                   ("csr_pass", la.KT_CSR_PASS))

    def class_times(names, solves):
        out = {}
        for name, cls in all_classes:
            if name in names:
                ms, cnt = ctx.kernel_time(cls)
                out[name] = {"ms_total_per_solve": ms / solves, "launches_per_solve": cnt / solves,
                             "ms_per_launch": ms / max(cnt, 1)}
        return out

    # time split of one untimed solve with every kernel class bracketed by HIP
    # events (each event pair adds a few microseconds between kernels, so the
    # timed solves below bracket only the finest-level sweeps)
    ctx.kernel_timing(True)
    solve(M32)
    ctx.synchronize()
    split = class_times([k for k, _ in all_classes], 1)
    ctx.kernel_timing(False)
    for k, v in split.items():
        log(f"time split {k}: {v['ms_total_per_solve']:.1f} ms/solve over {v['launches_per_solve']:.0f} "
            f"launches, {v['ms_per_launch'] * 1e3:.1f} us each")
    # timed region: every launch of the finest-level SOR sweeps (the dominant
    # kernel class) is bracketed by HIP events on the library's stream
    sor_names = ("stencil_sor_fwd", "stencil_sor_bwd")
    ctx.kernel_timing((1 << la.KT_STENCIL_SOR_FWD) | (1 << la.KT_STENCIL_SOR_BWD))
    t0 = time.perf_counter()
    its32 = []
    for _ in range(args.steps):
        r = solve(M32)
        its32.append(r.iterations)
        assert r.converged, r
    ctx.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    classes = class_times(sor_names, args.steps)
    ctx.kernel_timing(False)
    for k, v in classes.items():
        log(f"timed region {k}: {v['ms_total_per_solve']:.1f} ms/solve over {v['launches_per_solve']:.0f} "
            f"launches, {v['ms_per_launch'] * 1e3:.1f} us each")
    x_mixed = x.numpy()
    log(f"mixed: {t_mixed * 1e3:.1f} ms/solve, iterations {its32}")

    r = solve(M64)
    ctx.synchronize()
    t0 = time.perf_counter()
    its64 = []
    for _ in range(args.steps):
        r = solve(M64)
        its64.append(r.iterations)
        assert r.converged, r
    ctx.synchronize()
    t_fp64 = (time.perf_counter() - t0) / args.steps
    x_fp64 = x.numpy()
    log(f"fp64: {t_fp64 * 1e3:.1f} ms/solve, iterations {its64}")
    rel_err = float(np.abs(x_mixed - x_fp64).max() / np.abs(x_fp64).max())
    # the solution both 1e-8 solves approximate: the FP64 solve run to 1e-13
    # (untimed); each 1e-8 solve's error against it, max-norm relative
    x.fill(0.0)
    r_tight = la.ConjugateGradientsSolver(8, M64).solve(A, x, b, rel_tol=1e-13, max_it=1000)
    x_tight = x.numpy()
    scale = float(np.abs(x_tight).max())
    err_mixed = float(np.abs(x_mixed - x_tight).max() / scale)
    err_fp64 = float(np.abs(x_fp64 - x_tight).max() / scale)
    log(f"error against the FP64 solution at 1e-13 ({r_tight.iterations} it): mixed {err_mixed:.2e}, "
        f"fp64 {err_fp64:.2e}")

    # The dominant kernel of the mixed solve: the finest-level SOR sweep in
    # f32 (a forward sweep from zero as pre-smoother, a backward sweep from the
    # prolonged iterate as post-smoother).  Algorithmic bytes per launch (the
    # matrix-free figure): the f32 coefficient and right-hand side read, x_out
    # written, plus x_in read by the backward sweep, each cell once.  The
    # sweep reads its face weights and diagonal from the operator's cache
    # instead of the coefficients; kernel_bytes counts those.
    dom = max(("stencil_sor_fwd", "stencil_sor_bwd"), key=lambda k: classes[k]["ms_total_per_solve"])
    dom_bytes = (12 if dom == "stencil_sor_fwd" else 16) * N
    dom_ms = classes[dom]["ms_per_launch"]
    achieved = dom_bytes / dom_ms / 1e6
    kernel_bytes = dom_bytes - 4 * N + sor_cache_bytes(n, 4)
    # HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of
    # the same kernels at 256^3 (python -m lineal_amd.profile_kernels), FETCH_SIZE
    # scaled by the factor a plain copy kernel of known size gives
    traffic = TRAFFIC_256.get(dom) if n == 256 else None
    result = {
        "metric": "AMG-CG time-to-1e-8 at 256^3, mixed vs FP64 (s); SpMV HBM GB/s vs peak",
        "value": t_mixed,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64",
        "data": "synthetic",
        "config": {
            "workload": f"AMG-PCG to rel. residual 1e-8, {n}^3 7-point diffusion stencil, lognormal "
                        f"sigma={args.sigma} per-cell coefficients (seed 1), harmonic-mean faces, "
                        f"Dirichlet 1/0 on x, matrix-free finest level, SOR(1) V-cycle",
            "grid": [n, n, n],
            "fp64_seconds": t_fp64,
            "mixed_seconds": t_mixed,
            "mixed_over_fp64": t_mixed / t_fp64,
            "iterations_mixed": its32[-1],
            "iterations_fp64": its64[-1],
            "mixed_vs_fp64_rel_error": rel_err,
            "mixed_vs_converged_fp64_rel_error": err_mixed,
            "fp64_vs_converged_fp64_rel_error": err_fp64,
            "setup_seconds_mixed": M32.setup_seconds,
            "setup_seconds_fp64": M64.setup_seconds,
            "setup_split_fp64": M64.setup_split,
            "levels": [M32.level_size(l) for l in range(M32.level_num)],
            "kernels": kern,
            "timed_region_kernel_classes": classes,
            "time_split_untimed_solve": split,
            "spmv_roofline": {"bound": "hbm", "achieved": spmv_bytes / t_spmv / 1e6, "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": spmv_bytes / t_spmv / 1e6 / HBM_PEAK_GBS},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": {"stencil_sor_fwd": "sor_wavefront_r<float, forward, regular, x_in = 0> "
                                          "(finest-level pre-smoother)",
                       "stencil_sor_bwd": "sor_wavefront_r<float, backward, regular, x_in> "
                                          "(finest-level post-smoother)"}[dom],
            "algorithmic_bytes": dom_bytes,
            "kernel_bytes": kernel_bytes,
            "kernel_bytes_frac": kernel_bytes / dom_ms / 1e6 / HBM_PEAK_GBS,
            "ms_per_launch": dom_ms,
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (the oracle on the host's threads)")
        cb = cpu_baseline(n, args.sigma, args.cpu_runs, x_mixed, x_fp64)
        result["config"]["mixed_gpu_vs_cpu_fp64_rel_error"] = cb.pop("mixed_gpu_vs_cpu_fp64_rel_error")
        result["config"]["fp64_gpu_vs_cpu_fp64_rel_error"] = cb.pop("fp64_gpu_vs_cpu_fp64_rel_error")
        result["config"]["cpu_fp64_iterations"] = cb.pop("cpu_iterations")
        result["cpu_baseline"] = cb
    print(json.dumps(result), flush=True)


def timed_solves(ctx, solve, steps, warmup):
    for _ in range(warmup):
        solve()
    ctx.synchronize()
    t0 = time.perf_counter()
    res = None
    for _ in range(steps):
        res = solve()
        assert res.converged, res
    ctx.synchronize()
    return (time.perf_counter() - t0) / steps, res


def two_phase(args):
    """256^3 two-phase lookup field (u8 ids from seeded random spheres of radius
    4-12 until 35% solid; coefficients 1.0 pore / 1e-6 solid): FP64 AMG-CG to
    1e-8, and BiCGStab with first-order upwind advection v = (10 / L_x, 0, 0)
    (Peclet number v L_x / D_ref = 10 with D_ref = 1) preconditioned by the
    FP64 AMG of the advection operator (LU coarsest solve)."""
    import lineal_amd as la
    from lineal_amd import problems

    n = args.n
    dims = (n, n, n)
    N = n ** 3
    ctx = la.default_context()
    log(f"generating {n}^3 spheres (35% solid)")
    m = problems.spheres(dims, 3)
    mc = (1.0, 1e-6, 0.0, 0.0)
    A = la.StencilMatrix.lookup(la.BoxInfo(dims), m, mc, ctx)
    v = 10.0 / n
    Aadv = la.StencilMatrix.lookup(la.BoxInfo(dims, velocity_x=v), m, mc, ctx)
    M = la.MultiGridSolver(precision=np.float64).instantiate(A)
    log(f"AMG setup {M.setup_seconds:.2f} s, levels {[M.level_size(l) for l in range(M.level_num)]}")
    # BiCGStab's preconditioner: the AMG of the nonsymmetric advection
    # operator itself, with an LU coarsest solve
    Madv = la.MultiGridSolver(precision=np.float64).instantiate(Aadv)
    log(f"nonsymmetric AMG setup {Madv.setup_seconds:.2f} s, levels "
        f"{[Madv.level_size(l) for l in range(Madv.level_num)]}")
    b, badv = A.rhs(), Aadv.rhs()
    x = la.DenseVector(N, ctx=ctx)

    def cg():
        x.fill(0.0)
        return la.ConjugateGradientsSolver(8, M).solve(A, x, b, rel_tol=1e-8, max_it=1000)

    def bicg():
        x.fill(0.0)
        return la.BiCgStabSolver(8, Madv).solve(Aadv, x, badv, rel_tol=1e-8, max_it=1000)

    steps, warmup = max(1, min(args.steps, 3)), min(args.warmup, 1)
    t_cg, r_cg = timed_solves(ctx, cg, steps, warmup)
    log(f"AMG-CG: {t_cg * 1e3:.1f} ms, {r_cg.iterations} iterations")
    t_bi, r_bi = timed_solves(ctx, bicg, steps, warmup)
    r = la.DenseVector(N, ctx=ctx)
    Aadv.residual(badv, x, r)
    true_rel = float(np.sqrt(r.dot(r) / badv.dot(badv)))
    log(f"BiCGStab (Pe 10): {t_bi * 1e3:.1f} ms, {r_bi.iterations} iterations, true rel. residual {true_rel:.2e}")
    print(json.dumps({
        "metric": "two-phase 256^3 solve time (s): FP64 AMG-CG to 1e-8; BiCGStab with upwind advection Pe 10",
        "value": t_cg, "unit": "s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": t_cg * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n}^3 two-phase lookup stencil (spheres r 4-12, 35% solid, 1.0 / 1e-6), "
                               "FP64 AMG-PCG rel. tol 1e-8; BiCGStab + upwind advection v = 10/L_x",
                   "cg_seconds": t_cg, "cg_iterations": r_cg.iterations,
                   "bicgstab_seconds": t_bi, "bicgstab_iterations": r_bi.iterations,
                   "bicgstab_true_rel_residual": true_rel, "setup_seconds": M.setup_seconds,
                   "setup_seconds_nonsymmetric": Madv.setup_seconds,
                   "solid_fraction": float(m.mean())},
    }), flush=True)


def mixed512(args):
    """512^3 lognormal sigma = 2: FP32-hierarchy AMG inside FP64 CG against the
    plain FP64 AMG-CG, both to 1e-8; iterations, times, the mixed solution's
    relative difference to the FP64 one and the FP64 solution's true relative
    residual."""
    import lineal_amd as la
    from lineal_amd import problems

    n = 512 if args.n == 256 else args.n
    sigma = 2.0 if args.sigma == 1.0 else args.sigma
    dims = (n, n, n)
    N = n ** 3
    ctx = la.default_context()
    log(f"generating {n}^3 lognormal field sigma={sigma}")
    A = la.StencilMatrix.values(la.BoxInfo(dims), problems.lognormal(dims, sigma, 1), ctx)
    b = A.rhs()
    M32 = la.MultiGridSolver(precision=np.float32).instantiate(A)
    log(f"FP32 AMG setup {M32.setup_seconds:.1f} s, levels {[M32.level_size(l) for l in range(M32.level_num)]}")
    M64 = la.MultiGridSolver(precision=np.float64).instantiate(A)
    log(f"FP64 AMG setup {M64.setup_seconds:.1f} s")
    x = la.DenseVector(N, ctx=ctx)

    def make(M):
        def solve():
            x.fill(0.0)
            return la.ConjugateGradientsSolver(8, M).solve(A, x, b, rel_tol=1e-8, max_it=2000)
        return solve

    steps, warmup = max(1, min(args.steps, 2)), 0
    t32, r32 = timed_solves(ctx, make(M32), steps, warmup)
    x32 = x.numpy()
    log(f"mixed: {t32 * 1e3:.1f} ms, {r32.iterations} iterations")
    t64, r64 = timed_solves(ctx, make(M64), steps, warmup)
    x64 = x.numpy()
    r = la.DenseVector(N, ctx=ctx)
    A.residual(b, x, r)
    true_rel = float(np.sqrt(r.dot(r) / b.dot(b)))
    rel = float(np.abs(x32 - x64).max() / np.abs(x64).max())
    log(f"fp64: {t64 * 1e3:.1f} ms, {r64.iterations} iterations; mixed vs fp64 {rel:.2e}; "
        f"fp64 true rel. residual {true_rel:.2e}")
    print(json.dumps({
        "metric": "AMG-CG time-to-1e-8 at 512^3 lognormal sigma=2, mixed vs FP64 (s)",
        "value": t32, "unit": "s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": t32 * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32/f64", "data": "synthetic",
        "config": {"workload": f"{n}^3 lognormal sigma={sigma} values stencil, AMG-PCG rel. tol 1e-8",
                   "mixed_seconds": t32, "fp64_seconds": t64, "mixed_over_fp64": t32 / t64,
                   "iterations_mixed": r32.iterations, "iterations_fp64": r64.iterations,
                   "mixed_vs_fp64_rel_error": rel, "fp64_true_rel_residual": true_rel,
                   "setup_seconds_mixed": M32.setup_seconds, "setup_seconds_fp64": M64.setup_seconds},
    }), flush=True)


if __name__ == "__main__":
    main()
