#!/usr/bin/env python3
"""Per-kernel rooflines of one AMG-CG solve from a rocprofv3 kernel trace.

Run one solve under the profiler, then summarise its kernels:

    rocprofv3 --kernel-trace --output-format csv -d OUT -o run -- \
        python3 -m lineal_amd.profile_solve run --n 256 --precision f32 --levels OUT/levels.json
    python3 -m lineal_amd.profile_solve summary OUT/run_kernel_trace.csv OUT/levels.json

`run` builds the hierarchy, runs one warm-up solve, pauses for half a second
and runs the solve that `summary` reads (the dispatches after the last pause
of more than 0.3 s).  `summary` walks that solve's dispatches in order: every
multigrid cycle starts with the finest forward sweep and visits the levels
down (sweep, residual, restriction), solves the coarsest level and goes up
(prolongation, sweep), so each dispatch is assigned its level.  Each kernel
class gets the algorithmic bytes of one launch (every array it must read or
write once; gathered values counted once per row) and the achieved GB/s
against the 8 TB/s HBM peak.
"""

from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import sys
import time

HBM_PEAK_GBS = 8000.0


def run(args):
    import numpy as np

    import lineal_amd as la
    from lineal_amd import problems

    n = args.n
    dims = (n, n, n)
    ctx = la.default_context()
    A = la.StencilMatrix.values(la.BoxInfo(dims), problems.lognormal(dims, args.sigma, 1), ctx)
    b = A.rhs()
    prec = np.float32 if args.precision == "f32" else np.float64
    M = la.MultiGridSolver(precision=prec, coarse_segment_rows=args.coarse_segment_rows,
                           sequential_rows=args.sequential_rows).instantiate(A)
    x = la.DenseVector(n ** 3, ctx=ctx)

    def solve():
        x.fill(0.0)
        return la.ConjugateGradientsSolver(8, M).solve(A, x, b, rel_tol=1e-8, max_it=500)

    solve()
    ctx.synchronize()
    time.sleep(0.5)
    t0 = time.perf_counter()
    r = solve()
    ctx.synchronize()
    t = time.perf_counter() - t0
    info = {"n": n, "precision": args.precision, "iterations": r.iterations, "seconds": t,
            "levels": [{"n": M.level_size(l), "nnz": M.level_nnz(l) if l > 0 else 7 * n ** 3}
                       for l in range(M.level_num)]}
    with open(args.levels, "w") as f:
        json.dump(info, f)
    print(f"solve {t * 1e3:.1f} ms, {r.iterations} iterations", flush=True)


def short(name: str) -> str:
    k = name.replace("(anonymous namespace)::", "").replace("lgpu::", "").replace("void ", "")
    return re.sub(r"\(.*$", "", k)


def last_solve(rows):
    ev = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), short(r["Kernel_Name"])) for r in rows)
    start = 0
    for i in range(len(ev) - 1):
        if ev[i + 1][0] - ev[i][1] > 300_000_000:
            start = i + 1
    return ev[start:]


def esize(k: str) -> int:
    return 4 if "<float" in k else 8


def summary(args):
    info = json.load(open(args.levels))
    lv = info["levels"]
    L = len(lv)
    N = lv[0]["n"]
    ev = last_solve(list(csv.DictReader(open(args.trace))))
    span = (ev[-1][1] - ev[0][0]) / 1e6
    busy = sum(e - s for s, e, _ in ev) / 1e6
    # walk the cycles: position -> level
    stats = collections.defaultdict(lambda: [0, 0])  # (class, level) -> [ns, calls]
    level = None
    going_up = False
    for s, e, k in ev:
        base = k.split("<")[0]
        targs = [a.strip() for a in k[k.find("<") + 1: k.rfind(">")].split(",")] if "<" in k else []
        if base == "sor_wavefront_r" or base == "sor_wavefront":
            fwd = targs[1 if base == "sor_wavefront_r" else 2] == "true"
            if fwd:
                level, going_up = 0, False
            lvl = 0
        elif base in ("csr_sor_wave_kernel", "csr_sor_segment_kernel", "csr_sor_kernel",
                      "csr_sor_sched_kernel", "csr_sor_lds_kernel"):
            fwd = targs[1 if base == "csr_sor_lds_kernel" else 2] == "true"
            if fwd:
                level = (level or 0) + 1
                lvl = level
            else:
                lvl = level
        elif base in ("chol_solve_kernel", "lu_solve_kernel"):
            lvl = L - 1
            going_up = True
            level = L - 1
        elif base == "amg_tail_kernel":
            # the levels after the last one swept down, and the coarsest, in
            # one launch: the next prolongation is the last down level's
            lvl = "tail"
            going_up = True
            level = (level or 0) + 1
        elif base == "prolong_kernel":
            level = (level if level is not None else L - 1) - 1
            lvl = level
        elif base == "restrict_staged_kernel" or base == "restrict_kernel":
            lvl = (level or 0) + 1  # the level written
        elif base == "csr_staged_kernel" or base == "csr_kernel":
            lvl = level if level is not None else 0
        elif base == "stencil_kernel" and level == 0 and not going_up and targs[3] == "1":
            lvl = 0  # the finest residual of the cycle
        else:
            lvl = None  # Krylov vector passes: the finest size
        stats[(k, lvl)][0] += e - s
        stats[(k, lvl)][1] += 1

    def nbytes(k: str, lvl):
        base = k.split("<")[0]
        targs = [a.strip() for a in k[k.find("<") + 1: k.rfind(">")].split(",")] if "<" in k else []
        E = esize(k)
        if base.startswith("sor_wavefront"):
            has_x = targs[3 if base == "sor_wavefront_r" else 4] == "true"
            return (3 + has_x) * E * N
        if base == "stencil_kernel":
            kind, mode = int(targs[1]), int(targs[3])
            c = 1 if kind == 1 else E
            return (c + (2 if mode in (0, 2) else 3) * E) * N
        if lvl == "tail":
            return None
        if lvl is None:
            return {"cg_update_kernel": 52 * N, "xpby_scal_kernel": (4 if E == 4 else 8) * N + 16 * N,
                    "dot_mixed_kernel": 12 * N, "dot_kernel": 2 * E * N, "axpy_scal_kernel": 24 * N,
                    "convert_kernel": 12 * N, "fill_kernel": 8 * N}.get(base)
        n, nnz = lv[lvl]["n"], lv[lvl]["nnz"]
        if base in ("csr_sor_wave_kernel", "csr_sor_segment_kernel", "csr_sor_kernel",
                    "csr_sor_sched_kernel", "csr_sor_lds_kernel"):
            has_x = targs[3 if base == "csr_sor_lds_kernel" else 4] == "true"
            return nnz * (E + 4) + 8 * n + (2 + has_x) * E * n
        if base in ("csr_staged_kernel", "csr_kernel"):
            return nnz * (E + 4) + 8 * n + 3 * E * n
        if base.startswith("restrict"):
            nf = lv[lvl - 1]["n"]
            return nf * (E + 4) + 8 * n + E * n
        if base == "prolong_kernel":
            nc = lv[lvl + 1]["n"]
            return n * (3 * E + 4) + E * nc
        return None

    rows = []
    for (k, lvl), (ns, calls) in stats.items():
        avg_us = ns / calls / 1e3
        by = nbytes(k, lvl)
        gbs = by / (avg_us * 1e3) if by else None
        rows.append((ns, k, lvl, calls, avg_us, by, gbs))
    rows.sort(key=lambda r: -r[0])
    print(f"One {info['precision']} solve at {info['n']}^3: {info['iterations']} iterations, "
          f"{span:.1f} ms from the first to the last dispatch, {busy:.1f} ms in kernels.\n")
    print("| kernel | level (rows) | calls | ms per solve | us per call | MB per call | GB/s | of 8 TB/s |")
    print("|---|---|---:|---:|---:|---:|---:|---:|")
    for ns, k, lvl, calls, avg_us, by, gbs in rows:
        if lvl is None:
            where = "finest / Krylov"
        elif lvl == "tail":
            where = "tail (the small levels and the coarsest)"
        else:
            where = f"{lvl} ({lv[lvl]['n']})"
        mb = f"{by / 1e6:.1f}" if by else "-"
        g = f"{gbs:.0f}" if gbs else "-"
        frac = f"{gbs / HBM_PEAK_GBS * 100:.1f}%" if gbs else "-"
        print(f"| `{k}` | {where} | {calls} | {ns / 1e6:.2f} | {avg_us:.1f} | {mb} | {g} | {frac} |")


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--n", type=int, default=256)
    r.add_argument("--sigma", type=float, default=1.0)
    r.add_argument("--precision", choices=["f32", "f64"], default="f32")
    r.add_argument("--coarse-segment-rows", type=int, default=64)
    r.add_argument("--sequential-rows", type=int, default=None,
                   help="levels of at most this many rows swept in one segment (default: the library's)")
    r.add_argument("--levels", required=True)
    s = sub.add_parser("summary")
    s.add_argument("trace")
    s.add_argument("levels")
    args = ap.parse_args(argv)
    if args.cmd == "run":
        run(args)
    else:
        summary(args)


if __name__ == "__main__":
    sys.exit(main())
