#!/usr/bin/env python3
"""Launch single hot kernels repeatedly, for rocprofv3 kernel traces and
hardware-counter runs.

    python -m lineal_amd.profile_kernels --n 256 --reps 10 spmv_f64 sor_f32_zero copy_f64

Kernels (all on a lognormal sigma=1 values stencil of n^3 cells):
  spmv_f64 / spmv_f32      y = A x
  resid_f32                r = b - A x
  jacobi_f64               one Jacobi sweep (omega = 2/3)
  sor_f64 / sor_f32 / sor_f32_zero   one forward lexicographic SOR sweep (with x_in / from zero)
  sor_f64_bwd / sor_f32_bwd          one backward sweep with x_in (the V-cycle's post-smoother)
  csr_spmv_f64 / csr_spmv_f32   y = A x with the assembled CSR copy
  lookup_spmv_f64          y = A x with the u8 lookup (two-phase spheres) operator
  copy_f64 / copy_f32      a plain vector copy of n^3 entries (counter calibration)
Three distinct buffer sets rotate between launches so that repeated launches
cannot be served from the 256 MB infinity cache.
"""

from __future__ import annotations

import argparse
import sys

import numpy as np

import lineal_amd as la
from lineal_amd import problems


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("kernels", nargs="+")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--sor-cache", type=int, default=1,
                    help="1: SOR sweeps read cached face weights and diagonal; 0: evaluated on the fly")
    args = ap.parse_args(argv)
    n = args.n
    dims = (n, n, n)
    N = n ** 3
    ctx = la.default_context()
    A = la.StencilMatrix.values(la.BoxInfo(dims), problems.lognormal(dims, 1.0, 1), ctx)
    A.set_sor_cache(bool(args.sor_cache))
    rot = 3
    vec = {}

    def vs(dtype, seed):
        key = (np.dtype(dtype).str, seed)
        if key not in vec:
            vec[key] = [la.DenseVector(problems.uniform(N, seed + k, dtype), ctx=ctx) for k in range(rot)]
        return vec[key]

    csr = {}
    for name in args.kernels:
        if name.startswith("spmv_") or name.startswith("resid") or name.startswith("jacobi"):
            dt = np.float32 if name.endswith("f32") else np.float64
            x, y, b = vs(dt, 10), vs(dt, 20), vs(dt, 30)
            if name.startswith("spmv"):
                fn = lambda k: A.spmv(x[k % rot], y[k % rot])
            elif name.startswith("resid"):
                fn = lambda k: A.residual(b[k % rot], x[k % rot], y[k % rot])
            else:
                fn = lambda k: A.jacobi(x[k % rot], y[k % rot], b[k % rot], 2.0 / 3.0)
        elif name.startswith("sor_"):
            dt = np.float32 if "f32" in name else np.float64
            x, y, b = vs(dt, 10), vs(dt, 20), vs(dt, 30)
            zero = name.endswith("zero")
            fwd = "bwd" not in name
            fn = lambda k: A.sor(None if zero else x[k % rot], y[k % rot], b[k % rot], 1.0, fwd, False)
        elif name.startswith("csr_spmv"):
            dt = np.float32 if name.endswith("f32") else np.float64
            if dt not in csr:
                csr[dt] = A.to_csr(dt)
            C = csr[dt]
            x, y = vs(dt, 10), vs(dt, 20)
            fn = lambda k: C.spmv(x[k % rot], y[k % rot])
        elif name == "lookup_spmv_f64":
            L = la.StencilMatrix.lookup(la.BoxInfo(dims), problems.spheres(dims, 3), (1.0, 1e-6, 0.0, 0.0), ctx)
            x, y = vs(np.float64, 10), vs(np.float64, 20)
            fn = lambda k, L=L: L.spmv(x[k % rot], y[k % rot])
        elif name.startswith("copy_"):
            dt = np.float32 if name.endswith("f32") else np.float64
            x, y = vs(dt, 10), vs(dt, 20)
            fn = lambda k: y[k % rot].assign(x[k % rot])
        else:
            sys.exit(f"unknown kernel {name}")
        fn(0)
        ctx.synchronize()
        ctx.timer_start()
        for k in range(args.reps):
            fn(k)
        ms = ctx.timer_stop() / args.reps
        print(f"{name}: {ms:.4f} ms per launch (n={n}, HIP events, {args.reps} launches)", flush=True)


if __name__ == "__main__":
    main()
