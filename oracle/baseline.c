/*
 * CPU baseline of the solve phase: the oracle's PCG-AMG with its work spread
 * over host threads the way the reference's AdaptivePolicy spreads it
 * (parallel/execution-policy/adaptive.hpp:209-279).  TEST INFRASTRUCTURE
 * ONLY: bench.py's cpu_baseline leg times it; nothing in the product loads
 * it, and it is not a parity checker (the T = 1 oracle is).
 *
 * With T threads:
 *  - SOR sweeps use T thread segments (gauss-seidel-variant/sweep.hpp:153-169):
 *    one slab of axis-0 planes per thread on the matrix-free finest level
 *    (the UniformIndexSegmenter over planes, adaptive.hpp:209-234) and T
 *    contiguous row blocks on the CSR levels (adaptive.hpp:266-279); each
 *    segment is swept lexicographically with new values from itself only;
 *  - SpMV, residuals, transfers and vector updates split the rows evenly;
 *  - dot products sum per-thread partials (the reference's per-thread
 *    accumulators merged at the end, component-wise/sink/accumulator.hpp);
 *  - the coarsest Cholesky solve runs on one thread.
 * The hierarchy is the oracle's (one thread segment during setup), built
 * once and not timed; the V-cycle is DefaultCycle{REGULAR, 1} with relax 1.
 */
#include "oracle.c"

#include <omp.h>

static void bl_residual(const or_mat* a, const double* b, const double* x, double* y, int T) {
  const int64_t n = mat_rows_f64(a);
#pragma omp parallel num_threads(T)
  {
    row_t_f64 r = {0};
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      get_row_f64(a, i, &r);
      double sum = 0;
      for (int k = 0; k < r.n; ++k) sum = fma(r.vals[k], x[r.cols[k]], sum);
      y[i] = b ? b[i] - sum : sum;
    }
    free(r.dyn);
  }
}

static double bl_dot(const double* a, const double* b, int64_t n, int T) {
  double s = 0;
#pragma omp parallel for num_threads(T) schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* thread segments of level l: axis-0 slabs of the stencil, row blocks of a
 * CSR level */
static int64_t* bl_segments(const or_hierarchy* h, int l, int T, int* nseg) {
  int64_t* seg = (int64_t*)malloc(sizeof(int64_t) * (size_t)(T + 1));
  const int64_t n = h->size[l];
  if (l == 0 && !h->finest.is_csr) {
    const int64_t nx = h->finest.st->n[0], plane = n / nx;
    for (int t = 0; t <= T; ++t) seg[t] = (nx * t / T) * plane;
  } else {
    for (int t = 0; t <= T; ++t) seg[t] = n * t / T;
  }
  *nseg = T;
  return seg;
}

static void bl_sor(const or_mat* m, const double* x_in, double* x_out, const double* b, int fwd,
                   const int64_t* seg, int nseg, int T) {
#pragma omp parallel for num_threads(T) schedule(static, 1)
  for (int s = 0; s < nseg; ++s)
    if (seg[s + 1] > seg[s]) or_sor_f64(m, x_in, x_out, b, 1.0, fwd, 0, 1, seg + s);
}

static void bl_vcycle(const or_hierarchy* h, const double* r, double* z, double** rhs,
                      double** sol, double* tmp, double* aux, int T) {
  const int L = h->nlev;
  for (int l = 0; l + 1 < L; ++l) {
    or_mat m = h->finest;
    if (l > 0) {
      m.is_csr = 1;
      m.st = NULL;
      m.csr = &h->mats[l];
    }
    int nseg;
    int64_t* seg = bl_segments(h, l, T, &nseg);
    double* x = l == 0 ? z : sol[l];
    const double* b = l == 0 ? r : rhs[l];
    bl_sor(&m, NULL, x, b, 1, seg, nseg, T);
    free(seg);
    bl_residual(&m, b, x, tmp, T);
    const int64_t* ptr = h->c2f_ptr[l];
    const int64_t* idx = h->c2f_idx[l];
    double* out = rhs[l + 1];
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t c = 0; c < h->size[l + 1]; ++c) {
      double sum = 0;
      for (int64_t k = ptr[c]; k < ptr[c + 1]; ++k) sum += tmp[idx[k]];
      out[c] = sum;
    }
  }
  if (h->lu) or_lu_solve_dense(h->size[L - 1], h->lu_l, h->lu_u, rhs[L - 1], sol[L - 1]);
  else or_cholesky_solve(&h->chol, rhs[L - 1], sol[L - 1]);
  for (int l = L - 2; l >= 0; --l) {
    or_mat m = h->finest;
    if (l > 0) {
      m.is_csr = 1;
      m.st = NULL;
      m.csr = &h->mats[l];
    }
    const int64_t n = h->size[l];
    double* x = l == 0 ? z : sol[l];
    const double* b = l == 0 ? r : rhs[l];
    const uint32_t* map = h->agg[l];
    const double* xc = sol[l + 1];
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t v = 0; v < n; ++v) {
      const double refined = map[v] < OR_ISOLATED ? xc[map[v]] : 0.0;
      aux[v] = x[v] + refined;
    }
    int nseg;
    int64_t* seg = bl_segments(h, l, T, &nseg);
    bl_sor(&m, aux, x, b, 0, seg, nseg, T);
    free(seg);
  }
}

/* PCG-AMG (conjugate-gradients/solver.hpp:50-115) with the true residual
 * every 8th iteration, stopping at ||r|| <= rel_tol ||r_0||; returns the
 * iterations. */
int bl_pcg(const or_mat* a, const or_hierarchy* h, const double* b, double* x, double rel_tol,
           int max_it, int threads) {
  const int T = threads > 0 ? threads : omp_get_max_threads();
  const int64_t n = mat_rows_f64(a);
  const int L = h->nlev;
  double *r = malloc(sizeof(double) * (size_t)n), *p = malloc(sizeof(double) * (size_t)n),
         *q = malloc(sizeof(double) * (size_t)n), *tmp = malloc(sizeof(double) * (size_t)n),
         *aux = malloc(sizeof(double) * (size_t)n);
  double** rhs = calloc((size_t)L, sizeof(double*));
  double** sol = calloc((size_t)L, sizeof(double*));
  for (int l = 1; l < L; ++l) {
    rhs[l] = calloc((size_t)h->size[l], sizeof(double));
    sol[l] = calloc((size_t)h->size[l], sizeof(double));
  }
  or_hierarchy view = *h;
  view.finest = *a;
  bl_residual(a, b, x, r, T);
  double rr = bl_dot(r, r, n, T);
  bl_vcycle(&view, r, p, rhs, sol, tmp, aux, T);
  double rho = bl_dot(p, r, n, T);
  const double r0 = sqrt(rr);
  int it = 0;
  while (it < max_it && sqrt(rr) > rel_tol * r0 && rho != 0) {
    bl_residual(a, NULL, p, q, T);
    const double alpha = rho / bl_dot(p, q, n, T);
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + p[i] * alpha;
    if (it > 0 && it % 8 == 0) {
      bl_residual(a, b, x, r, T);
    } else {
#pragma omp parallel for num_threads(T) schedule(static)
      for (int64_t i = 0; i < n; ++i) r[i] = r[i] - q[i] * alpha;
    }
    rr = bl_dot(r, r, n, T);
    const double old = rho;
    bl_vcycle(&view, r, q, rhs, sol, tmp, aux, T);
    rho = bl_dot(r, q, n, T);
    const double beta = rho / old;
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t i = 0; i < n; ++i) p[i] = q[i] + p[i] * beta;
    ++it;
  }
  for (int l = 1; l < L; ++l) {
    free(rhs[l]);
    free(sol[l]);
  }
  free(rhs);
  free(sol);
  free(r);
  free(p);
  free(q);
  free(tmp);
  free(aux);
  return it;
}
