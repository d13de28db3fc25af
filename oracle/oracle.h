/*
 * CPU oracle for the lineal solve-phase hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker against which the HIP
 * implementation in lineal_amd/ is compared.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * Every routine restates the reference algorithm with an explicit operation
 * order (fused multiply-adds written as fma(), everything else compiled with
 * -ffp-contract=off), so a GPU kernel that performs the same operations in the
 * same order agrees with it bit for bit.  Each function cites the reference
 * file:line it follows (paths relative to the reference's include/lineal/).
 *
 * Parity status per function (how the restatement itself is pinned) is listed
 * at the top of oracle.c.
 */
#ifndef LINEAL_ORACLE_H
#define LINEAL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Matrix-free 7-point diffusion stencil on an nx*ny*nz box (z fastest,
 * index = (x*ny + y)*nz + z), flow axis 0 with Dirichlet borders.
 * kind 0: per-cell coefficients (StationaryDiffusionValue(s)Valuator)
 * kind 1: u8 material ids + coefficient table (StationaryDiffusionLookupValuator) */
typedef struct {
  int64_t n[3];
  double len[3];
  int kind;
  int zero_to_one;
  const double* coeff;       /* kind 0, N entries */
  const uint8_t* mat;        /* kind 1, N entries */
  double mat_coeffs[4];      /* kind 1 */
  double p_start, p_end;
  /* first-order upwind advection with velocity v >= 0 along +x (not in the
   * reference): every cell's +x face carries the outflow v*a (a = h_y h_z),
   * the -x coupling gains -v*a, the diagonal +v*a after the diffusion
   * entries, and the x = 0 rhs +v*a*p_start (inflow of the Dirichlet value) */
  double velocity;
} or_stencil;

/* CSR matrix, columns sorted, diagonal stored; values kept as double (an f32
 * level stores float-representable values and is flagged). */
typedef struct {
  int64_t n;
  const int64_t* rowptr;
  const int64_t* cols;
  const double* vals;
  int is_f32;
  int symmetric;
} or_csr;

/* A matrix: either a stencil or a CSR matrix. */
typedef struct {
  int is_csr;
  const or_stencil* st;
  const or_csr* csr;
} or_mat;

/* geometry / valuators */
void or_axis_quotients(const or_stencil* s, double q[3]);
double or_harmonic_mean_f64(double a, double b);
float or_harmonic_mean_f32(float a, float b);

/* row evaluation in the reference's unordered `iterate` order:
 * -x,-y,-z,+z,+y,+x (existing neighbours only), then the diagonal last.
 * Returns the entry count (<= 7). */
int or_stencil_row_f64(const or_stencil* s, int64_t i, int64_t* cols, double* vals);
int or_stencil_row_f32(const or_stencil* s, int64_t i, int64_t* cols, float* vals);

void or_stencil_rhs_f64(const or_stencil* s, double* b);
void or_stencil_rhs_f32(const or_stencil* s, float* b);

/* y = A x */
/* the reference's vectorised row order (banded_iterate): y = A x with fused
 * or separately rounded steps, and one-segment SOR (compute_base_simd) */
void or_spmv_banded_f64(const or_stencil* s, const double* x, double* y, int fused);
void or_spmv_banded_f32(const or_stencil* s, const float* x, float* y, int fused);
void or_sor_banded_f64(const or_stencil* s, const double* x_in, double* x_out, const double* b,
                       double w, int forward, int ultra);
void or_sor_banded_f32(const or_stencil* s, const float* x_in, float* x_out, const float* b,
                       float w, int forward, int ultra);
void or_spmv_f64(const or_mat* a, const double* x, double* y);
void or_spmv_f32(const or_mat* a, const float* x, float* y);
/* r = b - A x */
void or_residual_f64(const or_mat* a, const double* b, const double* x, double* r);
void or_residual_f32(const or_mat* a, const float* b, const float* x, float* r);

/* Jacobi: x_out = fma(w, (b - sum_{j!=i} a_ij x_in_j)/a_ii, (1-w) x_in_i); x_in may be NULL (= 0) */
void or_jacobi_f64(const or_mat* a, const double* x_in, double* x_out, const double* b, double w);
void or_jacobi_f32(const or_mat* a, const float* x_in, float* x_out, const float* b, float w);

/* SOR sweep with the reference's segment semantics.  seg has nseg+1 row
 * boundaries (nseg = 1, seg = {0, N} gives lexicographic SOR).  x_in may be NULL (= 0).
 * forward: 1 forward, 0 backward; ultra: 0 regular, 1 the l1 ("ultra") variant. */
void or_sor_f64(const or_mat* a, const double* x_in, double* x_out, const double* b, double w,
                int forward, int ultra, int nseg, const int64_t* seg);
void or_sor_f32(const or_mat* a, const float* x_in, float* x_out, const float* b, float w,
                int forward, int ultra, int nseg, const int64_t* seg);

/* SSOR: `iters` times a forward sweep into an intermediate vector and a
 * backward sweep from it (gauss-seidel-variant/solver.hpp:45-96), one segment;
 * x_in may be NULL (= 0) and may equal x_out */
void or_ssor_f64(const or_mat* a, const double* x_in, double* x_out, const double* b, double w,
                 int ultra, int iters);

/* sequential double accumulation of a[i]*b[i] */
double or_dot_f64(const double* a, const double* b, int64_t n);
double or_dot_f32(const float* a, const float* b, int64_t n);

/* CSR assembly from the stencil (tensor/dynamic/csr/create-from.hpp).
 * compress != 0 drops zero-diagonal rows (ZeroToZero handling) and fills cmap
 * (old index -> new index or -1).  Pass rowptr == NULL to query nnz/rows. */
int64_t or_csr_from_stencil(const or_stencil* s, int compress, int64_t* out_rows,
                            int64_t* rowptr, int64_t* cols, double* vals, double* rhs,
                            int64_t* cmap);

/* ---- AMG setup (hierarchy/wright.hpp, aggregation/, transport/matrix-shared.hpp) ---- */
typedef struct {
  int max_level_num;         /* 16 */
  int64_t min_coarsen_size;  /* 32 */
  double min_coarsen_factor; /* 1.2 */
  double strong_threshold;   /* 1/3 */
  double aggressive_strong_threshold;
  int64_t min_aggregate_size; /* 6 */
  int64_t max_aggregate_size; /* 9 */
  int coarse_f32;            /* coarse matrices stored in f32 (LoReal) */
} or_amg_params;

typedef struct or_hierarchy or_hierarchy;

or_hierarchy* or_amg_setup(const or_mat* finest, const or_amg_params* p);
/* strength graph of one level (dependency/criterion.hpp:35-78 with
 * graph/weight-transform.hpp:28-39): props[e] for the off-diagonal entries in
 * row order (bit 0 depends, bit 1 influences), iso[i] for each vertex;
 * returns the edge count (pass props == NULL to query it only) */
int64_t or_strength_graph(const or_mat* m, double theta, uint8_t* props, uint8_t* iso);
/* 1 if the coarsest level is solved by LU (nonsymmetric hierarchy) */
int or_amg_is_lu(const or_hierarchy* h);
/* dense row-wise LU without pivoting of a CSR matrix (unit lower factor) and
 * the matching forward/backward substitution */
int or_lu_decompose_dense(const or_csr* a, double* lo, double* up);
void or_lu_solve_dense(int64_t n, const double* lo, const double* up, const double* b,
                       double* x);
void or_amg_free(or_hierarchy* h);
int or_amg_level_num(const or_hierarchy* h);               /* including the finest */
int64_t or_amg_level_size(const or_hierarchy* h, int level);
/* aggregate map of level `level` (size = level size): coarse index, or
 * 0xFFFFFFFE (isolated); coarse->fine lists as (ptr, idx) */
const uint32_t* or_amg_agg_map(const or_hierarchy* h, int level);
const int64_t* or_amg_c2f_ptr(const or_hierarchy* h, int level);
const int64_t* or_amg_c2f_idx(const or_hierarchy* h, int level);
/* coarse matrix of level `level` (>= 1) */
const or_csr* or_amg_matrix(const or_hierarchy* h, int level);
/* lower Cholesky factor of the coarsest matrix (f64, CSR) */
const or_csr* or_amg_cholesky(const or_hierarchy* h);
/* aggregation statistics of level `level` (0..L-2): aggregates, multi, single,
 * merged, isolated */
void or_amg_stats(const or_hierarchy* h, int level, int64_t out[5]);

/* rows per thread segment of the smoothers on levels >= 1 (0: one segment) */
void or_amg_set_coarse_segment(or_hierarchy* h, int64_t rows);
/* rows per smoother segment of one coarse level (0: one segment), overriding
 * or_amg_set_coarse_segment there */
void or_amg_set_level_segment(or_hierarchy* h, int level, int64_t rows);
/* MultiGridSolver{relax, DefaultCycle{kind, iterations}} (multigrid/solver.hpp:183-186,
 * multigrid/cycle.hpp:24-79): kind 0 REGULAR, 1 DECREASING; iterations >= 1.
 * The default is relax 1 and DefaultCycle{REGULAR, 1} (a V-cycle). */
void or_amg_set_cycle(or_hierarchy* h, double relax, int kind, int iterations);

/* coarsest solve x = (L L^T)^{-1} b in f64 (direct/cholesky/solver.hpp:42-51) */
void or_cholesky_solve(const or_csr* lower, const double* b, double* x);
/* generic lower Cholesky (direct/cholesky/decomposition.hpp:44-99): output arrays
 * sized by the caller (nnz <= n*(n+1)/2).  Returns nnz or -1 if not SPD. */
int64_t or_cholesky_decompose(const or_csr* a, int64_t* rowptr, int64_t* cols, double* vals);

/* one cycle z = M r (multigrid/solver.hpp:56-176), x0 = 0, with the
 * hierarchy's relax and cycle (a V-cycle by default).  precision 64 or 32
 * selects the vector/smoother precision (32: FP32 hierarchy, FP64 coarsest). */
void or_vcycle_f64(const or_hierarchy* h, const double* r, double* z, int ultra);
void or_vcycle_f32(const or_hierarchy* h, const float* r, float* z, int ultra);
/* one cycle from the initial guess x_in (MultiGridSolver::Instance::apply,
 * multigrid/solver.hpp:41-53): the finest pre-smoothing starts from x_in
 * (NULL: zero, the same as or_vcycle) */
void or_cycle_f64(const or_hierarchy* h, const double* x_in, const double* r, double* z, int ultra);
void or_cycle_f32(const or_hierarchy* h, const float* x_in, const float* r, float* z, int ultra);

/* Krylov drivers (conjugate-gradients/solver.hpp, bicgstab/solver.hpp) with the
 * true residual recomputed at iterations 8, 16, ... and stopping on
 * ||r_k|| <= rel_tol ||r_0|| or max_it.  precon: 0 none, 1 AMG FP64, 2 AMG mixed
 * (FP32 V-cycle), 3 one SSOR step with omega = 1 (h may be NULL).  Returns
 * iterations; res_hist (max_it+1 entries) may be NULL. */
int or_pcg(const or_mat* a, const or_hierarchy* h, int precon, int ultra, const double* b,
           double* x, double rel_tol, int max_it, double* res_hist);
int or_pbicgstab(const or_mat* a, const or_hierarchy* h, int precon, int ultra,
                 const double* b, double* x, double rel_tol, int max_it, double* res_hist);

/* Outlet flux J = sum_{x = nx-1} (2 q_0 c_i)(p_i - p_end) of a per-cell
 * coefficient stencil, the adjoint right-hand side dJ/dp, and the gradient
 * dJ/dlog c given the forward solution p and the adjoint solution lambda
 * (not in the reference; derivation in oracle.c). */
double or_outlet_flux(const or_stencil* s, const double* p);
void or_outlet_flux_adjoint_rhs(const or_stencil* s, double* g);
void or_outlet_flux_gradient(const or_stencil* s, const double* p, const double* lam,
                             double* grad);

/* Effective diffusivity through the mid plane x = nx / 2 (EffDiffCalc,
 * component-wise/sink/effective-diffusion.hpp:229-327; needs nx >= 3 so
 * that the plane after the mid plane exists, NaN otherwise).
 * none_mats: 4 flags or NULL; cmap: the compress map of a ZeroToZero CSR
 * system (sol indexed by its rows) or NULL (sol indexed by cells). */
double or_eff_diff_area_part(const or_stencil* s, const uint8_t* none_mats);
double or_eff_diff(const or_stencil* s, const int64_t* cmap, const double* sol,
                   double area_part);

/* CSR with B x B block values (B <= 4, B*B row-major entries per block):
 * y = A x accumulated in f64 (tensor/fixed/compat.hpp:24-33,
 * component-wise/matrix-vector-product.hpp:71-82); the f32 variant rounds the
 * block values to f32 and stores y in f32 */
void or_bcsr_spmv_f64(int64_t nb, int bsz, const int64_t* rowptr, const int64_t* cols,
                      const double* vals, const double* x, double* y);
void or_bcsr_spmv_f32(int64_t nb, int bsz, const int64_t* rowptr, const int64_t* cols,
                      const double* vals, const float* x, float* y);

#ifdef __cplusplus
}
#endif
#endif
