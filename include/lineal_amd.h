/*
 * lineal_amd — C ABI of the MI355X (gfx950) implementation of lineal's
 * solve-phase hot path.  Plain pointers and sizes only; every function returns
 * an lgpu_status and no exception crosses the boundary.  Host arrays passed in
 * are borrowed for the duration of the call (copied to the device); the library
 * owns the device memory of the objects it creates.  All work of one context
 * is ordered on that context's HIP stream; a call returns after its work has
 * been enqueued unless it documents a host synchronisation.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference's include/lineal/).
 */
#ifndef LINEAL_AMD_H
#define LINEAL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  LGPU_OK = 0,
  LGPU_INVALID_ARG = 1,
  LGPU_OUT_OF_MEMORY = 2,
  LGPU_HIP_ERROR = 3,
  LGPU_BREAKDOWN = 4,      /* zero denominator in a Krylov recurrence */
  LGPU_NONFINITE = 5,      /* NaN/Inf residual */
  LGPU_NOT_CONVERGED = 6,  /* max_it reached before rel_tol */
  LGPU_NOT_SUPPORTED = 7,
  LGPU_SYNC_TIMEOUT = 8    /* a wavefront kernel's bounded wait expired */
} lgpu_status;

typedef enum { LGPU_F64 = 0, LGPU_F32 = 1 } lgpu_dtype;

typedef struct lgpu_ctx_s* lgpu_ctx;
typedef struct lgpu_vec_s* lgpu_vec;
typedef struct lgpu_op_s* lgpu_op;
typedef struct lgpu_amg_s* lgpu_amg;
typedef struct lgpu_bcsr_s* lgpu_bcsr;

/* ---- errors and context ------------------------------------------------ */
const char* lgpu_status_string(int status);
/* message of the last failing call on this host thread */
const char* lgpu_last_error(void);

/* One context per host thread/stream.  Replaces the execution policy of the
 * reference Env (parallel/execution-policy/adaptive.hpp:146-349): work that the
 * reference forks over a thread pool is launched on the context's stream. */
int lgpu_ctx_create(int device, lgpu_ctx* out);
int lgpu_ctx_destroy(lgpu_ctx ctx);
int lgpu_ctx_synchronize(lgpu_ctx ctx);
/* HIP events on the context stream: elapsed milliseconds between start and stop */
int lgpu_timer_start(lgpu_ctx ctx);
int lgpu_timer_stop(lgpu_ctx ctx, double* ms);
/* The context's sticky wait-error word (no counterpart in the reference:
 * the reference's sweeps run on host threads and cannot time out).  The
 * wavefront sweeps (SOR on a stencil, one-segment SOR on a CSR operator, and
 * the V-cycle's sweeps) wait on each other through bounded waits; one that
 * expires sets this word, after which every wait of the context's sweeps
 * gives up at once, so that every sweep still ends.  The entry points that
 * run such sweeps (lgpu_op_sor, lgpu_op_ssor, lgpu_amg_vcycle, the Krylov
 * solves) then return LGPU_SYNC_TIMEOUT and clear the word; the output of a
 * sweep whose waits gave up is undefined.  lgpu_ctx_set_wait_error sets the
 * word from the host, for tests of that path. */
int lgpu_ctx_set_wait_error(lgpu_ctx ctx);
/* Kernel-class timing (no counterpart in the reference; used by the
 * benchmark to time the dominant kernel inside a timed solve).  Every launch
 * of a class whose bit (1 << class) is set in the mask is bracketed by HIP
 * events on the context's stream; each event pair can add a few microseconds
 * between kernels, so time only the classes needed.  Setting a mask (0
 * disables) clears the records. */
enum {
  LGPU_KT_STENCIL_SOR_FWD = 0, /* forward SOR sweeps on a stencil operator */
  LGPU_KT_STENCIL_SOR_BWD = 1, /* backward SOR sweeps on a stencil operator */
  LGPU_KT_STENCIL_PASS = 2,    /* stencil SpMV / residual / Jacobi passes */
  LGPU_KT_CSR_SOR = 3,         /* CSR SOR sweeps (coarse levels included) */
  LGPU_KT_CSR_PASS = 4,        /* CSR SpMV / residual passes */
  LGPU_KT_COUNT = 5,
  LGPU_KT_ALL = (1 << 5) - 1
};
int lgpu_ctx_kernel_timing(lgpu_ctx ctx, int class_mask);
/* total milliseconds and launch count of a class since timing was enabled */
int lgpu_ctx_kernel_time(lgpu_ctx ctx, int kernel_class, double* total_ms, int64_t* launches);

/* ---- vectors: DenseVector (tensor/dynamic/dense/vector/shared.hpp:35-362) --- */
int lgpu_vec_create(lgpu_ctx ctx, int64_t n, int dtype, lgpu_vec* out);
/* borrow existing device memory (not freed by lgpu_vec_destroy) */
int lgpu_vec_wrap(lgpu_ctx ctx, void* device_ptr, int64_t n, int dtype, lgpu_vec* out);
int lgpu_vec_destroy(lgpu_vec v);
int lgpu_vec_size(lgpu_vec v, int64_t* n);
void* lgpu_vec_data(lgpu_vec v);
/* host <-> device; host array has the vector's dtype; synchronous */
int lgpu_vec_upload(lgpu_vec v, const void* host, int64_t n);
int lgpu_vec_download(lgpu_vec v, void* host, int64_t n);
/* assign(v, constant_like(v, value)) */
int lgpu_vec_fill(lgpu_vec v, double value);
/* assign(dst, src) with a precision conversion when the dtypes differ */
int lgpu_vec_copy(lgpu_vec src, lgpu_vec dst);
/* dot<f64>(a, b), component-wise/sink/accumulator.hpp:175-190 (synchronous) */
int lgpu_vec_dot(lgpu_vec a, lgpu_vec b, double* out);

/* ---- operators ----------------------------------------------------------- */
/* Box geometry and Dirichlet data, discretization/box-equidistant.hpp:57-111 and
 * tensor/dynamic/stencil/matrix/diffusion.hpp:52-96 (flow axis 0). */
typedef struct {
  int64_t n[3];        /* cells per axis, index = (x*ny + y)*nz + z */
  double len[3];       /* axis lengths */
  int zero_to_one;     /* ZeroToOne diagonal handling (1) or ZeroToZero (0) */
  double p_start;      /* solution at the x = 0 border */
  double p_end;        /* solution at the x = nx border */
  double velocity_x;   /* first-order upwind advection along +x (>= 0; 0: pure
                          diffusion).  Not in the reference: each cell's +x face
                          carries the outflow v*h_y*h_z (diagonal, after the
                          diffusion entries), the -x coupling gains -v*h_y*h_z and
                          the x = 0 rhs the inflow v*h_y*h_z*p_start. */
} lgpu_box;

/* AdjacentStencilMatrix over StationaryDiffusionValuesValuator
 * (tensor/dynamic/stencil/valuator/stationary-diffusion-values.hpp:31-211),
 * one per-cell coefficient for all axes.  coeff: host array of n[0]*n[1]*n[2]. */
int lgpu_stencil_values_create(lgpu_ctx ctx, const lgpu_box* box, const double* coeff,
                               lgpu_op* out);
/* AdjacentStencilMatrix over StationaryDiffusionLookupValuator
 * (tensor/dynamic/stencil/valuator/stationary-diffusion-lookup.hpp:29-151):
 * u8 material ids (< 4) and the material coefficient table. */
int lgpu_stencil_lookup_create(lgpu_ctx ctx, const lgpu_box* box, const uint8_t* materials,
                               const double mat_coeffs[4], lgpu_op* out);
/* CsrMatrix (tensor/dynamic/csr/shared.hpp:37-1103): sorted columns, diagonal
 * stored; values given in f64 and stored in `dtype`. */
int lgpu_csr_create(lgpu_ctx ctx, int64_t n, const int64_t* rowptr, const int32_t* cols,
                    const double* vals, int dtype, int symmetric, lgpu_op* out);
/* csr_from_adjacent_stencil (tensor/dynamic/csr/create-from.hpp:37-239) without
 * compression: CSR copy of a stencil operator, values stored in `dtype`. */
int lgpu_csr_from_stencil(lgpu_op stencil, int dtype, lgpu_op* out);
/* csr_from_adjacent_stencil<Mat, Vec> (tensor/dynamic/csr/create-from.hpp:37-239):
 * the whole system { lhs, rhs, compress_map }.  With compress != 0 the
 * ZeroToZero compression applies (create-from.hpp:54-55, 92-99, 197-229): rows
 * whose diagonal is zero are dropped and the others renumbered in order.
 * *lhs receives the CSR operator (values in `dtype`), *rhs a new vector of its
 * row count holding the Dirichlet right-hand side in `dtype`, and cmap (host,
 * n[0]*n[1]*n[2] entries, may be NULL) the row of every cell or -1.  A
 * compressed operator keeps its map for lgpu_eff_diff. */
int lgpu_csr_system_from_stencil(lgpu_op stencil, int dtype, int compress, lgpu_op* lhs,
                                 lgpu_vec* rhs, int64_t* cmap);
int lgpu_op_destroy(lgpu_op op);
int lgpu_op_rows(lgpu_op op, int64_t* n);

/* AdjacentStencilVector (tensor/dynamic/stencil/vector/shared.hpp:181-205):
 * Dirichlet right-hand side of a stencil operator, in b's dtype. */
int lgpu_op_rhs(lgpu_op op, lgpu_vec b);
/* y = A x, MatrixVectorProductExpr (component-wise/matrix-vector-product.hpp:53-84);
 * x and y share a dtype, which selects the arithmetic precision. */
int lgpu_op_spmv(lgpu_op op, lgpu_vec x, lgpu_vec y);
/* r = b - A x, subtract<Real>(rhs, multiply<Real>(lhs, x)) */
int lgpu_op_residual(lgpu_op op, lgpu_vec b, lgpu_vec x, lgpu_vec r);
/* One Jacobi sweep x_in -> x_out, JacobiSweepSink
 * (linear-solver/iterative-stationary/jacobi/sweep.hpp:66-87); x_in may be NULL (= 0). */
int lgpu_op_jacobi(lgpu_op op, lgpu_vec x_in, lgpu_vec x_out, lgpu_vec b, double omega);
/* One SOR sweep x_in -> x_out in lexicographic order (one thread segment),
 * SorSweepSink (linear-solver/iterative-stationary/gauss-seidel-variant/sweep.hpp:134-200);
 * forward 1/0, ultra 0 (regular) / 1 (l1 variant); x_in may be NULL (= 0) but
 * must not alias x_out.  Synchronises the context's stream to read the
 * wait-error word (LGPU_SYNC_TIMEOUT, see lgpu_ctx_set_wait_error); so do
 * lgpu_op_ssor and lgpu_amg_vcycle. */
int lgpu_op_sor(lgpu_op op, lgpu_vec x_in, lgpu_vec x_out, lgpu_vec b, double omega,
                int forward, int ultra);
/* Coefficient cache of the stencil SOR sweeps.  By default a stencil
 * operator's sweeps read its face weights and diagonal from a per-cell cache
 * (four values per cell in the sweep's precision, built by the first sweep of
 * that precision, kept with the operator; an AMG hierarchy builds its finest
 * level's at creation).  enable = 0 frees the caches and makes the sweeps
 * evaluate the weights on the fly; both ways give the same bits.  No
 * counterpart in the reference, whose sweeps evaluate the stencil on the fly
 * (tensor/dynamic/stencil/matrix/algorithm.hpp:91-227); for callers that
 * trade the memory against sweep speed.  No effect on CSR operators. */
int lgpu_op_set_sor_cache(lgpu_op op, int enable);
/* One SSOR step, SsorBroom::sweep
 * (linear-solver/iterative-stationary/gauss-seidel-variant/solver.hpp:45-88):
 * a forward sweep x_in -> aux, then a backward sweep aux -> x_out, both
 * lexicographic; x_in may be NULL (= 0) and may be x_out. */
int lgpu_op_ssor(lgpu_op op, lgpu_vec x_in, lgpu_vec x_out, lgpu_vec b, double omega, int ultra);
/* One SOR sweep of a CSR operator with the reference's thread-segment
 * semantics (gauss-seidel-variant/sweep.hpp:153-169; the segments of
 * parallel/execution-policy/adaptive.hpp:266-279): rows are cut into segments
 * of segment_rows (1..8192) consecutive rows, each swept in order with new
 * values only from its own segment. */
int lgpu_op_sor_segmented(lgpu_op op, lgpu_vec x_in, lgpu_vec x_out, lgpu_vec b, double omega,
                          int forward, int ultra, int64_t segment_rows);

/* ---- CSR with block values ------------------------------------------------
 * CsrMatrix over B x B block values, fix::DenseMatrix<Real, B, B>
 * (tensor/dynamic/csr/shared.hpp:37-1103, tensor/fixed/matrix.hpp:151; the
 * reference's test/shared-blocked.cpp uses B = 3 and 2).  block_rows block
 * rows, rowptr (block_rows + 1) and cols (block columns, any order) per block
 * entry, vals with B * B row-major entries per block given in f64 and stored
 * in `dtype`; B in 1..4.  Vectors have block_rows * B entries of that dtype. */
int lgpu_bcsr_create(lgpu_ctx ctx, int64_t block_rows, int block_size, const int64_t* rowptr,
                     const int32_t* cols, const double* vals, int dtype, lgpu_bcsr* out);
int lgpu_bcsr_destroy(lgpu_bcsr m);
int lgpu_bcsr_rows(lgpu_bcsr m, int64_t* block_rows, int* block_size);
/* y = A x, MatrixVectorProductExpr's row path
 * (component-wise/matrix-vector-product.hpp:71-82): per block row,
 * sum = compat::fmadd(A_ij, x_j, sum) = A_ij x_j + sum over the row's blocks
 * in stored order (tensor/fixed/compat.hpp:24-33), component r of A_ij x_j
 * being 0 + a_r0 x_0 + a_r1 x_1 + ... (tensor/fixed/expressions.hpp:22-44);
 * all in f64, the sums stored in the vectors' type. */
int lgpu_bcsr_spmv(lgpu_bcsr m, lgpu_vec x, lgpu_vec y);

/* ---- aggregation AMG ----------------------------------------------------- */
/* HierarchyWrightParams + AggregationParams + DuneDependencyCriterion
 * (multigrid/hierarchy/wright.hpp:46-55, aggregation/params.hpp, test/aux/default.hpp:170-188) */
typedef struct {
  int max_level_num;            /* 16 */
  int64_t min_coarsen_size;     /* 32 */
  double min_coarsen_factor;    /* 1.2 */
  double strong_threshold;      /* 1/3 */
  double aggressive_strong_threshold; /* 1/3 */
  int64_t min_aggregate_size;   /* 6 */
  int64_t max_aggregate_size;   /* 9 */
  int precision;                /* LGPU_F64: FP64 V-cycle; LGPU_F32: FP32 hierarchy,
                                   smoothers and level operators, FP64 coarsest solve */
  int ultra;                    /* SOR variant of the smoothers */
  int64_t coarse_segment_rows;  /* coarse-level (CSR) smoothers: rows per thread segment
                                   (default 64); 0 sweeps each coarse level in one
                                   segment (lexicographic). The finest level is always
                                   swept lexicographically. */
  /* MultiGridSolver{relax, DefaultCycle{cycle_kind, cycle_iterations}, wright}
   * (multigrid/solver.hpp:183-186, multigrid/cycle.hpp:24-79): the coarse
   * correction is added scaled by 1/relax (default 1); cycle_kind
   * LGPU_CYCLE_REGULAR or LGPU_CYCLE_DECREASING, cycle_iterations >= 1
   * (default REGULAR, 1: the V-cycle) */
  double relax;
  int cycle_kind;
  int cycle_iterations;
  /* where the order-independent setup phases run: 1 (default) the strength
   * criterion and the Galerkin products on the GPU, 0 on host threads; the
   * greedy aggregation runs on the host either way, and the hierarchy is the
   * same bit for bit */
  int setup_on_device;
  /* coarse levels with at most this many rows are swept in one segment
   * (lexicographically, as the reference with one thread) whatever
   * coarse_segment_rows says; default 512, 0: none.  Levels of at most 512
   * rows run a round program (one dependency level of rows per round, in one
   * workgroup); larger ones the scheduled sweep.  Fewer CG iterations
   * (256^3 lognormal: 98 -> 94 with the default) */
  int64_t sequential_rows;
} lgpu_amg_params;
enum { LGPU_CYCLE_REGULAR = 0, LGPU_CYCLE_DECREASING = 1 };

void lgpu_amg_default_params(lgpu_amg_params* p);
/* MultiGridSolver::instantiate -> HierarchyWrightInstance::create
 * (multigrid/solver.hpp:188-196, hierarchy/wright.hpp:145-309): the hierarchy
 * (strength criterion, aggregation with one thread segment, Galerkin products,
 * coarsest factor) is built on the host and uploaded once.  The coarsest level
 * is solved by Cholesky (direct/cholesky/decomposition.hpp:44-99) for a
 * symmetric operator and by row-wise LU without pivoting
 * (direct/lu/decomposition.hpp:30-111) for a nonsymmetric one (upwind
 * advection, the BiCgStab preconditioner). */
int lgpu_amg_create(lgpu_op finest, const lgpu_amg_params* params, lgpu_amg* out);
/* lgpu_op_ssor for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same sweeps and the same result bit for
 * bit; the forward sweep's output, which lgpu_op_ssor keeps in a buffer the
 * library allocates on the operator's first SSOR call, lives in a static
 * per-GPU device scratch of 2^22 f64 (32 MiB), so operators with more than
 * 4194304 rows are rejected (LGPU_INVALID_ARG).  Nothing is allocated and no
 * kernel-class timing is recorded.  The sweeps read the operator's SOR
 * coefficient cache when it exists and evaluate the stencil otherwise (both
 * give the same bits; building the cache would allocate).  f32 sweeps on a per-cell coefficient
 * stencil need the operator's f32 coefficient copy, which the first f32 pass
 * of another entry point makes (otherwise LGPU_INVALID_ARG).  Every call on
 * one GPU shares that GPU's scratch: calls that overlap on one GPU (other
 * contexts or host threads) race.  Synchronises the stream like lgpu_op_ssor. */
int lgpu_op_ssor_noalloc(lgpu_op op, lgpu_vec x_in, lgpu_vec x_out, lgpu_vec b, double omega,
                         int ultra);
/* One transfer of a hierarchy built elsewhere: the aggregate map of a level
 * (n fine rows; each row's aggregate, or 0xFFFFFFFE for an isolated row), the
 * coarse -> fine lists that fix the order of the restriction sums
 * (transport/vector-shared.hpp:49-56; c2f_ptr with coarse_n + 1 entries and
 * c2f_idx, or both NULL for ascending fine rows), and the coarse level's
 * matrix (CSR, sorted columns, diagonal stored, f64 values). */
typedef struct {
  const uint32_t* agg;
  const int64_t* c2f_ptr;
  const int32_t* c2f_idx;
  int64_t coarse_n;
  const int64_t* rowptr;
  const int32_t* cols;
  const double* vals;
} lgpu_amg_level;
/* MultiGridSolver over a hierarchy built elsewhere, e.g. by the reference's
 * own host setup (HierarchyWrightInstance::create, hierarchy/wright.hpp:145-309):
 * transfer_num transfers, levels[l] mapping level l (level 0: `finest`) to
 * level l + 1.  The library factors the coarsest matrix (Cholesky, or LU for a
 * nonsymmetric finest operator) and uploads everything; params supply the
 * precision (coarse values rounded to f32 for an f32 hierarchy), the
 * smoother variant, the coarse segments, relax and the cycle.  The maps and
 * matrices are validated (ranges, consistency of the lists with the map,
 * sorted columns with a diagonal). */
int lgpu_amg_create_from_levels(lgpu_op finest, const lgpu_amg_params* params, int transfer_num,
                                const lgpu_amg_level* levels, lgpu_amg* out);
/* host copy of transfer `level`'s coarse -> fine lists (level + 1 < level
 * count; c2f_ptr: coarse rows + 1 entries, c2f_idx: its last entry) */
int lgpu_amg_transfer(lgpu_amg amg, int level, int64_t* c2f_ptr, int32_t* c2f_idx);
/* SsorSolver<Real, FixedIterationManager, variant>{omega, IterMan{iterations}}
 * as a Krylov preconditioner (gauss-seidel-variant/solver.hpp:94-96,
 * broom.hpp:159-171; test/aux/iterative.hpp:143-165): a preconditioner handle
 * of the lgpu_amg type without a hierarchy.  Applying it runs `iterations`
 * SSOR steps from a zero initial guess with the Krylov operator; it has no
 * levels, and lgpu_amg_vcycle applies it to f64 vectors. */
int lgpu_ssor_precon_create(lgpu_op op, double omega, int ultra, int iterations, lgpu_amg* out);
int lgpu_amg_destroy(lgpu_amg amg);
int lgpu_amg_level_num(lgpu_amg amg, int* levels);
int lgpu_amg_level_size(lgpu_amg amg, int level, int64_t* n);
/* host copy of level `level`'s aggregate map (0xFFFFFFFE = isolated) */
int lgpu_amg_aggregates(lgpu_amg amg, int level, uint32_t* host_map);
/* host copy of coarse level `level`'s CSR (level >= 1); pass NULL arrays to
 * query nnz.  Values are returned in f64. */
int lgpu_amg_matrix(lgpu_amg amg, int level, int64_t* nnz, int64_t* rowptr, int32_t* cols,
                    double* vals);
/* seconds spent in the host setup */
int lgpu_amg_setup_seconds(lgpu_amg amg, double* seconds);
/* The host setup's wall seconds by phase (no counterpart in the reference):
 * [0] strength graph and criterion, [1] aggregation, [2] Galerkin products,
 * [3] coarsest factorisation, [4] the rest (row evaluation of the finest
 * operator, uploads). */
int lgpu_amg_setup_split(lgpu_amg amg, double seconds[5]);
/* z = M r, one cycle from zero (MultiGridSolver::Instance::apply with a
 * constant-zero in_sol, multigrid/solver.hpp:47-53); r and z are f64 (or
 * both f32 with an f32 hierarchy). */
int lgpu_amg_vcycle(lgpu_amg amg, lgpu_vec r, lgpu_vec z);
/* One cycle from the initial guess x_in: MultiGridSolver::Instance::apply
 * (multigrid/solver.hpp:41-53) with the finest pre-smoothing starting from
 * x_in instead of zero; x_in may be NULL (= lgpu_amg_vcycle) but must not
 * alias z.  All three vectors have the hierarchy's precision. */
int lgpu_amg_cycle(lgpu_amg amg, lgpu_vec x_in, lgpu_vec r, lgpu_vec z);

/* ---- Krylov solvers ------------------------------------------------------- */
typedef struct {
  double rel_tol;    /* stop when ||r_k||_2 <= rel_tol ||r_0||_2 */
  int max_it;
  int rere_every;    /* true residual every k-th iteration (8, test/aux/iterative.hpp:97-99); 0 never */
} lgpu_krylov_params;

/* run_iterative_solver (linear-solver/runner.hpp:17-84) over
 * ConjugateGradientsSolver (conjugate-gradients/solver.hpp:50-115), optionally
 * AMG-preconditioned.  The preconditioner may be built from another operator
 * of the same size and context (e.g. the diffusion part of an advection
 * operator).  x holds the initial guess and receives the solution (f64).
 * res_hist (max_it + 1 entries) may be NULL.  Returns LGPU_OK on convergence,
 * LGPU_NOT_CONVERGED, LGPU_BREAKDOWN or LGPU_NONFINITE. */
int lgpu_cg_solve(lgpu_op op, lgpu_amg precon, lgpu_vec b, lgpu_vec x,
                  const lgpu_krylov_params* params, int* iterations, double* res_hist);
/* BiCgStabSolver (bicgstab/solver.hpp:51-151), right-preconditioned. */
int lgpu_bicgstab_solve(lgpu_op op, lgpu_amg precon, lgpu_vec b, lgpu_vec x,
                        const lgpu_krylov_params* params, int* iterations, double* res_hist);

/* ---- Krylov system instances, one iteration per call ----------------------
 * ConjugateGradientsSolver / BiCgStabSolver ::Instance::instantiate(sol, rhs)
 * -> SystemInstance with iterate(env), residual_norm(), rho(), solution()
 * (conjugate-gradients/solver.hpp:38-198, bicgstab/solver.hpp:46-170), the
 * object run_iterative_solver hands to its iteration callback
 * (linear-solver/runner.hpp:52-68).  Creating one performs the
 * SystemInstance constructor's work (r = b - A x and its norm; CG: the
 * preconditioned residual and rho; BiCGStab: r0 = p = r, rho = ||r||^2) and
 * synchronises; so does every lgpu_krylov_iterate, which runs one iteration
 * with the true residual at iterations rere_every, 2 rere_every, ... (0:
 * never; test/aux/iterative.hpp:97-99).  x receives the iterates (the
 * solution() of the reference); b and x must outlive the handle.  Each
 * instance owns its work vectors and scalars, so several may be iterated in
 * turn; lgpu_cg_solve / lgpu_bicgstab_solve remain the fused loops. */
typedef struct lgpu_krylov_s* lgpu_krylov;
int lgpu_cg_create(lgpu_op op, lgpu_amg precon, lgpu_vec b, lgpu_vec x, int rere_every,
                   lgpu_krylov* out);
int lgpu_bicgstab_create(lgpu_op op, lgpu_amg precon, lgpu_vec b, lgpu_vec x, int rere_every,
                         lgpu_krylov* out);
/* SystemInstance::iterate (conjugate-gradients/solver.hpp:75-115,
 * bicgstab/solver.hpp:71-151): one iteration; returns LGPU_SYNC_TIMEOUT if a
 * wavefront sweep's wait expired */
int lgpu_krylov_iterate(lgpu_krylov k);
/* SystemInstance::residual_norm() = sqrt(||r||^2) and rho() of the last
 * iteration (or of the construction) */
int lgpu_krylov_residual_norm(lgpu_krylov k, double* out);
int lgpu_krylov_rho(lgpu_krylov k, double* out);
/* iterations performed since creation */
int lgpu_krylov_iteration(lgpu_krylov k, int* out);
int lgpu_krylov_destroy(lgpu_krylov k);

/* ---- matrix and vector files (io/tensor.hpp:25-263) -----------------------
 * SMF: u64 rows, u64 columns, u64 nonzeros, u64 row offsets[rows + 1], then
 * (u64 column, f64 value) pairs row by row.  DVF: u64 size, f64 values.
 * Little-endian.  Paths are host file names. */

/* SmfReader::row_num / col_num / nnz_num (io/tensor.hpp:35-46) */
int lgpu_smf_info(const char* path, int64_t* rows, int64_t* cols, int64_t* nnz);
/* SmfReader::read (io/tensor.hpp:47-75): a CSR operator, duplicate columns of a
 * row summed in f64 in file order and the row sorted by column (the row-wise
 * builder, tensor/dynamic/csr/shared.hpp:730-764), values stored in `dtype`.
 * The reference builds the matrix as nonsymmetric; symmetric != 0 declares it
 * symmetric instead (an AMG hierarchy then uses the symmetric strength
 * criterion and a Cholesky coarsest solve).  Needs a square matrix that
 * stores every diagonal entry. */
int lgpu_smf_read(lgpu_ctx ctx, const char* path, int dtype, int symmetric, lgpu_op* out);
/* SmfWriter::write (io/tensor.hpp:158-213) of a CSR or stencil operator: the
 * nonzero entries in column order; the header's nonzero count is the stored
 * entry count for CSR operators (CsrMatrix::non_zero_num) and the written
 * count for stencils. */
int lgpu_smf_write(lgpu_op op, const char* path);
/* DvfReader::size / read (io/tensor.hpp:215-233): a new vector in `dtype` */
int lgpu_dvf_size(const char* path, int64_t* n);
int lgpu_dvf_read(lgpu_ctx ctx, const char* path, int dtype, lgpu_vec* out);
/* DvfWriter::write (io/tensor.hpp:253-263): values written as f64 */
int lgpu_dvf_write(lgpu_vec v, const char* path);

/* ---- effective diffusivity (component-wise/sink/effective-diffusion.hpp) ----
 * Both need a stencil operator with n[0] >= 3: the mid plane x = n[0] / 2 and
 * the plane after it must exist. */

/* EffDiffCalc::area_part (effective-diffusion.hpp:230-250; AreaPartSink
 * :49-121): (number of filled cells among the mid-plane cells and their +x
 * neighbours) / (n[1] n[2]) * 0.5.  none_mats: 4 flags, nonzero marks a
 * "none" material (LookupSymmetricDiffusionInfo::none_materials,
 * tensor/dynamic/stencil/matrix/diffusion.hpp:66-81; the filled lookup
 * of tensor/dynamic/stencil/valuator/lookup-wright.hpp:98-106); NULL for none.  A
 * per-cell coefficient operator has no materials and counts every cell. */
int lgpu_eff_diff_area_part(lgpu_op stencil, const uint8_t* none_mats, double* area_part);
/* EffDiffCalc::operator() (effective-diffusion.hpp:298-327; EffDiffSink
 * :127-222): S = sum over mid-plane cells i of D(i, i') (sol[i'] - sol[i]),
 * i' = i + n[1] n[2], D the harmonic mean of the two coefficients (lookup:
 * H(C[m_i], C[m_i'])); *eff_diff = S q_0 / (area_part L_1 L_2) L_0 / (p_end - p_start).
 * map: NULL when sol is indexed by cells, else a compressed system operator
 * from lgpu_csr_system_from_stencil whose compress map indexes sol (pairs
 * with a dropped cell are skipped).  sol is f64 or f32 (differences taken in
 * its type, products and the sum in f64). */
int lgpu_eff_diff(lgpu_op stencil, lgpu_op map, lgpu_vec sol, double area_part, double* eff_diff);

/* ---- adjoint of the solve (no counterpart in the reference) ---------------
 * Objective: the total outlet flux of a per-cell coefficient stencil without
 * advection,
 *   J(p, c) = sum over the x = nx-1 cells of (2 q_0 c_i)(p_i - p_end),
 * built from the reference's Dirichlet border summands
 * (tensor/dynamic/stencil/valuator/stationary-diffusion-values.hpp:124-138),
 * the nearest reference functional being EffDiffCalc's mid-plane flux
 * (component-wise/sink/effective-diffusion.hpp:228-335).  Parameters: the
 * per-cell log coefficients theta_i = log c_i.  All vectors are f64. */

/* *flux = J(p, c) */
int lgpu_outlet_flux(lgpu_op op, lgpu_vec p, double* flux);
/* g = dJ/dp: 2 q_0 c_i on the x = nx-1 cells, 0 elsewhere */
int lgpu_outlet_flux_adjoint_rhs(lgpu_op op, lgpu_vec g);
/* grad = dJ/dtheta given the forward solution p (A p = b) and the adjoint
 * solution lambda (A lambda = dJ/dp), one fused pass over the faces:
 *   grad_k = c_k ([x = nx-1] 2 q_0 (p_k - p_end)
 *                 + lambda_k 2 q_0 ([x = 0](p_start - p_k) + [x = nx-1](p_end - p_k)))
 *          - sum over faces (k, j) H(c_k, c_j) q_d c_j / (c_k + c_j) (lambda_k - lambda_j)(p_k - p_j) */
int lgpu_outlet_flux_gradient(lgpu_op op, lgpu_vec p, lgpu_vec lambda, lgpu_vec grad);
/* Forward solve A p = b, adjoint solve A lambda = dJ/dp with the same
 * (optional) AMG preconditioner and CG parameters, then the gradient.  p and
 * lambda hold the initial guesses and receive the solutions; *flux receives
 * J(p, c); iterations[0] / iterations[1] the forward / adjoint CG counts (may
 * be NULL).  Returns the first non-OK status of the two solves. */
int lgpu_adjoint_grad(lgpu_op op, lgpu_amg precon, const lgpu_krylov_params* params, lgpu_vec p,
                      lgpu_vec lambda, lgpu_vec grad, double* flux, int* iterations);

#ifdef __cplusplus
}
#endif
#endif
