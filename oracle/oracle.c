/*
 * CPU oracle for the lineal solve-phase hot path.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
 * never by the product.  Compiled with -ffp-contract=off so that only the
 * fma() calls written below are fused.
 *
 * How each part of the restatement is pinned (tests/test_oracle.py):
 *  - geometry, harmonic mean, stencil rows, rhs: closed forms (H(a,a)=a,
 *    H(0,b)=0), an independent dense numpy assembly of the python reference's
 *    diffusion_matrix semantics, symmetry, and the exact linear solution for
 *    fields constant along x (test/shared-stencil.hpp:226-239).
 *  - SpMV, Jacobi, SOR: dense numpy brute force on small grids (the python
 *    reference's sor() semantics for lexicographic SOR), fixed points, the
 *    SOR-on-lower-triangular exactness invariant, and the 4x8x4 cylinder-gap
 *    SOR scenario (test/python-reference/sor.py).
 *  - CSR assembly: equals the dense operator; compressed rows match ZeroToZero.
 *  - Cholesky: scipy/numpy Cholesky of the 32x32 tridiagonal KAT
 *    (test/python-reference/cholesky.py) and of coarsest AMG matrices.
 *  - Galerkin / restriction / prolongation: P^T A P and P^T r with dense numpy
 *    products for the oracle's own aggregate maps.
 *  - Aggregation and strength criterion: aggregates traced by hand from the
 *    grow / round / single-vertex rules on chains, grids, a pendant vertex and
 *    a weak coupling (tests/aggregation_cases.py: ranking keys, ties,
 *    insertion order, merge); the nonsymmetric criterion by a hand-traced
 *    strength graph (reverse-entry weights, missing reverse entries, negative
 *    weights, the head's diagonal, one-way and influences-only edges, an
 *    isolated vertex) and nonsymmetric aggregates traced by hand (one-way
 *    ranking in grow, two-way candidates displacing one-way ones, the round
 *    rule counting dependencies only, the merge through an influences-only
 *    coupling); invariants on stencil hierarchies (every vertex in one
 *    aggregate or isolated, sizes <= max, strong-edge rules); the reference's
 *    convergence thresholds.  Maps of large graphs equal the reference's by
 *    construction only: the reference cannot be built here.
 *  - Galerkin products of nonsymmetric (upwind advection) levels: dense
 *    P^T A P on every level, with A_c != (P^T A^T P).
 *  - CG / BiCGStab / PCG-AMG: exact termination on tiny SPD systems, the
 *    reference's thresholds (test/shared-stencil.hpp:157-224), scipy solves.
 *  - Effective diffusivity: the reference's known answer for diagonal lines
 *    (test/aux/materials.hpp:235-248 via test/shared-stencil.hpp:226-239) on
 *    exact and solved fields, with anisotropic lengths and reversed borders;
 *    a brute-force count of the area part with "none" materials; and the
 *    compressed 4x7x6 system of test/shared-csr.cpp:193-252.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_UNAGGREGATED 0xFFFFFFFFu
#define OR_ISOLATED 0xFFFFFFFEu

struct or_hierarchy {
  int nlev;
  or_mat finest;
  int64_t size[32];
  uint32_t* agg[32];     /* level l -> l+1 map, size[l] entries */
  int64_t* c2f_ptr[32];  /* size[l+1]+1 */
  int64_t* c2f_idx[32];
  or_csr mats[32];       /* mats[l] for l >= 1 */
  int64_t stats[32][5];
  or_csr chol;
  int lu;             /* nonsymmetric coarsest matrix: dense LU factors instead of Cholesky */
  double* lu_l;       /* unit lower factor, dense row major */
  double* lu_u;       /* upper factor, dense row major */
  int64_t coarse_seg; /* rows per smoother segment on levels >= 1 (0: one segment) */
  /* per-level segment lengths overriding coarse_seg where level_seg_set[l] */
  int64_t level_seg[32];
  int level_seg_set[32];
  /* MultiGridSolver{relax, DefaultCycle{kind, iterations}}: 0 in relax means
   * 1 (the calloc'ed default), kind 0 REGULAR / 1 DECREASING, 0 in
   * iterations means 1 */
  double relax;
  int cycle_kind, cycle_iters;
  /* the coarse levels' right-hand sides and solutions, owned by the
   * hierarchy and kept between cycle applications as the reference's
   * hierarchy vectors are (hierarchy/wright.hpp:227-228); zero at setup
   * (the reference leaves them uninitialised).  A DECREASING cycle reads a
   * coarse solution its coarser level computed in an earlier application.
   * One set per precision, allocated at setup and shared by preconditioner
   * views of the hierarchy. */
  double* rhs64[32];
  double* sol64[32];
  float* rhs32[32];
  float* sol32[32];
};

/* Thread segments of level l's smoother (parallel/execution-policy/adaptive.hpp
 * segments rows into contiguous ranges; gauss-seidel-variant/sweep.hpp:153-169
 * uses new values only inside the row's own segment). */
static int64_t* or_level_segments(const or_hierarchy* h, int l, int* nseg) {
  const int64_t n = h->size[l];
  const int64_t rows = h->level_seg_set[l] ? h->level_seg[l] : h->coarse_seg;
  const int64_t len = (l == 0 || rows <= 0) ? n : rows;
  const int64_t k = (n + len - 1) / len;
  int64_t* seg = (int64_t*)malloc(sizeof(int64_t) * (size_t)(k + 1));
  for (int64_t s = 0; s < k; ++s) seg[s] = s * len;
  seg[k] = n;
  *nseg = (int)k;
  return seg;
}

void or_amg_set_coarse_segment(or_hierarchy* h, int64_t rows) { h->coarse_seg = rows; }
void or_amg_set_level_segment(or_hierarchy* h, int level, int64_t rows) {
  if (level >= 1 && level < 32) {
    h->level_seg[level] = rows;
    h->level_seg_set[level] = 1;
  }
}
void or_amg_set_cycle(or_hierarchy* h, double relax, int kind, int iterations) {
  h->relax = relax;
  h->cycle_kind = kind;
  h->cycle_iters = iterations;
}

/* discretization/box-equidistant.hpp:62-79: h_d = L_d / n_d,
 * q_d = (prod_{e != d} h_e) / h_d with the product left-folded from 1 */
void or_axis_quotients(const or_stencil* s, double q[3]) {
  double h[3];
  for (int d = 0; d < 3; ++d) h[d] = s->len[d] / (double)s->n[d];
  for (int d = 0; d < 3; ++d) {
    double area = 1.0;
    for (int e = 0; e < 3; ++e)
      if (e != d) area *= h[e];
    q[d] = area / h[d];
  }
}

/* advective face coefficient v * h_y * h_z (0 without advection) */
static double or_advection(const or_stencil* s) {
  if (s->velocity == 0.0) return 0.0;
  return s->velocity * (s->len[1] / (double)s->n[1]) * (s->len[2] / (double)s->n[2]);
}

#define REAL double
#define SUF(name) name##_f64
#define FMA fma
#include "oracle_real.inc"
#undef REAL
#undef SUF
#undef FMA

#define REAL float
#define SUF(name) name##_f32
#define FMA fmaf
#include "oracle_real.inc"
#undef REAL
#undef SUF
#undef FMA

/* ---------------------------------------------------------------------------
 * CSR from stencil, tensor/dynamic/csr/create-from.hpp:37-239: rows visited in
 * order; an off-diagonal is kept iff its value is nonzero; with compression a
 * row with a zero diagonal is dropped and indices are renumbered; the
 * (unique) row entries are then sorted by column (csr/shared.hpp:751-763). */
int64_t or_csr_from_stencil(const or_stencil* s, int compress, int64_t* out_rows,
                            int64_t* rowptr, int64_t* cols, double* vals, double* rhs,
                            int64_t* cmap) {
  const int64_t n = s->n[0] * s->n[1] * s->n[2];
  int64_t c[8];
  double v[8];
  int64_t* map = cmap;
  int64_t* own = NULL;
  if (compress && !map) map = own = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  int64_t rows = 0, nnz = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int k = or_stencil_row_f64(s, i, c, v);
    const double d = v[k - 1];
    if (compress) {
      if (d == 0) {
        map[i] = -1;
        continue;
      }
      map[i] = rows;
    }
    ++rows;
    for (int e = 0; e < k; ++e)
      if (e == k - 1 || v[e] != 0) ++nnz;
  }
  if (out_rows) *out_rows = rows;
  if (!rowptr) {
    free(own);
    return nnz;
  }
  double* b = (double*)malloc(sizeof(double) * (size_t)n);
  or_stencil_rhs_f64(s, b);
  int64_t r = 0, p = 0;
  rowptr[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int k = or_stencil_row_f64(s, i, c, v);
    if (compress && map[i] < 0) continue;
    int64_t rc[8];
    double rv[8];
    int m = 0;
    for (int e = 0; e < k; ++e) {
      if (e != k - 1 && v[e] == 0) continue;
      rc[m] = compress ? map[c[e]] : c[e];
      rv[m++] = v[e];
    }
    /* insertion sort by column (columns are unique) */
    for (int a = 1; a < m; ++a) {
      int64_t kc = rc[a];
      double kv = rv[a];
      int j = a - 1;
      while (j >= 0 && rc[j] > kc) {
        rc[j + 1] = rc[j];
        rv[j + 1] = rv[j];
        --j;
      }
      rc[j + 1] = kc;
      rv[j + 1] = kv;
    }
    for (int e = 0; e < m; ++e) {
      cols[p] = rc[e];
      vals[p++] = rv[e];
    }
    if (rhs) rhs[r] = b[i];
    rowptr[++r] = p;
  }
  free(b);
  free(own);
  return nnz;
}

/* ---------------------------------------------------------------------------
 * AMG setup */

typedef struct {
  int64_t n;
  int64_t* eptr;  /* off-diagonal edges of row i: eptr[i]..eptr[i+1] */
  int64_t* ecol;
  double* ew;     /* a_ij */
  double* diag;   /* a_ii */
  int64_t* erev;  /* index of the reverse edge or -1 */
  uint8_t* props; /* bit0 depends, bit1 influences */
  uint8_t* iso;
  int symmetric;
  int is_f32;
} graph_t;

static void graph_free(graph_t* g) {
  free(g->eptr);
  free(g->ecol);
  free(g->ew);
  free(g->diag);
  free(g->erev);
  free(g->props);
  free(g->iso);
}

/* graph/matrix-graph.hpp:63-190: edges are the off-diagonal entries in
 * iterate order (ascending column for both stencils and CSR); edge index =
 * start[row] + position */
static void graph_build(const or_mat* m, graph_t* g) {
  const int64_t n = m->is_csr ? m->csr->n : m->st->n[0] * m->st->n[1] * m->st->n[2];
  g->n = n;
  g->symmetric = m->is_csr ? m->csr->symmetric : m->st->velocity == 0.0;
  g->is_f32 = m->is_csr ? m->csr->is_f32 : 0;
  g->eptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
  g->diag = (double*)malloc(sizeof(double) * (size_t)n);
  g->iso = (uint8_t*)calloc((size_t)n, 1);
  int64_t ne = 0;
  row_t_f64 r = {0};
  g->eptr[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    get_row_f64(m, i, &r);
    ne += r.n - 1;
    g->eptr[i + 1] = ne;
  }
  g->ecol = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ne + 1));
  g->ew = (double*)malloc(sizeof(double) * (size_t)(ne + 1));
  g->erev = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ne + 1));
  g->props = (uint8_t*)calloc((size_t)(ne + 1), 1);
  for (int64_t i = 0; i < n; ++i) {
    get_row_f64(m, i, &r);
    int64_t e = g->eptr[i];
    for (int k = 0; k < r.n; ++k) {
      if (r.cols[k] == i) {
        g->diag[i] = r.vals[k];
      } else {
        g->ecol[e] = r.cols[k];
        g->ew[e++] = r.vals[k];
      }
    }
  }
  free(r.dyn);
  /* reverse edges (graph/matrix-graph.hpp:312-332) */
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t e = g->eptr[i]; e < g->eptr[i + 1]; ++e) {
      const int64_t j = g->ecol[e];
      int64_t lo = g->eptr[j], hi = g->eptr[j + 1], f = -1;
      while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (g->ecol[mid] < i) lo = mid + 1;
        else hi = mid;
      }
      if (lo < g->eptr[j + 1] && g->ecol[lo] == i) f = lo;
      g->erev[e] = f;
    }
  }
}

/* DUNE symmetric criterion, linear-solver/multigrid/dependency/criterion.hpp:35-78,
 * with weights w_ij = -a_ij (graph/weight-transform.hpp:28-39):
 * s_ij = max(w_ij,0) * w'_ji / a_jj where w'_ji = max(w_ij,0) for symmetric
 * matrices, else max(w_ji,0) (0 if absent); strong iff s_ij > theta max_j s_ij.
 * depends(i->j) and influences(j->i) record the decision; a vertex without
 * strong edges is isolated.  Arithmetic in the level's value precision. */
#define MAXF(a, b) (((a) < (b)) ? (b) : (a))
static void criterion(graph_t* g, double theta_d) {
  double* st = NULL;
  float* stf = NULL;
  int64_t cap = 0;
  for (int64_t i = 0; i < g->n; ++i) {
    const int64_t e0 = g->eptr[i], e1 = g->eptr[i + 1];
    if (e1 - e0 > cap) {
      cap = e1 - e0 + 16;
      st = (double*)realloc(st, sizeof(double) * (size_t)cap);
      stf = (float*)realloc(stf, sizeof(float) * (size_t)cap);
    }
    int any = 0;
    if (g->is_f32) {
      float mx = 0;
      for (int64_t e = e0; e < e1; ++e) {
        const float w = -(float)g->ew[e];
        const float nn = MAXF(w, 0.0f);
        const float ovw = (float)g->diag[g->ecol[e]];
        float ow;
        if (g->symmetric) ow = nn;
        else ow = g->erev[e] >= 0 ? MAXF(-(float)g->ew[g->erev[e]], 0.0f) : 0.0f;
        const float s = nn * ow / ovw;
        mx = MAXF(mx, s);
        stf[e - e0] = s;
      }
      const float lb = (float)theta_d * mx;
      for (int64_t e = e0; e < e1; ++e) {
        const int cond = stf[e - e0] > lb;
        any = any || cond;
        g->props[e] = (uint8_t)((g->props[e] & ~1u) | (cond ? 1u : 0u));
        if (g->erev[e] >= 0)
          g->props[g->erev[e]] = (uint8_t)((g->props[g->erev[e]] & ~2u) | (cond ? 2u : 0u));
      }
    } else {
      double mx = 0;
      for (int64_t e = e0; e < e1; ++e) {
        const double w = -g->ew[e];
        const double nn = MAXF(w, 0.0);
        const double ovw = g->diag[g->ecol[e]];
        double ow;
        if (g->symmetric) ow = nn;
        else ow = g->erev[e] >= 0 ? MAXF(-g->ew[g->erev[e]], 0.0) : 0.0;
        const double s = nn * ow / ovw;
        mx = MAXF(mx, s);
        st[e - e0] = s;
      }
      const double lb = theta_d * mx;
      for (int64_t e = e0; e < e1; ++e) {
        const int cond = st[e - e0] > lb;
        any = any || cond;
        g->props[e] = (uint8_t)((g->props[e] & ~1u) | (cond ? 1u : 0u));
        if (g->erev[e] >= 0)
          g->props[g->erev[e]] = (uint8_t)((g->props[g->erev[e]] & ~2u) | (cond ? 2u : 0u));
      }
    }
    g->iso[i] = (uint8_t)!any;
  }
  free(st);
  free(stf);
}

/* the strength graph of one level on its own (criterion.hpp:35-78): edge
 * flags in row order (off-diagonal entries in ascending column), isolated
 * flags per vertex; returns the edge count (props may be NULL to query it) */
int64_t or_strength_graph(const or_mat* m, double theta, uint8_t* props, uint8_t* iso) {
  graph_t g;
  memset(&g, 0, sizeof g);
  graph_build(m, &g);
  const int64_t ne = g.eptr[g.n];
  if (props) {
    criterion(&g, theta);
    memcpy(props, g.props, (size_t)ne);
    memcpy(iso, g.iso, (size_t)g.n);
  }
  graph_free(&g);
  return ne;
}

#define DEPENDS(p) (((p)&1u) != 0)
#define INFLUENCES(p) (((p)&2u) != 0)
#define ONE_WAY(p) (DEPENDS(p) && !INFLUENCES(p))
#define TWO_WAY(p) (DEPENDS(p) && INFLUENCES(p))
#define STRONG(p) (DEPENDS(p) || INFLUENCES(p))

typedef struct {
  uint32_t* agg;      /* per vertex */
  uint32_t* blk;      /* per aggregate: max_sz slots */
  uint8_t* blksz;
  int64_t nblk_cap;
  int64_t max_sz;
  /* front: sorted array + marker */
  int64_t* front;
  int64_t nfront, front_cap;
  uint8_t* in_front;
  uint32_t* connected;
  int64_t nconn, conn_cap;
  int64_t* cand;
  int64_t ncand, cand_cap;
} aggst_t;

static void ensure_block(aggst_t* a, int64_t idx) {
  if (idx < a->nblk_cap) return;
  int64_t nc = a->nblk_cap ? a->nblk_cap * 2 : 1024;
  while (nc <= idx) nc *= 2;
  a->blk = (uint32_t*)realloc(a->blk, sizeof(uint32_t) * (size_t)(nc * a->max_sz));
  a->blksz = (uint8_t*)realloc(a->blksz, (size_t)nc);
  memset(a->blksz + a->nblk_cap, 0, (size_t)(nc - a->nblk_cap));
  a->nblk_cap = nc;
}

static void front_insert(aggst_t* a, int64_t v) {
  if (a->nfront == a->front_cap) {
    a->front_cap = a->front_cap ? a->front_cap * 2 : 64;
    a->front = (int64_t*)realloc(a->front, sizeof(int64_t) * (size_t)a->front_cap);
  }
  int64_t k = a->nfront;
  while (k > 0 && a->front[k - 1] > v) {
    a->front[k] = a->front[k - 1];
    --k;
  }
  a->front[k] = v;
  ++a->nfront;
  a->in_front[v] = 1;
}

static void front_erase(aggst_t* a, int64_t v) {
  int64_t k = 0;
  while (k < a->nfront && a->front[k] != v) ++k;
  if (k < a->nfront) {
    memmove(a->front + k, a->front + k + 1, sizeof(int64_t) * (size_t)(a->nfront - k - 1));
    --a->nfront;
  }
  a->in_front[v] = 0;
}

static int is_connected(const aggst_t* a, uint32_t agg) {
  for (int64_t k = 0; k < a->nconn; ++k)
    if (a->connected[k] == agg) return 1;
  return 0;
}

/* AggregateManager::add_help, aggregation/base.hpp:246-267: store the vertex,
 * then visit every off-diagonal neighbour: aggregates become "connected"
 * (including the aggregate being built), unaggregated vertices join the front */
static void add_help(aggst_t* a, const graph_t* g, int64_t u, uint32_t cur) {
  a->agg[u] = cur;
  a->blk[(int64_t)cur * a->max_sz + a->blksz[cur]] = (uint32_t)u;
  a->blksz[cur]++;
  for (int64_t e = g->eptr[u]; e < g->eptr[u + 1]; ++e) {
    const int64_t t = g->ecol[e];
    const uint32_t ta = a->agg[t];
    if (ta < OR_ISOLATED && !is_connected(a, ta)) {
      if (a->nconn == a->conn_cap) {
        a->conn_cap = a->conn_cap ? a->conn_cap * 2 : 64;
        a->connected = (uint32_t*)realloc(a->connected, sizeof(uint32_t) * (size_t)a->conn_cap);
      }
      a->connected[a->nconn++] = ta;
    }
    if (ta == OR_UNAGGREGATED && !a->in_front[t]) front_insert(a, t);
  }
}

static void add_candidates(aggst_t* a, const graph_t* g, uint32_t cur) {
  for (int64_t k = 0; k < a->ncand; ++k) {
    front_erase(a, a->cand[k]);
    add_help(a, g, a->cand[k], cur);
  }
}

static void push_cand(aggst_t* a, int64_t v) {
  if (a->ncand == a->cand_cap) {
    a->cand_cap = a->cand_cap ? a->cand_cap * 2 : 64;
    a->cand = (int64_t*)realloc(a->cand, sizeof(int64_t) * (size_t)a->cand_cap);
  }
  a->cand[a->ncand++] = v;
}

/* lexicographic (Size, double, Size) comparison: 1 greater, 0 equivalent, -1 less */
static int tup_cmp(int64_t a0, double a1, int64_t a2, int64_t b0, double b1, int64_t b2) {
  if (a0 != b0) return a0 > b0 ? 1 : -1;
  if (a1 > b1) return 1;
  if (a1 < b1) return -1;
  if (a1 != b1) return -2; /* unordered (NaN) */
  if (a2 != b2) return a2 > b2 ? 1 : -1;
  return 0;
}

/* Greedy aggregation of one level with one thread segment,
 * aggregation/base.hpp:448-707 (handle_vertex, grow_aggregate,
 * round_aggregate, find_merge_aggregate) and aggregate-map/multithread.hpp.
 * Returns the aggregate count; fills agg (per vertex) and the coarse->fine lists. */
static int64_t aggregate(const graph_t* g, int64_t min_sz, int64_t max_sz, int aggressive,
                         uint32_t* aggmap, int64_t** c2f_ptr, int64_t** c2f_idx,
                         int64_t stats[5]) {
  aggst_t a;
  memset(&a, 0, sizeof a);
  a.agg = aggmap;
  a.max_sz = max_sz;
  a.in_front = (uint8_t*)calloc((size_t)g->n, 1);
  for (int64_t i = 0; i < g->n; ++i) a.agg[i] = OR_UNAGGREGATED;
  int64_t multi = 0, one = 0, merged = 0, iso = 0;
  uint32_t counter = 0;
  ensure_block(&a, 0);
  for (int64_t v = 0; v < g->n; ++v) {
    if (a.agg[v] != OR_UNAGGREGATED) continue;
    if (g->iso[v]) {
      a.agg[v] = OR_ISOLATED;
      ++iso;
      continue;
    }
    const uint32_t cur = counter;
    ensure_block(&a, cur);
    a.blksz[cur] = 0;
    a.nfront = 0;
    a.nconn = 0;
    add_help(&a, g, v, cur);
    /* grow_aggregate, aggregation/base.hpp:532-585 */
    while ((int64_t)a.blksz[cur] < min_sz) {
      int64_t m0 = 0, m2 = 0;
      double m1 = 0.0;
      int has_cons2 = 0;
      a.ncand = 0;
      for (int64_t fi = 0; fi < a.nfront; ++fi) {
        const int64_t vtx = a.front[fi];
        if (g->iso[vtx]) continue;
        int64_t con = 0, nbs = 0, cons2 = 0, cons1 = 0;
        for (int64_t e = g->eptr[vtx]; e < g->eptr[vtx + 1]; ++e) {
          const int64_t h = g->ecol[e];
          const uint32_t ha = a.agg[h];
          const uint8_t p = g->props[e];
          con += (ha < OR_ISOLATED) && is_connected(&a, ha);
          nbs += a.in_front[h];
          cons2 += (ha == cur) && TWO_WAY(p);
          cons1 += (ha == cur) && ONE_WAY(p);
        }
        const double conv = (double)con / (double)(g->eptr[vtx + 1] - g->eptr[vtx]);
        if (cons2 > 0 && !has_cons2) {
          has_cons2 = 1;
          m0 = cons2;
          m1 = conv;
          m2 = nbs;
          a.ncand = 0;
          push_cand(&a, vtx);
        } else {
          const int64_t t0 = has_cons2 ? cons2 : cons1;
          if (t0 > 0) {
            const int c = tup_cmp(t0, conv, nbs, m0, m1, m2);
            if (c == 1) {
              m0 = t0;
              m1 = conv;
              m2 = nbs;
              a.ncand = 0;
              push_cand(&a, vtx);
            } else if (c == 0) {
              push_cand(&a, vtx);
            }
          }
        }
      }
      if (a.ncand == 0) break;
      const int64_t room = max_sz - (int64_t)a.blksz[cur];
      if (a.ncand > room) a.ncand = room;
      add_candidates(&a, g, cur);
    }
    /* round_aggregate, aggregation/base.hpp:601-636 */
    while ((int64_t)a.blksz[cur] < max_sz) {
      a.ncand = 0;
      for (int64_t fi = 0; fi < a.nfront; ++fi) {
        const int64_t vtx = a.front[fi];
        if (g->iso[vtx]) continue;
        int64_t cons = 0, unagg = 0, inagg = 0;
        for (int64_t e = g->eptr[vtx]; e < g->eptr[vtx + 1]; ++e) {
          const uint32_t ha = a.agg[g->ecol[e]];
          cons += (ha == cur) && DEPENDS(g->props[e]);
          unagg += ha == OR_UNAGGREGATED;
          inagg += ha == cur;
        }
        if (cons == 0 || unagg >= inagg) continue;
        push_cand(&a, vtx);
        break;
      }
      if (a.ncand == 0) break;
      add_candidates(&a, g, cur);
    }
    /* clear the front markers (AggregateHelpers::clear) */
    for (int64_t fi = 0; fi < a.nfront; ++fi) a.in_front[a.front[fi]] = 0;
    a.nfront = 0;
    if (a.blksz[cur] == 1 && max_sz > 1) {
      /* find_merge_aggregate, aggregation/base.hpp:495-510 */
      uint32_t target = OR_UNAGGREGATED;
      for (int64_t e = g->eptr[v]; e < g->eptr[v + 1]; ++e) {
        if (!STRONG(g->props[e])) continue;
        const uint32_t ha = a.agg[g->ecol[e]];
        if (ha < OR_ISOLATED && (int64_t)a.blksz[ha] < max_sz) {
          target = ha;
          break;
        }
      }
      if (target != OR_UNAGGREGATED) {
        a.blksz[cur] = 0;
        a.agg[v] = target;
        a.blk[(int64_t)target * max_sz + a.blksz[target]] = (uint32_t)v;
        a.blksz[target]++;
        ++merged;
        continue;
      }
      if (aggressive) {
        a.blksz[cur] = 0;
        a.agg[v] = OR_ISOLATED;
        ++iso;
        continue;
      }
      ++one;
    } else {
      ++multi;
    }
    ++counter;
  }
  const int64_t nagg = counter;
  int64_t* ptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nagg + 1));
  ptr[0] = 0;
  for (int64_t c = 0; c < nagg; ++c) ptr[c + 1] = ptr[c] + a.blksz[c];
  int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ptr[nagg] + 1));
  for (int64_t c = 0; c < nagg; ++c)
    for (int64_t k = 0; k < a.blksz[c]; ++k) idx[ptr[c] + k] = a.blk[c * max_sz + k];
  *c2f_ptr = ptr;
  *c2f_idx = idx;
  stats[0] = nagg;
  stats[1] = multi;
  stats[2] = one;
  stats[3] = merged;
  stats[4] = iso;
  free(a.in_front);
  free(a.blk);
  free(a.blksz);
  free(a.front);
  free(a.connected);
  free(a.cand);
  return nagg;
}

/* Galerkin product A_c = P^T A P, transport/matrix-shared.hpp:136-150, 188-202
 * with the builder of csr/shared.hpp:733-764: for coarse row I, fine rows in
 * coarse->fine order, entries in iterate order; each value is cast to the
 * coarse value type, summed in f64 into the first slot of its coarse column,
 * then the row is sorted by column and cast to the coarse value type. */
static void galerkin(const or_mat* fine, const uint32_t* map, int64_t nc, const int64_t* ptr,
                     const int64_t* idx, int coarse_f32, or_csr* out) {
  int64_t cap = 1024, nnz = 0;
  int64_t* rp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc + 1));
  int64_t* cc = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
  double* cv = (double*)malloc(sizeof(double) * (size_t)cap);
  row_t_f64 r = {0};
  int64_t rcap = 64;
  int64_t* rc = (int64_t*)malloc(sizeof(int64_t) * (size_t)rcap);
  double* rv = (double*)malloc(sizeof(double) * (size_t)rcap);
  rp[0] = 0;
  for (int64_t c = 0; c < nc; ++c) {
    int64_t m = 0;
    for (int64_t k = ptr[c]; k < ptr[c + 1]; ++k) {
      get_row_f64(fine, idx[k], &r);
      for (int e = 0; e < r.n; ++e) {
        const uint32_t J = map[r.cols[e]];
        if (J >= OR_ISOLATED) continue;
        const double val = coarse_f32 ? (double)(float)r.vals[e] : r.vals[e];
        int64_t f = 0;
        while (f < m && rc[f] != (int64_t)J) ++f;
        if (f < m) {
          rv[f] += val;
        } else {
          if (m == rcap) {
            rcap *= 2;
            rc = (int64_t*)realloc(rc, sizeof(int64_t) * (size_t)rcap);
            rv = (double*)realloc(rv, sizeof(double) * (size_t)rcap);
          }
          rc[m] = J;
          rv[m++] = val;
        }
      }
    }
    for (int64_t a = 1; a < m; ++a) {
      const int64_t kc = rc[a];
      const double kv = rv[a];
      int64_t j = a - 1;
      while (j >= 0 && rc[j] > kc) {
        rc[j + 1] = rc[j];
        rv[j + 1] = rv[j];
        --j;
      }
      rc[j + 1] = kc;
      rv[j + 1] = kv;
    }
    if (nnz + m > cap) {
      while (nnz + m > cap) cap *= 2;
      cc = (int64_t*)realloc(cc, sizeof(int64_t) * (size_t)cap);
      cv = (double*)realloc(cv, sizeof(double) * (size_t)cap);
    }
    for (int64_t e = 0; e < m; ++e) {
      cc[nnz] = rc[e];
      cv[nnz++] = coarse_f32 ? (double)(float)rv[e] : rv[e];
    }
    rp[c + 1] = nnz;
  }
  free(r.dyn);
  free(rc);
  free(rv);
  out->n = nc;
  out->rowptr = rp;
  out->cols = cc;
  out->vals = cv;
  out->is_f32 = coarse_f32;
  out->symmetric = fine->is_csr ? fine->csr->symmetric : fine->st->velocity == 0.0;
}

/* Cholesky-Banachiewicz, linear-solver/direct/cholesky/decomposition.hpp:44-99:
 * row i: every column j between the row's first stored column and i (gaps
 * filled with 0) gets L_ij = (a_ij - dot(L_j, L_i so far)) / L_jj, kept if
 * nonzero, and diag -= L_ij^2; the diagonal entry a_ii is added when visited;
 * L_ii = sqrt(diag). */
int64_t or_cholesky_decompose(const or_csr* a, int64_t* rowptr, int64_t* cols, double* vals) {
  const int64_t n = a->n;
  int64_t nnz = 0;
  rowptr[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    double diag = 0;
    const int64_t rs = nnz;
    int64_t next_start = 0;
    int is_first = 1;
#define CHOL_LAMBDA(jj, base_in)                                                   \
  do {                                                                             \
    const int64_t j_ = (jj);                                                       \
    const int64_t b1 = rowptr[j_], e1 = rowptr[j_ + 1];                            \
    double dot = 0;                                                                \
    int64_t it1 = b1;                                                              \
    for (int64_t k2 = rs; k2 < nnz; ++k2) {                                        \
      while (it1 != e1 && cols[it1] < cols[k2]) ++it1;                             \
      if (it1 == e1) break;                                                        \
      if (cols[it1] == cols[k2]) dot += vals[k2] * vals[it1];                      \
    }                                                                              \
    const double value = ((base_in)-dot) / vals[e1 - 1];                           \
    if (value != 0) {                                                              \
      diag -= value * value;                                                       \
      cols[nnz] = j_;                                                              \
      vals[nnz++] = value;                                                         \
    }                                                                              \
  } while (0)
    for (int64_t k = a->rowptr[i]; k < a->rowptr[i + 1]; ++k) {
      const int64_t j = a->cols[k];
      if (j < i) {
        if (is_first) {
          is_first = 0;
        } else {
          for (int64_t g = next_start; g < j; ++g) CHOL_LAMBDA(g, 0.0);
        }
        CHOL_LAMBDA(j, a->vals[k]);
        next_start = j + 1;
      } else if (j == i) {
        diag += a->vals[k];
      }
    }
    for (int64_t g = next_start; g < i; ++g) CHOL_LAMBDA(g, 0.0);
#undef CHOL_LAMBDA
    if (!(diag > 0)) return -1;
    cols[nnz] = i;
    vals[nnz++] = sqrt(diag);
    rowptr[i + 1] = nnz;
  }
  return nnz;
}

/* forward substitution then transposed backward substitution,
 * linear-solver/direct/substitution.hpp:27-52, 82-113 */
void or_cholesky_solve(const or_csr* l, const double* b, double* x) {
  const int64_t n = l->n;
  double* aux = (double*)malloc(sizeof(double) * (size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    double sum = b[i], diag = 0;
    for (int64_t k = l->rowptr[i]; k < l->rowptr[i + 1]; ++k) {
      const int64_t j = l->cols[k];
      if (j == i) diag = l->vals[k];
      else sum -= l->vals[k] * aux[j];
    }
    aux[i] = sum / diag;
  }
  for (int64_t i = 0; i < n; ++i) x[i] = aux[i];
  for (int64_t j = n - 1; j >= 0; --j) {
    const int64_t b0 = l->rowptr[j], e0 = l->rowptr[j + 1];
    const double xj = x[j] / l->vals[e0 - 1];
    x[j] = xj;
    for (int64_t k = b0; k < e0 - 1; ++k) x[l->cols[k]] = x[l->cols[k]] - l->vals[k] * xj;
  }
  free(aux);
}

/* Row-wise LU without pivoting, linear-solver/direct/lu/decomposition.hpp:30-111,
 * restated with dense storage and structure flags: row i starts from its own
 * entries; every finished row k < i holding a (possibly filled-in) entry w_k
 * eliminates it in ascending k, l_ik = w_k / u_kk, and updates the entries
 * of its upper structure (w_j -= l_ik u_kj, or w_j = -(l_ik u_kj) as fill-in);
 * then u_ii = w_i and u_ij = w_j for the remaining j > i, l_ii = 1.
 * Returns 0, or -1 on a zero or missing pivot. */
int or_lu_decompose_dense(const or_csr* a, double* lo, double* up) {
  const int64_t n = a->n;
  double* w = (double*)malloc(sizeof(double) * (size_t)n);
  char* present = (char*)malloc((size_t)n);
  char* ustruct = (char*)calloc((size_t)(n * n), 1);
  int status = 0;
  for (int64_t t = 0; t < n * n; ++t) lo[t] = up[t] = 0.0;
  for (int64_t i = 0; i < n && status == 0; ++i) {
    for (int64_t j = 0; j < n; ++j) {
      w[j] = 0.0;
      present[j] = 0;
    }
    for (int64_t k = a->rowptr[i]; k < a->rowptr[i + 1]; ++k) {
      w[a->cols[k]] = a->vals[k];
      present[a->cols[k]] = 1;
    }
    for (int64_t k = 0; k < i; ++k) {
      if (!present[k]) continue;
      const double l = w[k] / up[k * n + k];
      lo[i * n + k] = l;
      for (int64_t j = k + 1; j < n; ++j) {
        if (!ustruct[k * n + j]) continue;
        const double sub = l * up[k * n + j];
        if (present[j]) {
          w[j] -= sub;
        } else {
          w[j] = -sub;
          present[j] = 1;
        }
      }
    }
    if (!present[i] || w[i] == 0.0) {
      status = -1;
      break;
    }
    lo[i * n + i] = 1.0;
    up[i * n + i] = w[i];
    ustruct[i * n + i] = 1;
    for (int64_t j = i + 1; j < n; ++j) {
      if (!present[j]) continue;
      up[i * n + j] = w[j];
      ustruct[i * n + j] = 1;
    }
  }
  free(w);
  free(present);
  free(ustruct);
  return status;
}

/* forward substitution with the unit lower factor, then backward
 * substitution with the upper factor (linear-solver/direct/substitution.hpp:27-80),
 * both column by column: after x_j is final, it is subtracted from the rows
 * below (forward) or above (backward) it */
void or_lu_solve_dense(int64_t n, const double* lo, const double* up, const double* b,
                       double* x) {
  for (int64_t i = 0; i < n; ++i) x[i] = b[i];
  for (int64_t j = 0; j < n; ++j) {
    const double xj = x[j];
    for (int64_t i = j + 1; i < n; ++i) x[i] = x[i] - lo[i * n + j] * xj;
  }
  for (int64_t j = n - 1; j >= 0; --j) {
    const double xj = x[j] / up[j * n + j];
    x[j] = xj;
    for (int64_t i = 0; i < j; ++i) x[i] = x[i] - up[i * n + j] * xj;
  }
}

/* hierarchy/wright.hpp:120-143, 145-309 */
or_hierarchy* or_amg_setup(const or_mat* finest, const or_amg_params* p) {
  or_hierarchy* h = (or_hierarchy*)calloc(1, sizeof(or_hierarchy));
  h->finest = *finest;
  h->size[0] = finest->is_csr ? finest->csr->n : finest->st->n[0] * finest->st->n[1] * finest->st->n[2];
  int level = 0, aggressive = 0;
  or_mat fine = *finest;
  for (;;) {
    graph_t g;
    memset(&g, 0, sizeof g);
    graph_build(&fine, &g);
    criterion(&g, aggressive ? p->aggressive_strong_threshold : p->strong_threshold);
    uint32_t* map = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)g.n);
    int64_t *ptr, *idx;
    const int64_t nc = aggregate(&g, p->min_aggregate_size, p->max_aggregate_size, aggressive,
                                 map, &ptr, &idx, h->stats[level]);
    graph_free(&g);
    h->agg[level] = map;
    h->c2f_ptr[level] = ptr;
    h->c2f_idx[level] = idx;
    galerkin(&fine, map, nc, ptr, idx, p->coarse_f32, &h->mats[level + 1]);
    h->size[level + 1] = nc;
    h->nlev = level + 2;
    int stop = 0;
    if (nc < p->min_coarsen_size) {
      stop = 1;
    } else if (level + 1 >= p->max_level_num) {
      stop = 1;
    } else if (p->min_coarsen_factor * (double)nc > (double)h->size[level]) {
      if (aggressive) stop = 1;
      else aggressive = 1;
    }
    if (stop || level + 2 >= 32) break;
    fine.is_csr = 1;
    fine.st = NULL;
    fine.csr = &h->mats[level + 1];
    ++level;
  }
  for (int l = 1; l < h->nlev; ++l) {
    h->rhs64[l] = (double*)calloc((size_t)h->size[l], sizeof(double));
    h->sol64[l] = (double*)calloc((size_t)h->size[l], sizeof(double));
    h->rhs32[l] = (float*)calloc((size_t)h->size[l], sizeof(float));
    h->sol32[l] = (float*)calloc((size_t)h->size[l], sizeof(float));
  }
  const or_csr* c = &h->mats[h->nlev - 1];
  const int64_t n = c->n;
  if (!c->symmetric) {
    /* BiCgStab's nonsymmetric hierarchies take an LU coarsest solver
     * (test/aux/iterative.hpp:122-181) */
    h->lu = 1;
    h->lu_l = (double*)malloc(sizeof(double) * (size_t)(n * n));
    h->lu_u = (double*)malloc(sizeof(double) * (size_t)(n * n));
    if (or_lu_decompose_dense(c, h->lu_l, h->lu_u) != 0) {
      or_amg_free(h);
      return NULL;
    }
    return h;
  }
  const int64_t cap = n * (n + 1) / 2 + 1;
  int64_t* rp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
  int64_t* cc = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
  double* cv = (double*)malloc(sizeof(double) * (size_t)cap);
  or_cholesky_decompose(c, rp, cc, cv);
  h->chol.n = n;
  h->chol.rowptr = rp;
  h->chol.cols = cc;
  h->chol.vals = cv;
  return h;
}

void or_amg_free(or_hierarchy* h) {
  if (!h) return;
  for (int l = 0; l < h->nlev; ++l) {
    free(h->agg[l]);
    free(h->c2f_ptr[l]);
    free(h->c2f_idx[l]);
    free((void*)h->mats[l].rowptr);
    free((void*)h->mats[l].cols);
    free((void*)h->mats[l].vals);
  }
  free((void*)h->chol.rowptr);
  free((void*)h->chol.cols);
  free((void*)h->chol.vals);
  free(h->lu_l);
  free(h->lu_u);
  for (int l = 0; l < 32; ++l) {
    free(h->rhs64[l]);
    free(h->sol64[l]);
    free(h->rhs32[l]);
    free(h->sol32[l]);
  }
  free(h);
}

int or_amg_level_num(const or_hierarchy* h) { return h->nlev; }
int64_t or_amg_level_size(const or_hierarchy* h, int level) { return h->size[level]; }
const uint32_t* or_amg_agg_map(const or_hierarchy* h, int level) { return h->agg[level]; }
const int64_t* or_amg_c2f_ptr(const or_hierarchy* h, int level) { return h->c2f_ptr[level]; }
const int64_t* or_amg_c2f_idx(const or_hierarchy* h, int level) { return h->c2f_idx[level]; }
const or_csr* or_amg_matrix(const or_hierarchy* h, int level) { return &h->mats[level]; }
const or_csr* or_amg_cholesky(const or_hierarchy* h) { return &h->chol; }
int or_amg_is_lu(const or_hierarchy* h) { return h->lu; }
void or_amg_stats(const or_hierarchy* h, int level, int64_t out[5]) {
  for (int k = 0; k < 5; ++k) out[k] = h->stats[level][k];
}

#define REAL double
#define SUF(name) name##_f64
#define STATE_RHS(h) ((h)->rhs64)
#define STATE_SOL(h) ((h)->sol64)
#include "oracle_vcycle.inc"
#undef REAL
#undef SUF
#undef STATE_RHS
#undef STATE_SOL

#define REAL float
#define SUF(name) name##_f32
#define STATE_RHS(h) ((h)->rhs32)
#define STATE_SOL(h) ((h)->sol32)
#include "oracle_vcycle.inc"
#undef REAL
#undef SUF
#undef STATE_RHS
#undef STATE_SOL

/* ---------------------------------------------------------------------------
 * Krylov drivers */

/* The hierarchy preconditions the Krylov operator `a`: level 0's smoothing
 * and residual use `a` itself, coarser levels the hierarchy's Galerkin
 * matrices.  For the operator the hierarchy was built from this is the
 * reference's V-cycle; for a related operator (diffusion plus advection with
 * a hierarchy built from the diffusion part) the finest smoother still sees
 * the operator being solved. */
static or_hierarchy* precon_view(const or_hierarchy* h, const or_mat* a) {
  if (!h) return NULL;
  or_hierarchy* v = (or_hierarchy*)malloc(sizeof(or_hierarchy));
  *v = *h;
  v->finest = *a;
  /* the solve's own coarse vectors (zero), so that solves with one
   * hierarchy may run concurrently */
  for (int l = 1; l < v->nlev; ++l) {
    v->rhs64[l] = (double*)calloc((size_t)v->size[l], sizeof(double));
    v->sol64[l] = (double*)calloc((size_t)v->size[l], sizeof(double));
    v->rhs32[l] = (float*)calloc((size_t)v->size[l], sizeof(float));
    v->sol32[l] = (float*)calloc((size_t)v->size[l], sizeof(float));
  }
  return v;
}

static void precon_view_free(or_hierarchy* v) {
  if (!v) return;
  for (int l = 1; l < v->nlev; ++l) {
    free(v->rhs64[l]);
    free(v->sol64[l]);
    free(v->rhs32[l]);
    free(v->sol32[l]);
  }
  free(v);
}

/* SsorBroom::sweep (gauss-seidel-variant/solver.hpp:45-88) repeated `iters`
 * times as DoubleSweepImpl does with a fixed iteration count
 * (broom.hpp:159-171): forward x_in -> aux, backward aux -> x_out, then
 * forward x_out -> aux, backward aux -> x_out; one thread segment. */
void or_ssor_f64(const or_mat* a, const double* x_in, double* x_out, const double* b, double w,
                 int ultra, int iters) {
  const int64_t n = a->is_csr ? a->csr->n : a->st->n[0] * a->st->n[1] * a->st->n[2];
  const int64_t seg[2] = {0, n};
  double* aux = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  or_sor_f64(a, x_in, aux, b, w, 1, ultra, 1, seg);
  or_sor_f64(a, aux, x_out, b, w, 0, ultra, 1, seg);
  for (int k = 1; k < iters; ++k) {
    or_sor_f64(a, x_out, aux, b, w, 1, ultra, 1, seg);
    or_sor_f64(a, aux, x_out, b, w, 0, ultra, 1, seg);
  }
  free(aux);
}

/* precon 1: FP64 V-cycle, 2: FP32 V-cycle, 3: one SSOR step with omega = 1
 * from zero (test/aux/iterative.hpp:143-165), else none */
static void apply_precon(const or_mat* a, const or_hierarchy* h, int precon, int ultra,
                         const double* r, double* z, int64_t n) {
  if (precon == 3) {
    or_ssor_f64(a, NULL, z, r, 1.0, ultra, 1);
  } else if (precon == 1) {
    or_vcycle_f64(h, r, z, ultra);
  } else if (precon == 2) {
    float* r32 = (float*)calloc((size_t)n, sizeof(float));
    float* z32 = (float*)calloc((size_t)n, sizeof(float));
    for (int64_t i = 0; i < n; ++i) r32[i] = (float)r[i];
    or_vcycle_f32(h, r32, z32, ultra);
    for (int64_t i = 0; i < n; ++i) z[i] = (double)z32[i];
    free(r32);
    free(z32);
  } else {
    memcpy(z, r, sizeof(double) * (size_t)n);
  }
}

static int rere(int iteration) { return iteration > 0 && iteration % 8 == 0; }

/* conjugate-gradients/solver.hpp:50-115 */
int or_pcg(const or_mat* a, const or_hierarchy* h, int precon, int ultra, const double* b,
           double* x, double rel_tol, int max_it, double* res_hist) {
  const int64_t n = a->is_csr ? a->csr->n : a->st->n[0] * a->st->n[1] * a->st->n[2];
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  or_hierarchy* hv = precon_view(h, a);
  h = hv;
  or_residual_f64(a, b, x, r);
  double rr = or_dot_f64(r, r, n), rho;
  if (precon) {
    apply_precon(a, h, precon, ultra, r, p, n);
    rho = or_dot_f64(p, r, n);
  } else {
    memcpy(p, r, sizeof(double) * (size_t)n);
    rho = rr;
  }
  const double r0 = sqrt(rr);
  if (res_hist) res_hist[0] = r0;
  int it = 0;
  while (it < max_it && sqrt(rr) > rel_tol * r0 && rho != 0) {
    or_spmv_f64(a, p, q);
    const double pq = or_dot_f64(p, q, n);
    const double alpha = rho / pq;
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + p[i] * alpha;
    if (rere(it)) {
      or_residual_f64(a, b, x, r);
    } else {
      for (int64_t i = 0; i < n; ++i) r[i] = r[i] - q[i] * alpha;
    }
    rr = or_dot_f64(r, r, n);
    const double old = rho;
    if (precon) {
      apply_precon(a, h, precon, ultra, r, q, n);
      rho = or_dot_f64(r, q, n);
      const double beta = rho / old;
      for (int64_t i = 0; i < n; ++i) p[i] = q[i] + p[i] * beta;
    } else {
      rho = rr;
      const double beta = rho / old;
      for (int64_t i = 0; i < n; ++i) p[i] = r[i] + p[i] * beta;
    }
    ++it;
    if (res_hist) res_hist[it] = sqrt(rr);
  }
  free(r);
  free(p);
  free(q);
  precon_view_free(hv);
  return it;
}

/* bicgstab/solver.hpp:51-151 (right preconditioning) */
int or_pbicgstab(const or_mat* a, const or_hierarchy* h, int precon, int ultra,
                 const double* b, double* x, double rel_tol, int max_it, double* res_hist) {
  const int64_t n = a->is_csr ? a->csr->n : a->st->n[0] * a->st->n[1] * a->st->n[2];
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* r0 = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* v = (double*)malloc(sizeof(double) * (size_t)n);
  double* t = (double*)malloc(sizeof(double) * (size_t)n);
  double* y = (double*)malloc(sizeof(double) * (size_t)n);
  or_hierarchy* hv = precon_view(h, a);
  h = hv;
  or_residual_f64(a, b, x, r);
  memcpy(r0, r, sizeof(double) * (size_t)n);
  memcpy(p, r, sizeof(double) * (size_t)n);
  double rr = or_dot_f64(r, r, n);
  double rho = rr;
  const double rn0 = sqrt(rr);
  if (res_hist) res_hist[0] = rn0;
  int it = 0;
  while (it < max_it && sqrt(rr) > rel_tol * rn0 && rho != 0) {
    apply_precon(a, h, precon, ultra, p, y, n);
    or_spmv_f64(a, y, v);
    const double alpha = rho / or_dot_f64(r0, v, n);
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + y[i] * alpha;
    for (int64_t i = 0; i < n; ++i) r[i] = r[i] - v[i] * alpha;
    rr = or_dot_f64(r, r, n);
    apply_precon(a, h, precon, ultra, r, y, n);
    or_spmv_f64(a, y, t);
    const double omega = or_dot_f64(t, r, n) / or_dot_f64(t, t, n);
    for (int64_t i = 0; i < n; ++i) x[i] = x[i] + y[i] * omega;
    const double old = rho;
    if (rere(it)) {
      or_residual_f64(a, b, x, r);
    } else {
      for (int64_t i = 0; i < n; ++i) r[i] = r[i] - t[i] * omega;
    }
    rho = or_dot_f64(r, r0, n);
    rr = or_dot_f64(r, r, n);
    const double beta = (rho / old) * (alpha / omega);
    for (int64_t i = 0; i < n; ++i) p[i] = r[i] + (p[i] - v[i] * omega) * beta;
    ++it;
    if (res_hist) res_hist[it] = sqrt(rr);
  }
  free(r);
  free(r0);
  free(p);
  free(v);
  free(t);
  free(y);
  precon_view_free(hv);
  return it;
}

/* ---------------------------------------------------------------------------
 * Outlet flux and its adjoint gradient (not in the reference).
 *
 * J(p, c) = sum over the cells with x = nx-1 of (2 q_0 c_i) (p_i - p_end): the
 * flux leaving through the outlet's Dirichlet half cells, built from the
 * reference's border summands 2 q_0 c_i (valuator/stationary-diffusion-values.hpp:124-138).
 * With A(c) p = b(c) and A^T lambda = dJ/dp (A symmetric: g_i = [x = nx-1] 2 q_0 c_i),
 *   dJ/dlog c_k = c_k [ dJ/dc_k + lambda^T (db/dc_k - (dA/dc_k) p) ]
 *              = c_k ( [x = nx-1] 2 q_0 (p_k - p_end)
 *                    + lambda_k 2 q_0 ([x = 0](p_start - p_k) + [x = nx-1](p_end - p_k)) )
 *              - sum over faces (k, j): H(c_k, c_j) q_d c_j / (c_k + c_j)
 *                                      (lambda_k - lambda_j)(p_k - p_j),
 * using c_k dH(c_k, c_j)/dc_k = H(c_k, c_j) c_j / (c_k + c_j).  Per-cell
 * coefficient stencils (kind 0) without advection only. */

double or_outlet_flux(const or_stencil* s, const double* p) {
  double q[3];
  or_axis_quotients(s, q);
  const int64_t plane = s->n[1] * s->n[2], base = (s->n[0] - 1) * plane;
  double sum = 0;
  for (int64_t t = 0; t < plane; ++t) {
    const int64_t i = base + t;
    sum += ((2.0 * q[0]) * s->coeff[i]) * (p[i] - s->p_end);
  }
  return sum;
}

void or_outlet_flux_adjoint_rhs(const or_stencil* s, double* g) {
  double q[3];
  or_axis_quotients(s, q);
  const int64_t plane = s->n[1] * s->n[2], n = s->n[0] * plane, base = n - plane;
  for (int64_t i = 0; i < n; ++i) g[i] = i >= base ? (2.0 * q[0]) * s->coeff[i] : 0.0;
}

void or_outlet_flux_gradient(const or_stencil* s, const double* p, const double* lam,
                             double* grad) {
  double q[3];
  or_axis_quotients(s, q);
  const int64_t nx = s->n[0], ny = s->n[1], nz = s->n[2];
  const int64_t off[3] = {ny * nz, nz, 1};
  const double b2 = 2.0 * q[0];
  for (int64_t x = 0; x < nx; ++x)
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t z = 0; z < nz; ++z) {
        const int64_t k = (x * ny + y) * nz + z;
        const int64_t pos[3] = {x, y, z};
        const double ck = s->coeff[k], pk = p[k], lk = lam[k];
        double acc = 0;
        if (x + 1 == nx) acc = b2 * (pk - s->p_end);
        if (x == 0) acc += lk * (b2 * (s->p_start - pk));
        if (x + 1 == nx) acc += lk * (b2 * (s->p_end - pk));
        acc = ck * acc;
        /* faces in the row order -x, -y, -z, +z, +y, +x */
        for (int e = 0; e < 6; ++e) {
          const int d = e < 3 ? e : 5 - e;
          const int up = e >= 3;
          if (up ? pos[d] + 1 >= s->n[d] : pos[d] == 0) continue;
          const int64_t j = up ? k + off[d] : k - off[d];
          const double cj = s->coeff[j];
          const double w = or_harmonic_mean_f64(ck, cj) * q[d];
          if (w == 0.0) continue;
          const double t = w * (cj / (ck + cj));
          acc -= t * ((lk - lam[j]) * (pk - p[j]));
        }
        grad[k] = acc;
      }
}

/* ---------------------------------------------------------------------------
 * Effective diffusivity, EffDiffCalc (component-wise/sink/effective-diffusion.hpp:
 * 229-327), one thread: the mid plane x = nx / 2 holds the cells
 * [mid ny nz, (mid + 1) ny nz) and each pairs with its +x neighbour.
 *
 * area part (AreaPartSink, :49-121, and :230-250): the integer count of
 * filled cells among both cells of every pair, divided by ny nz, times 0.5.
 * A lookup cell is filled unless its material is "none"
 * (lookup-wright.hpp:98-106); per-cell coefficients have no materials and
 * every cell counts. */
double or_eff_diff_area_part(const or_stencil* s, const uint8_t* none_mats) {
  if (s->n[0] < 3) return NAN;
  const int64_t other = s->n[1] * s->n[2], mid = s->n[0] / 2;
  int64_t sum = 0;
  for (int64_t i1 = mid * other; i1 < (mid + 1) * other; ++i1) {
    const int64_t i2 = i1 + other;
    int64_t f1 = 1, f2 = 1;
    if (s->kind == 1 && none_mats) {
      f1 = none_mats[s->mat[i1]] ? 0 : 1;
      f2 = none_mats[s->mat[i2]] ? 0 : 1;
    }
    sum += f1 + f2;
  }
  return (double)sum / (double)other * 0.5;
}

/* the sink's sum (EffDiffSink, :127-222): D(i1, i2) (sol[i2] - sol[i1]) in index
 * order, D the valuator's diffusion coefficient without the axis quotient
 * (stationary-diffusion-value.hpp:91-96, stationary-diffusion-lookup.hpp:83-89);
 * with cmap sol is indexed by compressed rows and pairs with a dropped cell
 * are skipped (:175-190).  Then the normalisation of :320-326. */
double or_eff_diff(const or_stencil* s, const int64_t* cmap, const double* sol,
                   double area_part) {
  if (s->n[0] < 3) return NAN;
  double q[3];
  or_axis_quotients(s, q);
  const int64_t other = s->n[1] * s->n[2], mid = s->n[0] / 2;
  double sum = 0.0;
  for (int64_t i1 = mid * other; i1 < (mid + 1) * other; ++i1) {
    const int64_t i2 = i1 + other;
    int64_t k1 = i1, k2 = i2;
    if (cmap) {
      k1 = cmap[i1];
      k2 = cmap[i2];
      if (k1 < 0 || k2 < 0) continue;
    }
    const double d = s->kind == 1
                         ? or_harmonic_mean_f64(s->mat_coeffs[s->mat[i1]], s->mat_coeffs[s->mat[i2]])
                         : or_harmonic_mean_f64(s->coeff[i1], s->coeff[i2]);
    sum += d * (sol[k2] - sol[k1]);
  }
  double out = sum;
  out *= q[0];
  out /= area_part * ((1.0 * s->len[1]) * s->len[2]);
  out *= s->len[0];
  return out / (s->p_end - s->p_start);
}

/* ---------------------------------------------------------------------------
 * CSR with B x B block values: y = A x with MatrixVectorProductExpr's row path
 * (component-wise/matrix-vector-product.hpp:71-82): sum = A_ij x_j + sum over
 * the row's blocks in stored order (compat::fmadd, tensor/fixed/compat.hpp:24-33),
 * component r of A_ij x_j the left fold 0 + a_r0 x_0 + a_r1 x_1 + ...
 * (tensor/fixed/expressions.hpp:22-44), in f64; y rounded to the vectors'
 * type.  vals: B*B row-major entries per block. */
void or_bcsr_spmv_f64(int64_t nb, int bsz, const int64_t* rowptr, const int64_t* cols,
                      const double* vals, const double* x, double* y) {
  for (int64_t i = 0; i < nb; ++i) {
    double sum[4] = {0, 0, 0, 0};
    for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const double* a = vals + p * bsz * bsz;
      const double* xj = x + cols[p] * bsz;
      for (int r = 0; r < bsz; ++r) {
        double t = 0.0;
        for (int c = 0; c < bsz; ++c) t = t + a[r * bsz + c] * xj[c];
        sum[r] = t + sum[r];
      }
    }
    for (int r = 0; r < bsz; ++r) y[i * bsz + r] = sum[r];
  }
}

void or_bcsr_spmv_f32(int64_t nb, int bsz, const int64_t* rowptr, const int64_t* cols,
                      const double* vals, const float* x, float* y) {
  for (int64_t i = 0; i < nb; ++i) {
    double sum[4] = {0, 0, 0, 0};
    for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const double* a = vals + p * bsz * bsz;
      const float* xj = x + cols[p] * bsz;
      for (int r = 0; r < bsz; ++r) {
        double t = 0.0;
        for (int c = 0; c < bsz; ++c) t = t + (double)(float)a[r * bsz + c] * (double)xj[c];
        sum[r] = t + sum[r];
      }
    }
    for (int r = 0; r < bsz; ++r) y[i * bsz + r] = (float)sum[r];
  }
}
