// Host-side AMG setup: strength criterion, greedy aggregation, Galerkin
// products and the coarsest Cholesky factor.  The hierarchy it produces is
// uploaded once and consumed by the GPU V-cycle.
#pragma once

#include <cstdint>
#include <vector>

namespace lgpu::setup {

// Rows of a level's matrix with the entries of each row in the reference's
// iterate order (stencil rows: -x,-y,-z,+z,+y,+x, diagonal last; CSR rows:
// column order with the diagonal inline).
struct RowMatrix {
  int64_t n = 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> cols;
  std::vector<double> vals;  // values of the level's type (f32 levels hold floats)
  bool f32 = false;
  bool symmetric = true;
};

struct AggregationStats {
  int64_t aggregates = 0, multi = 0, single = 0, merged = 0, isolated = 0;
};

struct Params {
  int max_level_num = 16;
  int64_t min_coarsen_size = 32;
  double min_coarsen_factor = 1.2;
  double strong_threshold = 1.0 / 3.0;
  double aggressive_strong_threshold = 1.0 / 3.0;
  int64_t min_aggregate_size = 6;
  int64_t max_aggregate_size = 9;
  bool coarse_f32 = false;
  int threads = 0;  // host threads for the order-independent phases (0: hardware)
};

struct Level {
  std::vector<uint32_t> agg;      // fine vertex -> aggregate (0xFFFFFFFE isolated)
  std::vector<int64_t> c2f_ptr;   // aggregate -> fine vertices, insertion order
  std::vector<int32_t> c2f_idx;
  AggregationStats stats;
};

struct Hierarchy {
  std::vector<int64_t> sizes;     // level sizes, finest first
  std::vector<Level> transfer;    // transfer[l] maps level l to level l+1
  std::vector<RowMatrix> coarse;  // coarse[l] is level l+1's matrix (column order)
  std::vector<double> chol_dense; // coarsest lower factor, dense row major: Cholesky
                                  // (symmetric) or the unit lower LU factor
  std::vector<double> upper_dense;  // coarsest upper LU factor (nonsymmetric), dense row major
  bool lu = false;                  // nonsymmetric coarsest matrix: LU instead of Cholesky
  bool spd = true;                  // the coarsest factorisation succeeded
  // wall seconds of the setup phases, summed over levels: strength graph and
  // criterion, aggregation, Galerkin products, coarsest factorisation
  double seconds_criterion = 0, seconds_aggregation = 0, seconds_galerkin = 0, seconds_coarsest = 0;
};

constexpr uint32_t kUnaggregated = 0xFFFFFFFFu;
constexpr uint32_t kIsolated = 0xFFFFFFFEu;

// Strength flags of a level: per off-diagonal entry in row order (depends,
// influences) and per vertex (isolated).
struct StrengthFlags {
  std::vector<uint8_t> depends, influences, isolated;
};

// Another executor of the setup phases that do not depend on a visiting
// order: the strength criterion and the Galerkin product.  Their results
// must equal the host's bit for bit (the same operations in the same order
// per vertex / per coarse row); the greedy aggregation stays on the host.
struct Accel {
  virtual ~Accel() = default;
  virtual void criterion(const RowMatrix& m, double theta, StrengthFlags& out) = 0;
  virtual RowMatrix galerkin(const RowMatrix& fine, const Level& lv, bool coarse_f32) = 0;
};

Hierarchy build(const RowMatrix& finest, const Params& p, Accel* accel = nullptr);

// The coarsest factor of a hierarchy whose levels are set (Cholesky or, for a
// nonsymmetric coarsest matrix, LU); sets lu, spd, chol_dense, upper_dense.
void factor_coarsest(Hierarchy& h);

// Sparse Cholesky-Banachiewicz of a CSR matrix (column order rows); returns
// false if a pivot is not positive.  Output rows in column order, diagonal last.
bool cholesky(const RowMatrix& a, RowMatrix& lower);

// Row-wise LU without pivoting (linear-solver/direct/lu/decomposition.hpp:30-111)
// into dense row-major unit-lower and upper factors; false on a zero pivot.
bool lu_dense(const RowMatrix& a, std::vector<double>& lower, std::vector<double>& upper);

}  // namespace lgpu::setup
