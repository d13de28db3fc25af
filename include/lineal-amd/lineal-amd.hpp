// Header-only C++23 classes over the lineal_amd C ABI.  They mirror the
// reference's matrix, vector and solver interfaces (paths relative to the
// reference's include/lineal/) so that code written against those interfaces
// can switch to the GPU classes:
//   GpuDenseVector<Real>        DenseVector        tensor/dynamic/dense/vector/shared.hpp:35-362
//   GpuStencilValueMatrix       AdjacentStencilMatrix over StationaryDiffusionValuesValuator
//                               tensor/dynamic/stencil/matrix/shared.hpp:35-360,
//                               tensor/dynamic/stencil/valuator/stationary-diffusion-values.hpp:31-211
//   GpuStencilLookupMatrix      ... over StationaryDiffusionLookupValuator
//                               tensor/dynamic/stencil/valuator/stationary-diffusion-lookup.hpp:29-151
//   GpuCsrMatrix                CsrMatrix          tensor/dynamic/csr/shared.hpp:37-1103
//   GpuJacobiSolver             JacobiSolver       linear-solver/iterative-stationary/jacobi/solver.hpp:21-39
//   GpuSorSolver                SorSolver          linear-solver/iterative-stationary/gauss-seidel-variant/solver.hpp:22-44
//   GpuMultiGridSolver          MultiGridSolver    linear-solver/multigrid/solver.hpp:29-196
//   GpuConjugateGradients       ConjugateGradientsSolver  linear-solver/conjugate-gradients/solver.hpp:25-221
//   GpuBiCgStab                 BiCgStabSolver     linear-solver/bicgstab/solver.hpp:26-259
//   run_iterative_solver        linear-solver/runner.hpp:17-84
// Errors: every failing C call throws GpuError (the reference throws C++
// exceptions for structural errors; NaN propagation becomes a status).
// Objects of one GpuContext are used from one host thread; all work is
// ordered on the context's stream.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "lineal_amd.h"

namespace lineal::gpu {

class GpuError : public std::runtime_error {
 public:
  GpuError(int status, const std::string& what)
      : std::runtime_error(std::string(lgpu_status_string(status)) + ": " + what), status_(status) {}
  [[nodiscard]] int status() const { return status_; }

 private:
  int status_;
};

inline void check(int status) {
  if (status != LGPU_OK) throw GpuError(status, lgpu_last_error());
}

// One device and one HIP stream: the GPU counterpart of the reference Env's
// execution policy (parallel/execution-policy/adaptive.hpp:146-349).
class GpuContext {
 public:
  explicit GpuContext(int device = 0) {
    lgpu_ctx h = nullptr;
    check(lgpu_ctx_create(device, &h));
    h_.reset(h);
  }
  void synchronize() const { check(lgpu_ctx_synchronize(h_.get())); }
  [[nodiscard]] lgpu_ctx handle() const { return h_.get(); }

 private:
  struct Del {
    void operator()(lgpu_ctx c) const { lgpu_ctx_destroy(c); }
  };
  std::unique_ptr<lgpu_ctx_s, Del> h_;
};

template <typename TReal>
class GpuDenseVector {
  static_assert(std::is_same_v<TReal, double> || std::is_same_v<TReal, float>);

 public:
  using Value = TReal;
  using Size = std::size_t;
  static constexpr int dtype = std::is_same_v<TReal, double> ? LGPU_F64 : LGPU_F32;

  GpuDenseVector(Size size, const GpuContext& ctx) : ctx_(&ctx) {
    lgpu_vec h = nullptr;
    check(lgpu_vec_create(ctx.handle(), int64_t(size), dtype, &h));
    h_.reset(h);
  }
  GpuDenseVector(Size size, Value value, const GpuContext& ctx) : GpuDenseVector(size, ctx) {
    fill(value);
  }
  GpuDenseVector(std::span<const Value> host, const GpuContext& ctx)
      : GpuDenseVector(host.size(), ctx) {
    upload(host);
  }
  GpuDenseVector(GpuDenseVector&&) noexcept = default;
  GpuDenseVector& operator=(GpuDenseVector&&) noexcept = default;
  // copies are deep, as for the reference's owned vectors
  GpuDenseVector(const GpuDenseVector& o) : GpuDenseVector(o.size(), *o.ctx_) { assign(o); }
  GpuDenseVector& operator=(const GpuDenseVector& o) {
    if (this != &o) assign(o);
    return *this;
  }

  [[nodiscard]] Size size() const {
    int64_t n = 0;
    check(lgpu_vec_size(h_.get(), &n));
    return Size(n);
  }
  void fill(Value v) { check(lgpu_vec_fill(h_.get(), double(v))); }
  template <typename TOther>
  void assign(const GpuDenseVector<TOther>& src) {
    check(lgpu_vec_copy(src.handle(), h_.get()));
  }
  void upload(std::span<const Value> host) {
    check(lgpu_vec_upload(h_.get(), host.data(), int64_t(host.size())));
  }
  void download(std::span<Value> host) const {
    check(lgpu_vec_download(h_.get(), host.data(), int64_t(host.size())));
  }
  [[nodiscard]] std::vector<Value> to_host() const {
    std::vector<Value> v(size());
    download(v);
    return v;
  }
  // device pointer of the storage
  [[nodiscard]] Value* data() { return static_cast<Value*>(lgpu_vec_data(h_.get())); }
  [[nodiscard]] const Value* data() const {
    return static_cast<const Value*>(lgpu_vec_data(h_.get()));
  }
  [[nodiscard]] lgpu_vec handle() const { return h_.get(); }
  [[nodiscard]] const GpuContext& context() const { return *ctx_; }
  // takes ownership of a vector of this dtype that the library created
  static GpuDenseVector adopt(lgpu_vec h, const GpuContext& ctx) {
    return GpuDenseVector(h, ctx, AdoptTag{});
  }

 private:
  struct AdoptTag {};
  GpuDenseVector(lgpu_vec h, const GpuContext& ctx, AdoptTag) : ctx_(&ctx), h_(h) {}
  struct Del {
    void operator()(lgpu_vec v) const { lgpu_vec_destroy(v); }
  };
  const GpuContext* ctx_;
  std::unique_ptr<lgpu_vec_s, Del> h_;
};

// dot<double>(a, b), component-wise/sink/accumulator.hpp:175-190
template <typename T>
inline double dot(const GpuDenseVector<T>& a, const GpuDenseVector<T>& b) {
  double out = 0;
  check(lgpu_vec_dot(a.handle(), b.handle(), &out));
  return out;
}

// Box geometry and Dirichlet data (discretization/box-equidistant.hpp:23-111,
// tensor/dynamic/stencil/matrix/diffusion.hpp:52-96), flow axis 0.
struct BoxInfo {
  std::array<int64_t, 3> axis_sizes;
  std::array<double, 3> axis_lengths;
  double solution_start = 1.0;
  double solution_end = 0.0;
  bool zero_to_one = true;
  double velocity_x = 0.0;  // upwind advection along +x (not in the reference)

  explicit BoxInfo(std::array<int64_t, 3> sizes) : axis_sizes(sizes) {
    for (int d = 0; d < 3; ++d) axis_lengths[size_t(d)] = double(sizes[size_t(d)]);
  }
  BoxInfo(std::array<int64_t, 3> sizes, std::array<double, 3> lengths)
      : axis_sizes(sizes), axis_lengths(lengths) {}
  [[nodiscard]] int64_t cell_num() const { return axis_sizes[0] * axis_sizes[1] * axis_sizes[2]; }
  [[nodiscard]] lgpu_box c_box() const {
    lgpu_box b{};
    for (int d = 0; d < 3; ++d) {
      b.n[d] = axis_sizes[size_t(d)];
      b.len[d] = axis_lengths[size_t(d)];
    }
    b.zero_to_one = zero_to_one ? 1 : 0;
    b.p_start = solution_start;
    b.p_end = solution_end;
    b.velocity_x = velocity_x;
    return b;
  }
};

// Common operations of the GPU matrices: the reference's SharedMatrix
// queries plus the products the solvers use.
class GpuMatrixBase {
 public:
  [[nodiscard]] std::size_t row_num() const {
    int64_t n = 0;
    check(lgpu_op_rows(h_.get(), &n));
    return std::size_t(n);
  }
  [[nodiscard]] std::size_t column_num() const { return row_num(); }
  [[nodiscard]] bool is_symmetric() const { return symmetric_; }
  [[nodiscard]] lgpu_op handle() const { return h_.get(); }
  [[nodiscard]] const GpuContext& context() const { return *ctx_; }

  // y = A x, MatrixVectorProductExpr (component-wise/matrix-vector-product.hpp:34-140)
  template <typename T>
  void multiply(const GpuDenseVector<T>& x, GpuDenseVector<T>& y) const {
    check(lgpu_op_spmv(h_.get(), x.handle(), y.handle()));
  }
  template <typename T>
  GpuDenseVector<T> operator*(const GpuDenseVector<T>& x) const {
    GpuDenseVector<T> y(x.size(), *ctx_);
    multiply(x, y);
    return y;
  }
  // r = b - A x
  template <typename T>
  void residual(const GpuDenseVector<T>& b, const GpuDenseVector<T>& x, GpuDenseVector<T>& r) const {
    check(lgpu_op_residual(h_.get(), b.handle(), x.handle(), r.handle()));
  }

 protected:
  GpuMatrixBase(const GpuContext& ctx, lgpu_op h, bool symmetric)
      : ctx_(&ctx), h_(h), symmetric_(symmetric) {}
  struct Del {
    void operator()(lgpu_op o) const { lgpu_op_destroy(o); }
  };
  const GpuContext* ctx_;
  std::unique_ptr<lgpu_op_s, Del> h_;
  bool symmetric_;
};

// Matrix-free stencil with the Dirichlet right-hand side of its valuator
// (AdjacentStencilVector, tensor/dynamic/stencil/vector/shared.hpp:181-205).
class GpuStencilMatrixBase : public GpuMatrixBase {
 public:
  template <typename T = double>
  [[nodiscard]] GpuDenseVector<T> rhs() const {
    GpuDenseVector<T> b(row_num(), *ctx_);
    check(lgpu_op_rhs(h_.get(), b.handle()));
    return b;
  }
  [[nodiscard]] const BoxInfo& info() const { return box_; }
  // SOR sweeps over cached per-cell face weights and diagonal (the default)
  // or with them evaluated on the fly (false frees the caches); same bits
  void set_sor_cache(bool enable) { check(lgpu_op_set_sor_cache(h_.get(), enable ? 1 : 0)); }

 protected:
  GpuStencilMatrixBase(const GpuContext& ctx, lgpu_op h, BoxInfo box)
      : GpuMatrixBase(ctx, h, box.velocity_x == 0.0), box_(box) {}
  BoxInfo box_;

  static lgpu_op make_values(const GpuContext& ctx, const BoxInfo& box,
                             std::span<const double> coeff) {
    if (std::int64_t(coeff.size()) != box.cell_num())
      throw GpuError(LGPU_INVALID_ARG, "one coefficient per cell");
    lgpu_box b = box.c_box();
    lgpu_op h = nullptr;
    check(lgpu_stencil_values_create(ctx.handle(), &b, coeff.data(), &h));
    return h;
  }
  static lgpu_op make_lookup(const GpuContext& ctx, const BoxInfo& box,
                             std::span<const std::uint8_t> materials,
                             const std::array<double, 4>& coeffs) {
    if (std::int64_t(materials.size()) != box.cell_num())
      throw GpuError(LGPU_INVALID_ARG, "one material id per cell");
    lgpu_box b = box.c_box();
    lgpu_op h = nullptr;
    check(lgpu_stencil_lookup_create(ctx.handle(), &b, materials.data(), coeffs.data(), &h));
    return h;
  }
};

class GpuStencilValueMatrix : public GpuStencilMatrixBase {
 public:
  GpuStencilValueMatrix(const BoxInfo& box, std::span<const double> coeff, const GpuContext& ctx)
      : GpuStencilMatrixBase(ctx, make_values(ctx, box, coeff), box) {}
};

class GpuStencilLookupMatrix : public GpuStencilMatrixBase {
 public:
  GpuStencilLookupMatrix(const BoxInfo& box, std::span<const std::uint8_t> materials,
                         const std::array<double, 4>& material_coeffs, const GpuContext& ctx)
      : GpuStencilMatrixBase(ctx, make_lookup(ctx, box, materials, material_coeffs), box) {}
};

// CSR with sorted columns and the diagonal stored; values kept in TReal.
template <typename TReal = double>
class GpuCsrMatrix : public GpuMatrixBase {
 public:
  using Value = TReal;
  GpuCsrMatrix(std::span<const int64_t> rowptr, std::span<const int32_t> cols,
               std::span<const double> vals, bool symmetric, const GpuContext& ctx)
      : GpuMatrixBase(ctx, make(ctx, rowptr, cols, vals, symmetric), symmetric) {}
  // csr_from_adjacent_stencil (tensor/dynamic/csr/create-from.hpp:37-239)
  explicit GpuCsrMatrix(const GpuStencilMatrixBase& stencil)
      : GpuMatrixBase(stencil.context(), from_stencil(stencil), stencil.is_symmetric()) {}
  // takes ownership of a CSR operator with TReal values that the library created
  GpuCsrMatrix(lgpu_op h, bool symmetric, const GpuContext& ctx) : GpuMatrixBase(ctx, h, symmetric) {}

 private:
  static constexpr int dtype = std::is_same_v<TReal, double> ? LGPU_F64 : LGPU_F32;
  static lgpu_op make(const GpuContext& ctx, std::span<const int64_t> rowptr,
                      std::span<const int32_t> cols, std::span<const double> vals, bool symmetric) {
    if (rowptr.empty()) throw GpuError(LGPU_INVALID_ARG, "empty row pointer");
    lgpu_op h = nullptr;
    check(lgpu_csr_create(ctx.handle(), int64_t(rowptr.size()) - 1, rowptr.data(), cols.data(),
                          vals.data(), dtype, symmetric ? 1 : 0, &h));
    return h;
  }
  static lgpu_op from_stencil(const GpuStencilMatrixBase& s) {
    lgpu_op h = nullptr;
    check(lgpu_csr_from_stencil(s.handle(), dtype, &h));
    return h;
  }
};

// The system { lhs, rhs, compress_map } of csr_from_adjacent_stencil<Mat, Vec>
// (tensor/dynamic/csr/create-from.hpp:37-239).  With compression the rows
// whose diagonal is zero (ZeroToZero) are dropped and the rest renumbered in
// order (create-from.hpp:54-55, 92-99, 197-229); compress_map holds the row of
// every cell or -1, and is empty without compression.
template <typename TReal = double>
struct GpuCsrSystem {
  GpuCsrMatrix<TReal> lhs;
  GpuDenseVector<TReal> rhs;
  std::vector<int64_t> compress_map;
};

template <typename TReal = double>
inline GpuCsrSystem<TReal> csr_from_adjacent_stencil(const GpuStencilMatrixBase& stencil,
                                                     bool compress) {
  constexpr int dtype = std::is_same_v<TReal, double> ? LGPU_F64 : LGPU_F32;
  std::vector<int64_t> map(stencil.row_num());
  lgpu_op lhs = nullptr;
  lgpu_vec rhs = nullptr;
  check(lgpu_csr_system_from_stencil(stencil.handle(), dtype, compress ? 1 : 0, &lhs, &rhs,
                                     map.data()));
  const GpuContext& ctx = stencil.context();
  GpuCsrSystem<TReal> sys{GpuCsrMatrix<TReal>(lhs, stencil.is_symmetric(), ctx),
                          GpuDenseVector<TReal>::adopt(rhs, ctx), {}};
  if (compress) sys.compress_map = std::move(map);
  return sys;
}

// Matrix and vector files (io/tensor.hpp:25-263).  SmfReader builds the CSR
// matrix as the reference does (nonsymmetric unless declared symmetric).
template <typename TReal = double>
class SmfReader {
 public:
  explicit SmfReader(std::string path) : path_(std::move(path)) {
    check(lgpu_smf_info(path_.c_str(), &rows_, &cols_, &nnz_));
  }
  [[nodiscard]] std::uint64_t row_num() const { return std::uint64_t(rows_); }
  [[nodiscard]] std::uint64_t col_num() const { return std::uint64_t(cols_); }
  [[nodiscard]] std::uint64_t nnz_num() const { return std::uint64_t(nnz_); }
  [[nodiscard]] GpuCsrMatrix<TReal> read(const GpuContext& ctx, bool symmetric = false) const {
    constexpr int dtype = std::is_same_v<TReal, double> ? LGPU_F64 : LGPU_F32;
    lgpu_op h = nullptr;
    check(lgpu_smf_read(ctx.handle(), path_.c_str(), dtype, symmetric ? 1 : 0, &h));
    return GpuCsrMatrix<TReal>(h, symmetric, ctx);
  }

 private:
  std::string path_;
  int64_t rows_ = 0, cols_ = 0, nnz_ = 0;
};

struct SmfWriter {
  explicit SmfWriter(std::string path) : path(std::move(path)) {}
  void write(const GpuMatrixBase& mat) const { check(lgpu_smf_write(mat.handle(), path.c_str())); }
  std::string path;
};

template <typename TReal = double>
class DvfReader {
 public:
  explicit DvfReader(std::string path) : path_(std::move(path)) {
    check(lgpu_dvf_size(path_.c_str(), &size_));
  }
  [[nodiscard]] std::uint64_t size() const { return std::uint64_t(size_); }
  [[nodiscard]] GpuDenseVector<TReal> read(const GpuContext& ctx) const {
    lgpu_vec h = nullptr;
    check(lgpu_dvf_read(ctx.handle(), path_.c_str(), GpuDenseVector<TReal>::dtype, &h));
    return GpuDenseVector<TReal>::adopt(h, ctx);
  }

 private:
  std::string path_;
  int64_t size_ = 0;
};

struct DvfWriter {
  explicit DvfWriter(std::string path) : path(std::move(path)) {}
  template <typename T>
  void write(const GpuDenseVector<T>& vec) const {
    check(lgpu_dvf_write(vec.handle(), path.c_str()));
  }
  std::string path;
};

// Effective diffusivity through the mid plane x = nx / 2, EffDiffCalc
// (component-wise/sink/effective-diffusion.hpp:229-327).  The area part is
// computed at construction (:292-295); none_materials flags the "none"
// materials of a lookup operator (LookupSymmetricDiffusionInfo::none_materials).
class GpuEffDiffCalc {
 public:
  explicit GpuEffDiffCalc(const GpuStencilMatrixBase& valuator) {
    check(lgpu_eff_diff_area_part(valuator.handle(), nullptr, &area_part_));
  }
  GpuEffDiffCalc(const GpuStencilMatrixBase& valuator, const std::array<bool, 4>& none_materials) {
    std::array<std::uint8_t, 4> flags{};
    for (std::size_t m = 0; m < 4; ++m) flags[m] = none_materials[m] ? 1 : 0;
    check(lgpu_eff_diff_area_part(valuator.handle(), flags.data(), &area_part_));
  }
  [[nodiscard]] double area_part() const { return area_part_; }

  // sol indexed by cells
  template <typename T>
  double operator()(const GpuStencilMatrixBase& valuator, const GpuDenseVector<T>& sol) const {
    double out = 0.0;
    check(lgpu_eff_diff(valuator.handle(), nullptr, sol.handle(), area_part_, &out));
    return out;
  }
  // sol indexed by the rows of a compressed system (effdiff(valuator, compress_map, sol))
  template <typename TR, typename T>
  double operator()(const GpuStencilMatrixBase& valuator, const GpuCsrSystem<TR>& sys,
                    const GpuDenseVector<T>& sol) const {
    double out = 0.0;
    const lgpu_op map = sys.compress_map.empty() ? nullptr : sys.lhs.handle();
    check(lgpu_eff_diff(valuator.handle(), map, sol.handle(), area_part_, &out));
    return out;
  }

 private:
  double area_part_ = 0.0;
};

// CsrMatrix over fix::DenseMatrix<Real, tB, tB> block values
// (tensor/dynamic/csr/shared.hpp:37-1103, tensor/fixed/matrix.hpp:151):
// blocks row-major, TReal the stored value type; products accumulate in f64.
template <typename TReal, int tB>
class GpuBlockCsrMatrix {
  static_assert(tB >= 1 && tB <= 4);

 public:
  GpuBlockCsrMatrix(std::span<const int64_t> rowptr, std::span<const int32_t> cols,
                    std::span<const double> vals, const GpuContext& ctx)
      : ctx_(&ctx) {
    lgpu_bcsr h = nullptr;
    check(lgpu_bcsr_create(ctx.handle(), int64_t(rowptr.size()) - 1, tB, rowptr.data(), cols.data(),
                           vals.data(), GpuDenseVector<TReal>::dtype, &h));
    h_.reset(h);
  }
  [[nodiscard]] std::size_t row_num() const {
    int64_t nb = 0;
    check(lgpu_bcsr_rows(h_.get(), &nb, nullptr));
    return std::size_t(nb) * tB;
  }
  void multiply(const GpuDenseVector<TReal>& x, GpuDenseVector<TReal>& y) const {
    check(lgpu_bcsr_spmv(h_.get(), x.handle(), y.handle()));
  }
  [[nodiscard]] GpuDenseVector<TReal> operator*(const GpuDenseVector<TReal>& x) const {
    GpuDenseVector<TReal> y(x.size(), *ctx_);
    multiply(x, y);
    return y;
  }

 private:
  struct Del {
    void operator()(lgpu_bcsr m) const { lgpu_bcsr_destroy(m); }
  };
  const GpuContext* ctx_;
  std::unique_ptr<lgpu_bcsr_s, Del> h_;
};

// ---------------------------------------------------------------------------
// stationary iterative solvers: instantiate(lhs) -> Instance with ex-situ
// apply(sol_in, sol_out, rhs) (broom.hpp:188-244); a null sol_in is the
// constant-zero initial guess.

enum class SorDirection : bool { BACKWARD = false, FORWARD = true };
enum class SorVariant : bool { REGULAR = false, ULTRA = true };

namespace detail {
// In-situ application of a sweep (SingleSweepImpl::apply(sol, rhs, aux),
// iterative-stationary/broom.hpp:43-51, 92-123): each sweep reads the
// current iterate from an auxiliary copy.
template <typename TReal>
class InSitu {
 protected:
  template <typename TSweep>
  void in_situ(GpuDenseVector<TReal>& sol, int iterations, TSweep&& sweep) const {
    if (!aux_ || aux_->size() != sol.size())
      aux_ = std::make_unique<GpuDenseVector<TReal>>(sol.size(), sol.context());
    for (int k = 0; k < iterations; ++k) {
      aux_->assign(sol);
      sweep(*aux_, sol);
    }
  }

 private:
  mutable std::unique_ptr<GpuDenseVector<TReal>> aux_;
};
}  // namespace detail

// JacobiSolver{relax[, IterMan{iterations}]} (jacobi/solver.hpp:21-39)
template <typename TReal>
class GpuJacobiSolver {
 public:
  explicit GpuJacobiSolver(TReal relax, int iterations = 1) : relax_(relax), iterations_(iterations) {}
  class Instance : detail::InSitu<TReal> {
   public:
    Instance(const GpuMatrixBase& lhs, TReal relax, int iterations)
        : lhs_(&lhs), relax_(relax), iterations_(iterations) {}
    // ex-situ: iterations sweeps from sol_in (null: zero) into sol_out
    void apply(const GpuDenseVector<TReal>* sol_in, GpuDenseVector<TReal>& sol_out,
               const GpuDenseVector<TReal>& rhs) const {
      sweep(sol_in, sol_out, rhs);
      if (iterations_ > 1) in_situ_n(sol_out, rhs, iterations_ - 1);
    }
    // in-situ: iterations sweeps from sol
    void apply(GpuDenseVector<TReal>& sol, const GpuDenseVector<TReal>& rhs) const {
      in_situ_n(sol, rhs, iterations_);
    }

   private:
    void sweep(const GpuDenseVector<TReal>* in, GpuDenseVector<TReal>& out,
               const GpuDenseVector<TReal>& rhs) const {
      check(lgpu_op_jacobi(lhs_->handle(), in ? in->handle() : nullptr, out.handle(), rhs.handle(),
                           double(relax_)));
    }
    void in_situ_n(GpuDenseVector<TReal>& sol, const GpuDenseVector<TReal>& rhs, int n) const {
      this->in_situ(sol, n, [&](const GpuDenseVector<TReal>& in, GpuDenseVector<TReal>& out) {
        sweep(&in, out, rhs);
      });
    }
    const GpuMatrixBase* lhs_;
    TReal relax_;
    int iterations_;
  };
  [[nodiscard]] Instance instantiate(const GpuMatrixBase& lhs) const {
    return Instance(lhs, relax_, iterations_);
  }

 private:
  TReal relax_;
  int iterations_;
};

// SorSolver<Real, IterMan, dir, variant>{relax[, IterMan{iterations}]}
// (gauss-seidel-variant/solver.hpp:22-44, broom.hpp:232-239)
template <typename TReal, SorDirection tDir = SorDirection::FORWARD,
          SorVariant tVariant = SorVariant::REGULAR>
class GpuSorSolver {
 public:
  explicit GpuSorSolver(TReal relax, int iterations = 1) : relax_(relax), iterations_(iterations) {}
  class Instance : detail::InSitu<TReal> {
   public:
    Instance(const GpuMatrixBase& lhs, TReal relax, int iterations = 1)
        : lhs_(&lhs), relax_(relax), iterations_(iterations) {}
    // ex-situ: iterations lexicographic sweeps (one thread segment) from
    // sol_in (null: zero) into sol_out
    void apply(const GpuDenseVector<TReal>* sol_in, GpuDenseVector<TReal>& sol_out,
               const GpuDenseVector<TReal>& rhs) const {
      sweep(sol_in, sol_out, rhs);
      if (iterations_ > 1)
        this->in_situ(sol_out, iterations_ - 1,
                      [&](const GpuDenseVector<TReal>& in, GpuDenseVector<TReal>& out) {
                        sweep(&in, out, rhs);
                      });
    }
    // in-situ: iterations sweeps from sol
    void apply(GpuDenseVector<TReal>& sol, const GpuDenseVector<TReal>& rhs) const {
      this->in_situ(sol, iterations_, [&](const GpuDenseVector<TReal>& in, GpuDenseVector<TReal>& out) {
        sweep(&in, out, rhs);
      });
    }
    void sweep(const GpuDenseVector<TReal>* sol_in, GpuDenseVector<TReal>& sol_out,
               const GpuDenseVector<TReal>& rhs) const {
      check(lgpu_op_sor(lhs_->handle(), sol_in ? sol_in->handle() : nullptr, sol_out.handle(),
                        rhs.handle(), double(relax_), tDir == SorDirection::FORWARD ? 1 : 0,
                        tVariant == SorVariant::ULTRA ? 1 : 0));
    }
    // the reference's thread-segment semantics with segments of segment_rows
    // rows (CSR operators)
    void apply_segmented(const GpuDenseVector<TReal>* sol_in, GpuDenseVector<TReal>& sol_out,
                         const GpuDenseVector<TReal>& rhs, int64_t segment_rows) const {
      check(lgpu_op_sor_segmented(lhs_->handle(), sol_in ? sol_in->handle() : nullptr,
                                  sol_out.handle(), rhs.handle(), double(relax_),
                                  tDir == SorDirection::FORWARD ? 1 : 0,
                                  tVariant == SorVariant::ULTRA ? 1 : 0, segment_rows));
    }

   private:
    const GpuMatrixBase* lhs_;
    TReal relax_;
    int iterations_;
  };
  [[nodiscard]] Instance instantiate(const GpuMatrixBase& lhs) const {
    return Instance(lhs, relax_, iterations_);
  }

 private:
  TReal relax_;
  int iterations_;
};

// SsorSolver<Real, FixedIterationManager, variant>{relax, IterMan{iterations}}
// (gauss-seidel-variant/solver.hpp:45-96): apply runs `iterations` SSOR steps,
// each a forward and a backward lexicographic sweep.  The instance is also a
// Krylov preconditioner (GpuConjugateGradients<GpuSsorSolver<...>>), applied
// from a zero initial guess.
template <typename TReal, SorVariant tVariant = SorVariant::REGULAR>
class GpuSsorSolver {
 public:
  explicit GpuSsorSolver(TReal relax, int iterations = 1) : relax_(relax), iterations_(iterations) {}
  class Instance {
   public:
    Instance(const GpuMatrixBase& lhs, TReal relax, int iterations)
        : lhs_(&lhs), relax_(relax), iterations_(iterations) {
      lgpu_amg h = nullptr;
      check(lgpu_ssor_precon_create(lhs.handle(), double(relax), ultra(), iterations, &h));
      h_.reset(h);
    }
    // sol_in may be null (zero) or &sol_out
    void apply(const GpuDenseVector<TReal>* sol_in, GpuDenseVector<TReal>& sol_out,
               const GpuDenseVector<TReal>& rhs) const {
      check(lgpu_op_ssor(lhs_->handle(), sol_in ? sol_in->handle() : nullptr, sol_out.handle(),
                         rhs.handle(), double(relax_), ultra()));
      for (int k = 1; k < iterations_; ++k)
        check(lgpu_op_ssor(lhs_->handle(), sol_out.handle(), sol_out.handle(), rhs.handle(),
                           double(relax_), ultra()));
    }
    // in-situ (DoubleSweepImpl, broom.hpp:125-186)
    void apply(GpuDenseVector<TReal>& sol, const GpuDenseVector<TReal>& rhs) const {
      apply(&sol, sol, rhs);
    }
    [[nodiscard]] lgpu_amg handle() const { return h_.get(); }

   private:
    static constexpr int ultra() { return tVariant == SorVariant::ULTRA ? 1 : 0; }
    struct Del {
      void operator()(lgpu_amg a) const { lgpu_amg_destroy(a); }
    };
    const GpuMatrixBase* lhs_;
    TReal relax_;
    int iterations_;
    std::unique_ptr<lgpu_amg_s, Del> h_;
  };
  [[nodiscard]] Instance instantiate(const GpuMatrixBase& lhs) const {
    return Instance(lhs, relax_, iterations_);
  }

 private:
  TReal relax_;
  int iterations_;
};

// ---------------------------------------------------------------------------
// aggregation AMG

struct HierarchyWrightParams {
  int max_level_num = 16;
  int64_t min_coarsen_size = 32;
  double min_coarsen_factor = 1.2;
  double strong_threshold = 1.0 / 3.0;
  double aggressive_strong_threshold = 1.0 / 3.0;
  int64_t min_aggregate_size = 6;
  int64_t max_aggregate_size = 9;
  bool setup_on_device = true;  // criterion and Galerkin products on the GPU
};

// DefaultCycle{kind, iterations} (multigrid/cycle.hpp:24-79)
enum struct CycleKind : int { REGULAR = LGPU_CYCLE_REGULAR, DECREASING = LGPU_CYCLE_DECREASING };
struct DefaultCycle {
  CycleKind kind = CycleKind::REGULAR;
  int iterations = 1;
};

// MultiGridSolver{relax, DefaultCycle{kind, iterations}, wright}
// (multigrid/solver.hpp:183-186) with SOR(1) pre/post smoothers and a
// Cholesky coarsest solver (LU for a nonsymmetric operator).  TLevelReal
// selects the hierarchy's precision: float gives the mixed-precision
// preconditioner (FP32 hierarchy, smoothers and level operators; FP64
// coarsest solve).
template <typename TLevelReal = double>
class GpuMultiGridSolver {
 public:
  explicit GpuMultiGridSolver(HierarchyWrightParams params = {},
                              SorVariant variant = SorVariant::REGULAR,
                              int64_t coarse_segment_rows = 64)
      : params_(params), variant_(variant), coarse_segment_rows_(coarse_segment_rows) {}
  GpuMultiGridSolver(double relax, DefaultCycle cycle, HierarchyWrightParams params = {},
                     SorVariant variant = SorVariant::REGULAR, int64_t coarse_segment_rows = 64)
      : params_(params), variant_(variant), coarse_segment_rows_(coarse_segment_rows),
        relax_(relax), cycle_(cycle) {}

  class Instance {
   public:
    Instance(const GpuMatrixBase& lhs, const GpuMultiGridSolver& s) : lhs_(&lhs) {
      const lgpu_amg_params p = s.c_params();
      lgpu_amg h = nullptr;
      check(lgpu_amg_create(lhs.handle(), &p, &h));
      h_.reset(h);
    }
    // over a hierarchy built elsewhere (e.g. the reference's own host setup):
    // levels[l] maps level l to level l + 1 (lgpu_amg_create_from_levels)
    Instance(const GpuMatrixBase& lhs, const GpuMultiGridSolver& s,
             std::span<const lgpu_amg_level> levels)
        : lhs_(&lhs) {
      const lgpu_amg_params p = s.c_params();
      lgpu_amg h = nullptr;
      check(lgpu_amg_create_from_levels(lhs.handle(), &p, int(levels.size()), levels.data(), &h));
      h_.reset(h);
    }
    // ex-situ: one cycle from sol_in (null: the constant-zero initial guess,
    // the preconditioner's use) into sol_out (multigrid/solver.hpp:47-53)
    void apply(const GpuDenseVector<TLevelReal>* sol_in, GpuDenseVector<TLevelReal>& sol_out,
               const GpuDenseVector<TLevelReal>& rhs) const {
      if (!sol_in && std::is_same_v<TLevelReal, double>)
        check(lgpu_amg_vcycle(h_.get(), rhs.handle(), sol_out.handle()));
      else
        check(lgpu_amg_cycle(h_.get(), sol_in ? sol_in->handle() : nullptr, rhs.handle(),
                             sol_out.handle()));
    }
    // in-situ: one cycle from sol (multigrid/solver.hpp:41-46)
    void apply(GpuDenseVector<TLevelReal>& sol, const GpuDenseVector<TLevelReal>& rhs) const {
      if (!aux_ || aux_->size() != sol.size())
        aux_ = std::make_unique<GpuDenseVector<TLevelReal>>(sol.size(), sol.context());
      aux_->assign(sol);
      check(lgpu_amg_cycle(h_.get(), aux_->handle(), rhs.handle(), sol.handle()));
    }
    [[nodiscard]] int level_num() const {
      int l = 0;
      check(lgpu_amg_level_num(h_.get(), &l));
      return l;
    }
    [[nodiscard]] std::size_t level_size(int level) const {
      int64_t n = 0;
      check(lgpu_amg_level_size(h_.get(), level, &n));
      return std::size_t(n);
    }
    [[nodiscard]] lgpu_amg handle() const { return h_.get(); }
    [[nodiscard]] const GpuMatrixBase& lhs() const { return *lhs_; }

   private:
    struct Del {
      void operator()(lgpu_amg a) const { lgpu_amg_destroy(a); }
    };
    const GpuMatrixBase* lhs_;
    std::unique_ptr<lgpu_amg_s, Del> h_;
    mutable std::unique_ptr<GpuDenseVector<TLevelReal>> aux_;
  };
  [[nodiscard]] Instance instantiate(const GpuMatrixBase& lhs) const { return Instance(lhs, *this); }
  [[nodiscard]] lgpu_amg_params c_params() const {
    lgpu_amg_params p;
    lgpu_amg_default_params(&p);
    p.max_level_num = params_.max_level_num;
    p.min_coarsen_size = params_.min_coarsen_size;
    p.min_coarsen_factor = params_.min_coarsen_factor;
    p.strong_threshold = params_.strong_threshold;
    p.aggressive_strong_threshold = params_.aggressive_strong_threshold;
    p.min_aggregate_size = params_.min_aggregate_size;
    p.max_aggregate_size = params_.max_aggregate_size;
    p.precision = std::is_same_v<TLevelReal, float> ? LGPU_F32 : LGPU_F64;
    p.ultra = variant_ == SorVariant::ULTRA ? 1 : 0;
    p.coarse_segment_rows = coarse_segment_rows_;
    p.relax = relax_;
    p.cycle_kind = int(cycle_.kind);
    p.cycle_iterations = cycle_.iterations;
    p.setup_on_device = params_.setup_on_device ? 1 : 0;
    if (sequential_rows_ >= 0) p.sequential_rows = sequential_rows_;
    return p;
  }
  // coarse levels with at most `rows` rows are swept in one segment (0:
  // none; the library's default otherwise)
  GpuMultiGridSolver& sequential_rows(int64_t rows) {
    sequential_rows_ = rows;
    return *this;
  }

 private:
  HierarchyWrightParams params_;
  SorVariant variant_;
  int64_t coarse_segment_rows_;
  int64_t sequential_rows_ = -1;
  double relax_ = 1.0;
  DefaultCycle cycle_{};
};

// ---------------------------------------------------------------------------
// Krylov solvers: instantiate(lhs) -> Instance; instance.instantiate(sol, rhs)
// -> SystemInstance with the reference's iterate(), residual_norm(), rho()
// and solution() (conjugate-gradients/solver.hpp:75-121), plus solve(), the
// fused loop that iterates on the device until a relative-residual stop rule
// holds.

struct StopRule {
  double rel_tol = 1e-8;
  int max_it = 1000;
};

struct SolveReport {
  int status = LGPU_OK;
  int iterations = 0;
  std::vector<double> residual_norms;  // ||r_k||_2, k = 0..iterations
  [[nodiscard]] bool converged() const { return status == LGPU_OK; }
  [[nodiscard]] double residual_norm() const {
    return residual_norms.empty() ? 0.0 : residual_norms.back();
  }
};

namespace detail {
using KrylovFn = int (*)(lgpu_op, lgpu_amg, lgpu_vec, lgpu_vec, const lgpu_krylov_params*, int*,
                         double*);
using KrylovCreateFn = int (*)(lgpu_op, lgpu_amg, lgpu_vec, lgpu_vec, int, lgpu_krylov*);

template <KrylovFn tFn, KrylovCreateFn tCreate, typename TPrecon>
class KrylovSolver {
 public:
  // ConjugateGradientsSolver{type_tag<Defs>, rere_controller, precon}: the
  // true residual every `rere_every` iterations (test/aux/iterative.hpp:97-99)
  explicit KrylovSolver(int rere_every = 8, const TPrecon* precon = nullptr)
      : rere_every_(rere_every), precon_(precon) {}

  class Instance {
   public:
    Instance(const GpuMatrixBase& lhs, const KrylovSolver& s)
        : lhs_(&lhs), rere_every_(s.rere_every_) {
      if (s.precon_) precon_.emplace_back(s.precon_->instantiate(lhs));
    }
    class SystemInstance {
     public:
      // the reference's constructor work (the initial residual, and for CG
      // the preconditioned residual and rho) runs on the first call of
      // iterate(), residual_norm() or rho(), which solve() never needs
      SystemInstance(const Instance& inst, GpuDenseVector<double>& sol,
                     const GpuDenseVector<double>& rhs)
          : inst_(&inst), sol_(&sol), rhs_(&rhs) {}
      // SystemInstance::iterate (conjugate-gradients/solver.hpp:75-115)
      void iterate() { check(lgpu_krylov_iterate(system())); }
      [[nodiscard]] double residual_norm() const {
        double v = 0;
        check(lgpu_krylov_residual_norm(system(), &v));
        return v;
      }
      [[nodiscard]] double rho() const {
        double v = 0;
        check(lgpu_krylov_rho(system(), &v));
        return v;
      }
      [[nodiscard]] std::size_t iteration() const {
        int v = 0;
        check(lgpu_krylov_iteration(system(), &v));
        return std::size_t(v);
      }
      // the fused loop from the current solution
      SolveReport solve(StopRule stop) {
        lgpu_krylov_params p{stop.rel_tol, stop.max_it, inst_->rere_every_};
        SolveReport r;
        r.residual_norms.resize(size_t(stop.max_it) + 1);
        const int st = tFn(inst_->lhs_->handle(),
                           inst_->precon_.empty() ? nullptr : inst_->precon_.front().handle(),
                           rhs_->handle(), sol_->handle(), &p, &r.iterations,
                           r.residual_norms.data());
        if (st != LGPU_OK && st != LGPU_NOT_CONVERGED && st != LGPU_BREAKDOWN &&
            st != LGPU_NONFINITE)
          check(st);
        r.status = st;
        r.residual_norms.resize(size_t(r.iterations) + 1);
        return r;
      }
      [[nodiscard]] const GpuDenseVector<double>& solution() const { return *sol_; }

     private:
      lgpu_krylov system() const {
        if (!sys_) {
          lgpu_krylov h = nullptr;
          check(tCreate(inst_->lhs_->handle(),
                        inst_->precon_.empty() ? nullptr : inst_->precon_.front().handle(),
                        rhs_->handle(), sol_->handle(), inst_->rere_every_, &h));
          sys_.reset(h);
        }
        return sys_.get();
      }
      struct Del {
        void operator()(lgpu_krylov k) const { lgpu_krylov_destroy(k); }
      };
      const Instance* inst_;
      GpuDenseVector<double>* sol_;
      const GpuDenseVector<double>* rhs_;
      mutable std::unique_ptr<lgpu_krylov_s, Del> sys_;
    };
    [[nodiscard]] SystemInstance instantiate(GpuDenseVector<double>& sol,
                                             const GpuDenseVector<double>& rhs) const {
      return SystemInstance(*this, sol, rhs);
    }

   private:
    const GpuMatrixBase* lhs_;
    int rere_every_;
    std::vector<typename TPrecon::Instance> precon_;  // zero or one
  };
  [[nodiscard]] Instance instantiate(const GpuMatrixBase& lhs) const { return Instance(lhs, *this); }

 private:
  int rere_every_;
  const TPrecon* precon_;
};
}  // namespace detail

template <typename TPrecon = GpuMultiGridSolver<double>>
using GpuConjugateGradients = detail::KrylovSolver<&lgpu_cg_solve, &lgpu_cg_create, TPrecon>;
template <typename TPrecon = GpuMultiGridSolver<double>>
using GpuBiCgStab = detail::KrylovSolver<&lgpu_bicgstab_solve, &lgpu_bicgstab_create, TPrecon>;

// run_iterative_solver with the relative residual stop rule, run as the
// fused device loop.
template <typename TSolver>
inline SolveReport run_iterative_solver(const TSolver& solver, const GpuMatrixBase& lhs,
                                        GpuDenseVector<double>& sol,
                                        const GpuDenseVector<double>& rhs, StopRule stop) {
  auto inst = solver.instantiate(lhs);
  auto sys = inst.instantiate(sol, rhs);
  return sys.solve(stop);
}

// run_iterative_solver(solver, lhs, sol, rhs, aux, env, iter_op)
// (linear-solver/runner.hpp:17-84): iter_op(iteration, instance) returns an
// std::optional; the first value it returns is the result.  A Krylov
// solver's system instance is built from (sol, rhs) and iter_op sees it at
// iteration 0 and after every iterate(); a stationary solver's instance is
// applied in-situ to sol and iter_op sees it after every apply, from 1.
template <typename TSolver, typename TVector, typename TIterOp>
requires(!std::is_same_v<std::decay_t<TIterOp>, StopRule>)
inline auto run_iterative_solver(const TSolver& solver, const GpuMatrixBase& lhs, TVector& sol,
                                 const TVector& rhs, TIterOp&& iter_op) {
  auto inst = solver.instantiate(lhs);
  if constexpr (requires { inst.instantiate(sol, rhs); }) {
    auto sys = inst.instantiate(sol, rhs);
    if (auto res = iter_op(std::size_t{0}, sys); res.has_value()) return *res;
    for (std::size_t i = 1;; ++i) {
      sys.iterate();
      if (auto res = iter_op(i, sys); res.has_value()) return *res;
    }
  } else {
    for (std::size_t i = 1;; ++i) {
      inst.apply(sol, rhs);
      if (auto res = iter_op(i, inst); res.has_value()) return *res;
    }
  }
}

// ---------------------------------------------------------------------------
// adjoint of the solve (not in the reference): outlet flux J and dJ/dlog c

struct AdjointReport {
  int status = LGPU_OK;
  double flux = 0.0;
  int forward_iterations = 0, adjoint_iterations = 0;
  [[nodiscard]] bool converged() const { return status == LGPU_OK; }
};

template <typename TPreconInstance>
inline AdjointReport outlet_flux_gradient(const GpuStencilValueMatrix& lhs,
                                          const TPreconInstance* precon,
                                          GpuDenseVector<double>& p, GpuDenseVector<double>& lambda,
                                          GpuDenseVector<double>& grad, StopRule stop,
                                          int rere_every = 8) {
  lgpu_krylov_params prm{stop.rel_tol, stop.max_it, rere_every};
  AdjointReport r;
  int its[2] = {0, 0};
  const int st = lgpu_adjoint_grad(lhs.handle(), precon ? precon->handle() : nullptr, &prm,
                                   p.handle(), lambda.handle(), grad.handle(), &r.flux, its);
  if (st != LGPU_OK && st != LGPU_NOT_CONVERGED && st != LGPU_BREAKDOWN && st != LGPU_NONFINITE)
    check(st);
  r.status = st;
  r.forward_iterations = its[0];
  r.adjoint_iterations = its[1];
  return r;
}

}  // namespace lineal::gpu
