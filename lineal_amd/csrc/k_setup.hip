// AMG setup phases on the GPU: the strength criterion and the Galerkin
// product, the parts of HierarchyWrightInstance::create_impl
// (multigrid/hierarchy/wright.hpp:145-309) that do not depend on a visiting
// order.  Each vertex / coarse row performs the host setup's operations in
// the host's order, so the flags and the coarse matrices equal the host's
// bit for bit; the greedy aggregation between them stays on the host.
#include <memory>

#include "internal.hpp"
#include "setup.hpp"

namespace lgpu {

namespace {

constexpr int NT = 256;

unsigned blocks_for(int64_t n) { return unsigned((n + NT - 1) / NT); }

// std::max(a, b) of the host: (a < b) ? b : a
template <typename W>
__device__ __forceinline__ W wmax(W a, W b) {
  return a < b ? b : a;
}

// Level rows are in column order, except that a stencil row keeps its
// diagonal last (the reference's iterate order, -x, -y, -z, +z, +y, +x,
// diagonal).  Position of column c in row j, or -1: a diagonal in the last
// place is set aside first, so the search runs over sorted columns.
__device__ __forceinline__ int64_t find_col(const int64_t* rp, const int32_t* cols, int64_t j,
                                            int32_t c) {
  int64_t lo = rp[j], hi = rp[j + 1];
  if (hi > lo && cols[hi - 1] == j) {
    if (c == j) return hi - 1;
    --hi;
  }
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (cols[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return (lo < rp[j + 1] && cols[lo] == c) ? lo : -1;
}

__global__ void diag_kernel(int64_t n, const int64_t* __restrict__ rp,
                            const int32_t* __restrict__ cols, const double* __restrict__ vals,
                            double* __restrict__ diag, int64_t* __restrict__ dpos) {
  const int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x;
  if (i >= n) return;
  // every level row stores its diagonal (checked where matrices enter)
  const int64_t p = find_col(rp, cols, i, int32_t(i));
  dpos[i] = p;
  diag[i] = p >= 0 ? vals[p] : 0.0;
}

// DuneDependencyCriterion::handle (dependency/criterion.hpp:35-78) with the
// NegativeOffdiagonalTransform, in the level's value type W: s = max(w, 0) *
// w' / a_jj with w' = max(w, 0) (symmetric) or max(w_ji, 0) (0 without a
// reverse entry); strong iff s > theta * max s.  Edge e of row i is the
// off-diagonal entry at position p: e = p - i - [p > diagonal position].
template <typename W>
__global__ void criterion_kernel(int64_t n, const int64_t* __restrict__ rp,
                                 const int32_t* __restrict__ cols, const double* __restrict__ vals,
                                 const double* __restrict__ diag, const int64_t* __restrict__ dpos,
                                 int symmetric, W theta, uint8_t* __restrict__ depends,
                                 uint8_t* __restrict__ influences, uint8_t* __restrict__ iso) {
  const int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x;
  if (i >= n) return;
  const int64_t p0 = rp[i], p1 = rp[i + 1], pd = dpos[i];
  auto strength = [&](int64_t p, int64_t& rev) {
    const int32_t j = cols[p];
    const W w = -W(vals[p]);
    const W nn = wmax(w, W(0));
    const W head_weight = W(diag[j]);
    rev = find_col(rp, cols, j, int32_t(i));
    W opposite;
    if (symmetric) opposite = nn;
    else opposite = rev >= 0 ? wmax(-W(vals[rev]), W(0)) : W(0);
    return nn * opposite / head_weight;
  };
  W mx = W(0);
  for (int64_t p = p0; p < p1; ++p) {
    if (p == pd) continue;
    int64_t rev;
    mx = wmax(mx, strength(p, rev));
  }
  const W lb = theta * mx;
  bool any = false;
  for (int64_t p = p0; p < p1; ++p) {
    if (p == pd) continue;
    int64_t rev;
    const bool strong = strength(p, rev) > lb;
    any = any || strong;
    depends[p - i - (p > pd ? 1 : 0)] = strong;
    if (rev >= 0) {
      const int64_t j = cols[p];
      influences[rev - j - (rev > dpos[j] ? 1 : 0)] = strong;
    }
  }
  iso[i] = !any;
}

// Galerkin product (transport/matrix-shared.hpp:188-202 with the non-unique
// row builder of csr/shared.hpp:733-764): coarse row I collects (J, a_ij)
// over its members i in coarse -> fine list order and their entries in
// column order, skipping isolated J; entries of equal J are summed in that
// order and the row is sorted by J.  Here the sequence is written out, sorted
// stably by J and folded per J, which is the same sum in the same order.
__global__ void galerkin_count(int64_t nc, const int64_t* __restrict__ c2f_ptr,
                               const int32_t* __restrict__ c2f_idx, const int64_t* __restrict__ rp,
                               const int32_t* __restrict__ cols, const uint32_t* __restrict__ agg,
                               int64_t* __restrict__ cnt) {
  const int64_t c = blockIdx.x * int64_t(NT) + threadIdx.x;
  if (c >= nc) return;
  int64_t k = 0;
  for (int64_t m = c2f_ptr[c]; m < c2f_ptr[c + 1]; ++m) {
    const int32_t v = c2f_idx[m];
    for (int64_t p = rp[v]; p < rp[v + 1]; ++p) k += agg[cols[p]] < 0xFFFFFFFEu;
  }
  cnt[c] = k;
}

__global__ void galerkin_fill(int64_t nc, const int64_t* __restrict__ c2f_ptr,
                              const int32_t* __restrict__ c2f_idx, const int64_t* __restrict__ rp,
                              const int32_t* __restrict__ cols, const double* __restrict__ vals,
                              const uint32_t* __restrict__ agg, const int64_t* __restrict__ off,
                              int coarse_f32, uint32_t* __restrict__ key,
                              double* __restrict__ val) {
  const int64_t c = blockIdx.x * int64_t(NT) + threadIdx.x;
  if (c >= nc) return;
  int64_t k = off[c];
  for (int64_t m = c2f_ptr[c]; m < c2f_ptr[c + 1]; ++m) {
    const int32_t v = c2f_idx[m];
    for (int64_t p = rp[v]; p < rp[v + 1]; ++p) {
      const uint32_t J = agg[cols[p]];
      if (J >= 0xFFFFFFFEu) continue;
      key[k] = J;
      val[k] = coarse_f32 ? double(float(vals[p])) : vals[p];
      ++k;
    }
  }
}

// stable insertion sort of each coarse row's sequence by J; distinct J count
__global__ void galerkin_sort(int64_t nc, const int64_t* __restrict__ off, uint32_t* __restrict__ key,
                              double* __restrict__ val, int64_t* __restrict__ distinct) {
  const int64_t c = blockIdx.x * int64_t(NT) + threadIdx.x;
  if (c >= nc) return;
  const int64_t b = off[c], e = off[c + 1];
  for (int64_t k = b + 1; k < e; ++k) {
    const uint32_t kk = key[k];
    const double kv = val[k];
    int64_t j = k - 1;
    while (j >= b && key[j] > kk) {
      key[j + 1] = key[j];
      val[j + 1] = val[j];
      --j;
    }
    key[j + 1] = kk;
    val[j + 1] = kv;
  }
  int64_t d = 0;
  for (int64_t k = b; k < e; ++k) d += (k == b || key[k] != key[k - 1]);
  distinct[c] = d;
}

__global__ void galerkin_fold(int64_t nc, const int64_t* __restrict__ off,
                              const uint32_t* __restrict__ key, const double* __restrict__ val,
                              const int64_t* __restrict__ crp, int coarse_f32,
                              int32_t* __restrict__ ccols, double* __restrict__ cvals) {
  const int64_t c = blockIdx.x * int64_t(NT) + threadIdx.x;
  if (c >= nc) return;
  const int64_t b = off[c], e = off[c + 1];
  int64_t o = crp[c];
  int64_t k = b;
  while (k < e) {
    const uint32_t J = key[k];
    double acc = val[k++];
    while (k < e && key[k] == J) acc += val[k++];
    ccols[o] = int32_t(J);
    cvals[o] = coarse_f32 ? double(float(acc)) : acc;
    ++o;
  }
}

// A level matrix on the device (f64 values; an f32 level holds values that
// f32 represents exactly, as the host's RowMatrix does).
struct DevMatrix {
  int64_t n = 0, nnz = 0;
  DevBuf rp, cols, vals, diag, dpos;
  void upload(const setup::RowMatrix& m) {
    n = m.n;
    nnz = int64_t(m.cols.size());
    rp.alloc(size_t(n + 1) * sizeof(int64_t));
    cols.alloc(size_t(nnz) * sizeof(int32_t));
    vals.alloc(size_t(nnz) * sizeof(double));
    LGPU_HIP_CHECK(hipMemcpy(rp.ptr, m.rowptr.data(), size_t(n + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    LGPU_HIP_CHECK(hipMemcpy(cols.ptr, m.cols.data(), size_t(nnz) * sizeof(int32_t), hipMemcpyHostToDevice));
    LGPU_HIP_CHECK(hipMemcpy(vals.ptr, m.vals.data(), size_t(nnz) * sizeof(double), hipMemcpyHostToDevice));
  }
  void find_diagonals(hipStream_t st) {
    diag.alloc(size_t(n) * sizeof(double));
    dpos.alloc(size_t(n) * sizeof(int64_t));
    if (n == 0) return;
    hipLaunchKernelGGL(diag_kernel, dim3(blocks_for(n)), dim3(NT), 0, st, n, rp.as<int64_t>(),
                       cols.as<int32_t>(), vals.as<double>(), diag.as<double>(), dpos.as<int64_t>());
    LGPU_HIP_CHECK(hipGetLastError());
  }
};

template <typename T>
void to_device(DevBuf& buf, const std::vector<T>& v) {
  buf.alloc(v.size() * sizeof(T));
  if (!v.empty()) LGPU_HIP_CHECK(hipMemcpy(buf.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}
// after the work queued on st (the context's stream is non-blocking, so a
// plain hipMemcpy does not wait for it)
template <typename T>
void to_host(std::vector<T>& v, const DevBuf& buf, size_t n, hipStream_t st) {
  v.resize(n);
  LGPU_HIP_CHECK(hipStreamSynchronize(st));
  if (n) LGPU_HIP_CHECK(hipMemcpy(v.data(), buf.ptr, n * sizeof(T), hipMemcpyDeviceToHost));
}

class DeviceSetup final : public setup::Accel {
 public:
  explicit DeviceSetup(hipStream_t st) : st_(st) {}

  void criterion(const setup::RowMatrix& m, double theta, setup::StrengthFlags& out) override {
    resident(m);
    const int64_t n = m.n, ne = int64_t(m.cols.size()) - n;
    DevBuf dep{size_t(ne)}, inf{size_t(ne)}, iso{size_t(n)};
    if (ne > 0) LGPU_HIP_CHECK(hipMemsetAsync(inf.ptr, 0, size_t(ne), st_));
    if (n > 0) {
      if (m.f32)
        hipLaunchKernelGGL(criterion_kernel<float>, dim3(blocks_for(n)), dim3(NT), 0, st_, n,
                           cur_.rp.as<int64_t>(), cur_.cols.as<int32_t>(), cur_.vals.as<double>(),
                           cur_.diag.as<double>(), cur_.dpos.as<int64_t>(), int(m.symmetric),
                           float(theta), dep.as<uint8_t>(), inf.as<uint8_t>(), iso.as<uint8_t>());
      else
        hipLaunchKernelGGL(criterion_kernel<double>, dim3(blocks_for(n)), dim3(NT), 0, st_, n,
                           cur_.rp.as<int64_t>(), cur_.cols.as<int32_t>(), cur_.vals.as<double>(),
                           cur_.diag.as<double>(), cur_.dpos.as<int64_t>(), int(m.symmetric),
                           theta, dep.as<uint8_t>(), inf.as<uint8_t>(), iso.as<uint8_t>());
      LGPU_HIP_CHECK(hipGetLastError());
    }
    LGPU_HIP_CHECK(hipStreamSynchronize(st_));
    to_host(out.depends, dep, size_t(ne), st_);
    to_host(out.influences, inf, size_t(ne), st_);
    to_host(out.isolated, iso, size_t(n), st_);
  }

  setup::RowMatrix galerkin(const setup::RowMatrix& fine, const setup::Level& lv,
                            bool coarse_f32) override {
    resident(fine);
    const int64_t nc = int64_t(lv.c2f_ptr.size()) - 1;
    setup::RowMatrix out;
    out.n = nc;
    out.f32 = coarse_f32;
    out.symmetric = fine.symmetric;
    DevBuf agg, cptr, cidx, cnt(size_t(nc + 1) * sizeof(int64_t));
    to_device(agg, lv.agg);
    to_device(cptr, lv.c2f_ptr);
    to_device(cidx, lv.c2f_idx);
    const dim3 g(blocks_for(nc)), b(NT);
    if (nc > 0) {
      hipLaunchKernelGGL(galerkin_count, g, b, 0, st_, nc, cptr.as<int64_t>(), cidx.as<int32_t>(),
                         cur_.rp.as<int64_t>(), cur_.cols.as<int32_t>(), agg.as<uint32_t>(),
                         cnt.as<int64_t>());
      LGPU_HIP_CHECK(hipGetLastError());
    }
    std::vector<int64_t> off;
    to_host(off, cnt, size_t(nc), st_);
    exclusive_scan(off);
    DevBuf doff, key(size_t(off.back()) * sizeof(uint32_t)), val(size_t(off.back()) * sizeof(double));
    DevBuf distinct(size_t(nc + 1) * sizeof(int64_t));
    to_device(doff, off);
    if (nc > 0) {
      hipLaunchKernelGGL(galerkin_fill, g, b, 0, st_, nc, cptr.as<int64_t>(), cidx.as<int32_t>(),
                         cur_.rp.as<int64_t>(), cur_.cols.as<int32_t>(), cur_.vals.as<double>(),
                         agg.as<uint32_t>(), doff.as<int64_t>(), int(coarse_f32),
                         key.as<uint32_t>(), val.as<double>());
      LGPU_HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(galerkin_sort, g, b, 0, st_, nc, doff.as<int64_t>(), key.as<uint32_t>(),
                         val.as<double>(), distinct.as<int64_t>());
      LGPU_HIP_CHECK(hipGetLastError());
    }
    to_host(out.rowptr, distinct, size_t(nc), st_);
    exclusive_scan(out.rowptr);
    const int64_t nnz = out.rowptr.back();
    DevMatrix next;
    next.n = nc;
    next.nnz = nnz;
    to_device(next.rp, out.rowptr);
    next.cols.alloc(size_t(nnz) * sizeof(int32_t));
    next.vals.alloc(size_t(nnz) * sizeof(double));
    if (nc > 0) {
      hipLaunchKernelGGL(galerkin_fold, g, b, 0, st_, nc, doff.as<int64_t>(), key.as<uint32_t>(),
                         val.as<double>(), next.rp.as<int64_t>(), int(coarse_f32),
                         next.cols.as<int32_t>(), next.vals.as<double>());
      LGPU_HIP_CHECK(hipGetLastError());
    }
    LGPU_HIP_CHECK(hipStreamSynchronize(st_));
    to_host(out.cols, next.cols, size_t(nnz), st_);
    to_host(out.vals, next.vals, size_t(nnz), st_);
    // the coarse matrix stays on the device for the next level's criterion
    next_ = std::move(next);
    has_next_ = true;
    return out;
  }

 private:
  // n + 1 entries: the exclusive prefix sums of the n counts
  static void exclusive_scan(std::vector<int64_t>& v) {
    v.push_back(0);
    int64_t s = 0;
    for (auto& x : v) {
      const int64_t t = x;
      x = s;
      s += t;
    }
  }

  // the device copy of m: the coarse matrix the last Galerkin product left,
  // or an upload
  void resident(const setup::RowMatrix& m) {
    if (cur_.n == m.n && cur_.nnz == int64_t(m.cols.size()) && cur_.rp.ptr && !has_next_) return;
    if (has_next_ && next_.n == m.n && next_.nnz == int64_t(m.cols.size())) {
      cur_ = std::move(next_);
    } else {
      cur_ = DevMatrix{};
      cur_.upload(m);
    }
    has_next_ = false;
    cur_.find_diagonals(st_);
  }

  hipStream_t st_;
  DevMatrix cur_, next_;
  bool has_next_ = false;
};

}  // namespace

std::unique_ptr<setup::Accel> make_device_setup(hipStream_t st) {
  return std::make_unique<DeviceSetup>(st);
}

}  // namespace lgpu
