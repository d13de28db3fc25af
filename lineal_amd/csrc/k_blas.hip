// Vector kernels: fills, conversions, deterministic two-pass reductions, the
// fused Krylov updates, AMG transfer operators and the coarsest direct solve.
#include "device.hpp"

namespace lgpu {

namespace {

constexpr int NT = 256;
constexpr int64_t kMaxBlocks = 2048;

int64_t grid_for(int64_t n) {
  const int64_t b = (n + NT - 1) / NT;
  return b < 1 ? 1 : (b > kMaxBlocks ? kMaxBlocks : b);
}

template <typename T>
__global__ void fill_kernel(T* x, int64_t n, T v) {
  for (int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * NT)
    x[i] = v;
}

template <typename S, typename D>
__global__ void convert_kernel(const S* s, D* d, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * NT)
    d[i] = D(s[i]);
}

template <typename T>
__global__ __launch_bounds__(NT) void dot_kernel(const T* a, const T* b, int64_t n, double* part) {
  double s = 0;
  for (int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * NT)
    s += double(a[i]) * double(b[i]);
  const double t = block_sum<NT>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// <a, b> with an f32 b: b widens exactly, so the result is the f64 dot of a
// with b's f64 copy, bit for bit (same grid, same order)
__global__ __launch_bounds__(NT) void dot_mixed_kernel(const double* a, const float* b, int64_t n,
                                                       double* part) {
  double s = 0;
  for (int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * NT)
    s += double(a[i]) * double(b[i]);
  const double t = block_sum<NT>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(NT) void dot2_kernel(const double* a, const double* b, int64_t n,
                                                  double* part) {
  double s0 = 0, s1 = 0;
  for (int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * NT) {
    const double ai = a[i];
    s0 += ai * b[i];
    s1 += ai * ai;
  }
  const double t0 = block_sum<NT>(s0);
  __syncthreads();
  const double t1 = block_sum<NT>(s1);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = t0;
    part[gridDim.x + blockIdx.x] = t1;
  }
}

// single block: fixed-order reduction of np partials
__global__ __launch_bounds__(NT) void reduce_kernel(const double* part, int64_t np, double* out,
                                                    const double* part2, double* out2) {
  double s = 0;
  for (int64_t i = threadIdx.x; i < np; i += NT) s += part[i];
  const double t = block_sum<NT>(s);
  if (threadIdx.x == 0) *out = t;
  if (part2) {
    __syncthreads();
    double s2 = 0;
    for (int64_t i = threadIdx.x; i < np; i += NT) s2 += part2[i];
    const double t2 = block_sum<NT>(s2);
    if (threadIdx.x == 0) *out2 = t2;
  }
}

// CG: alpha = scal[num] / scal[den]; x += alpha p; r -= alpha q; partial += r^2
// (conjugate-gradients/solver.hpp:84-95: sol + dir*alpha, res - prod*alpha)
__global__ __launch_bounds__(NT) void cg_update_kernel(double* __restrict__ x, double* __restrict__ r,
                                                       float* __restrict__ r32,
                                                       const double* __restrict__ p,
                                                       const double* __restrict__ q,
                                                       const double* __restrict__ scal, int num,
                                                       int den, int64_t n, double* __restrict__ part) {
  const double alpha = scal[num] / scal[den];
  double s = 0;
  const int64_t stride = int64_t(gridDim.x) * NT;
  int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x;
  // two elements per trip with all eight loads issued before the stores (the
  // arrays are distinct); the partial sum still takes the thread's elements
  // in order
  for (; i + stride < n; i += 2 * stride) {
    const int64_t j = i + stride;
    const double xa = x[i], pa = p[i], ra = r[i], qa = q[i];
    const double xb = x[j], pb = p[j], rb = r[j], qb = q[j];
    x[i] = xa + pa * alpha;
    x[j] = xb + pb * alpha;
    const double ua = ra - qa * alpha, ub = rb - qb * alpha;
    r[i] = ua;
    r[j] = ub;
    if (r32) {  // the FP32 V-cycle's input, as vec_convert rounds it
      r32[i] = float(ua);
      r32[j] = float(ub);
    }
    s += ua * ua;
    s += ub * ub;
  }
  if (i < n) {
    x[i] = x[i] + p[i] * alpha;
    const double ri = r[i] - q[i] * alpha;
    r[i] = ri;
    if (r32) r32[i] = float(ri);
    s += ri * ri;
  }
  const double t = block_sum<NT>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void axpy_scal_kernel(double* x, const double* p, const double* scal, int num, int den,
                                 double sign, int64_t n) {
  const double alpha = scal[num] / scal[den];
  for (int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * NT) {
    if (sign > 0) x[i] = x[i] + p[i] * alpha;
    else x[i] = x[i] - p[i] * alpha;
  }
}

// p = z + p * beta, beta = scal[num] / scal[den] (conjugate-gradients/solver.hpp:106-107)
template <typename Z>
__global__ void xpby_scal_kernel(double* p, const Z* z, const double* scal, int num, int den,
                                 int64_t n) {
  const double beta = scal[num] / scal[den];
  for (int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * NT)
    p[i] = double(z[i]) + p[i] * beta;
}

// bicgstab/solver.hpp:145-146: beta = (rho/old_rho) * (alpha/omega); p = r + (p - v*omega)*beta
__global__ void bicg_p_kernel(double* p, const double* r, const double* v, const double* scal,
                              int rho_new, int rho_old, int64_t n) {
  const double omega = scal[SC_OMEGA];
  const double beta = (scal[rho_new] / scal[rho_old]) * (scal[SC_ALPHA] / omega);
  for (int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * NT)
    p[i] = r[i] + (p[i] - v[i] * omega) * beta;
}

// r = r - v * a with a = scal[num]/scal[den]; the thread of block 0 stores a into scal[out]
__global__ __launch_bounds__(NT) void bicg_r_kernel(double* r, const double* v, double* scal,
                                                    int num, int den, int out, int64_t n,
                                                    double* part) {
  const double a = scal[num] / scal[den];
  double s = 0;
  for (int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * NT) {
    const double ri = r[i] - v[i] * a;
    r[i] = ri;
    s += ri * ri;
  }
  const double t = block_sum<NT>(s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void store_ratio_kernel(double* scal, int num, int den, int out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) scal[out] = scal[num] / scal[den];
}

// transport/vector-shared.hpp:49-56: sum in coarse->fine list order
template <typename T>
__global__ void restrict_kernel(const T* fine, const int64_t* ptr, const int32_t* idx, int64_t nc,
                                T* coarse) {
  for (int64_t c = blockIdx.x * int64_t(NT) + threadIdx.x; c < nc; c += int64_t(gridDim.x) * NT) {
    T s = T(0);
    for (int64_t k = ptr[c]; k < ptr[c + 1]; ++k) s += fine[idx[k]];
    coarse[c] = s;
  }
}

// The same sums with the block's list entries staged: the fine values of the
// 256 coarse rows' lists are gathered into LDS with coalesced index loads, then
// each thread adds its list in order.
template <typename T>
__global__ __launch_bounds__(NT) void restrict_staged_kernel(const T* fine, const int64_t* ptr,
                                                             const int32_t* idx, int64_t nc,
                                                             T* coarse) {
  extern __shared__ double restrict_lds[];
  T* const sv = reinterpret_cast<T*>(restrict_lds);
  const int64_t c0 = blockIdx.x * int64_t(NT);
  const int64_t c1 = c0 + NT < nc ? c0 + NT : nc;
  const int64_t e0 = ptr[c0], e1 = ptr[c1];
  // batches of 8 entries per thread: the index loads, then the gathers, in
  // flight together
  constexpr int B = 8;
  for (int64_t kb = e0 + threadIdx.x; kb < e1; kb += B * NT) {
    int32_t j[B];
    T v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) j[u] = idx[kb + u * NT < e1 ? kb + u * NT : e1 - 1];
#pragma unroll
    for (int u = 0; u < B; ++u) v[u] = fine[j[u]];
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (kb + u * NT < e1) sv[kb + u * NT - e0] = v[u];
  }
  __syncthreads();
  const int64_t c = c0 + threadIdx.x;
  if (c < nc) {
    T s = T(0);
    const int64_t k1 = ptr[c + 1] - e0;
    for (int64_t k = ptr[c] - e0; k < k1; ++k) s += sv[k];
    coarse[c] = s;
  }
}

// multigrid/solver.hpp:125-137 with transport/vector-shared.hpp:121-125
template <typename T>
__global__ void prolong_kernel(const T* x, const T* xc, const uint32_t* agg, int64_t n,
                               T inv_relax, T* out) {
  for (int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * NT) {
    const uint32_t g = agg[i];
    const T refined = g < 0xFFFFFFFEu ? xc[g] : T(0);
    out[i] = x[i] + refined * inv_relax;
  }
}

// Coarsest solves, one block (block_chol_solve / block_lu_solve)
template <typename T>
__global__ __launch_bounds__(NT) void chol_solve_kernel(const double* L, int64_t n, const T* b,
                                                        T* x) {
  extern __shared__ double sh[];
  block_chol_solve(L, n, b, x, sh);
}

template <typename T>
__global__ __launch_bounds__(NT) void lu_solve_kernel(const double* L, const double* U, int64_t n,
                                                      const T* b, T* x) {
  extern __shared__ double sh[];
  block_lu_solve(L, U, n, b, x, sh);
}

}  // namespace

void dense_lu_solve(const double* L, const double* U, int64_t n, const void* b, void* x, bool f32,
                    hipStream_t st) {
  LGPU_REQUIRE(n <= 4096, "coarsest level too large for the dense direct solve");
  const size_t sh = size_t(n) * sizeof(double);
  if (f32) hipLaunchKernelGGL(lu_solve_kernel<float>, dim3(1), dim3(NT), sh, st, L, U, n, static_cast<const float*>(b), static_cast<float*>(x));
  else hipLaunchKernelGGL(lu_solve_kernel<double>, dim3(1), dim3(NT), sh, st, L, U, n, static_cast<const double*>(b), static_cast<double*>(x));
  LGPU_HIP_CHECK(hipGetLastError());
}

void vec_fill(void* x, int64_t n, bool f32, double v, hipStream_t st) {
  if (f32) hipLaunchKernelGGL(fill_kernel<float>, dim3(grid_for(n)), dim3(NT), 0, st, static_cast<float*>(x), n, float(v));
  else hipLaunchKernelGGL(fill_kernel<double>, dim3(grid_for(n)), dim3(NT), 0, st, static_cast<double*>(x), n, v);
  LGPU_HIP_CHECK(hipGetLastError());
}

void vec_fill_int(int* x, int64_t n, int v, hipStream_t st) {
  hipLaunchKernelGGL(fill_kernel<int>, dim3(grid_for(n)), dim3(NT), 0, st, x, n, v);
  LGPU_HIP_CHECK(hipGetLastError());
}

void vec_convert(const void* src, bool src_f32, void* dst, bool dst_f32, int64_t n, hipStream_t st) {
  const dim3 g(unsigned(grid_for(n))), b(NT);
  if (src_f32 && dst_f32) hipLaunchKernelGGL((convert_kernel<float, float>), g, b, 0, st, static_cast<const float*>(src), static_cast<float*>(dst), n);
  else if (src_f32) hipLaunchKernelGGL((convert_kernel<float, double>), g, b, 0, st, static_cast<const float*>(src), static_cast<double*>(dst), n);
  else if (dst_f32) hipLaunchKernelGGL((convert_kernel<double, float>), g, b, 0, st, static_cast<const double*>(src), static_cast<float*>(dst), n);
  else hipLaunchKernelGGL((convert_kernel<double, double>), g, b, 0, st, static_cast<const double*>(src), static_cast<double*>(dst), n);
  LGPU_HIP_CHECK(hipGetLastError());
}

int64_t dot_partials(int64_t n) { return grid_for(n); }

void vec_dot(const void* a, const void* b, bool f32, int64_t n, double* partials, hipStream_t st) {
  if (f32) hipLaunchKernelGGL(dot_kernel<float>, dim3(grid_for(n)), dim3(NT), 0, st, static_cast<const float*>(a), static_cast<const float*>(b), n, partials);
  else hipLaunchKernelGGL(dot_kernel<double>, dim3(grid_for(n)), dim3(NT), 0, st, static_cast<const double*>(a), static_cast<const double*>(b), n, partials);
  LGPU_HIP_CHECK(hipGetLastError());
}

void vec_dot2(const double* a, const double* b, int64_t n, double* partials, hipStream_t st) {
  hipLaunchKernelGGL(dot2_kernel, dim3(grid_for(n)), dim3(NT), 0, st, a, b, n, partials);
  LGPU_HIP_CHECK(hipGetLastError());
}

void reduce_partials(const double* partials, int64_t np, double* out, hipStream_t st) {
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(NT), 0, st, partials, np, out,
                     static_cast<const double*>(nullptr), static_cast<double*>(nullptr));
  LGPU_HIP_CHECK(hipGetLastError());
}

void reduce_partials2(const double* partials, int64_t np, double* out0, double* out1,
                      hipStream_t st) {
  hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(NT), 0, st, partials, np, out0, partials + np, out1);
  LGPU_HIP_CHECK(hipGetLastError());
}

void cg_update(double* x, double* r, float* r32, const double* p, const double* q,
               const double* scal, int num, int den, int64_t n, double* partials, hipStream_t st) {
  hipLaunchKernelGGL(cg_update_kernel, dim3(grid_for(n)), dim3(NT), 0, st, x, r, r32, p, q, scal, num, den, n, partials);
  LGPU_HIP_CHECK(hipGetLastError());
}

void axpy_scal(double* x, const double* p, const double* scal, int num, int den, double sign,
               int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(axpy_scal_kernel, dim3(grid_for(n)), dim3(NT), 0, st, x, p, scal, num, den, sign, n);
  LGPU_HIP_CHECK(hipGetLastError());
}

void xpby_scal(double* p, const double* z, const double* scal, int num, int den, int64_t n,
               hipStream_t st) {
  hipLaunchKernelGGL(xpby_scal_kernel<double>, dim3(grid_for(n)), dim3(NT), 0, st, p, z, scal, num, den, n);
  LGPU_HIP_CHECK(hipGetLastError());
}

void xpby_scal(double* p, const float* z, const double* scal, int num, int den, int64_t n,
               hipStream_t st) {
  hipLaunchKernelGGL(xpby_scal_kernel<float>, dim3(grid_for(n)), dim3(NT), 0, st, p, z, scal, num, den, n);
  LGPU_HIP_CHECK(hipGetLastError());
}

void vec_dot_mixed(const double* a, const float* b, int64_t n, double* partials, hipStream_t st) {
  hipLaunchKernelGGL(dot_mixed_kernel, dim3(grid_for(n)), dim3(NT), 0, st, a, b, n, partials);
  LGPU_HIP_CHECK(hipGetLastError());
}

void bicg_p_update(double* p, const double* r, const double* v, const double* scal, int rho_new,
                   int rho_old, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(bicg_p_kernel, dim3(grid_for(n)), dim3(NT), 0, st, p, r, v, scal, rho_new, rho_old, n);
  LGPU_HIP_CHECK(hipGetLastError());
}

void bicg_r_update(double* r, const double* v, double* scal, int num, int den, int out, int64_t n,
                   double* partials, hipStream_t st) {
  hipLaunchKernelGGL(store_ratio_kernel, dim3(1), dim3(64), 0, st, scal, num, den, out);
  hipLaunchKernelGGL(bicg_r_kernel, dim3(grid_for(n)), dim3(NT), 0, st, r, v, scal, num, den, out, n, partials);
  LGPU_HIP_CHECK(hipGetLastError());
}

void restrict_sum(const void* fine, bool f32, const int64_t* ptr, const int32_t* idx, int64_t nc,
                  int64_t blk_entries, void* coarse, hipStream_t st) {
  static_assert(NT == kCsrBlockRows, "blk_entries counts blocks of NT coarse rows");
  const size_t lds = size_t(blk_entries) * (f32 ? sizeof(float) : sizeof(double));
  if (nc > 0 && blk_entries > 0 && lds <= 48 * 1024) {
    const dim3 g(unsigned((nc + NT - 1) / NT)), b(NT);
    if (f32)
      hipLaunchKernelGGL(restrict_staged_kernel<float>, g, b, lds, st, static_cast<const float*>(fine), ptr, idx, nc,
                         static_cast<float*>(coarse));
    else
      hipLaunchKernelGGL(restrict_staged_kernel<double>, g, b, lds, st, static_cast<const double*>(fine), ptr, idx,
                         nc, static_cast<double*>(coarse));
    LGPU_HIP_CHECK(hipGetLastError());
    return;
  }
  if (f32) hipLaunchKernelGGL(restrict_kernel<float>, dim3(grid_for(nc)), dim3(NT), 0, st, static_cast<const float*>(fine), ptr, idx, nc, static_cast<float*>(coarse));
  else hipLaunchKernelGGL(restrict_kernel<double>, dim3(grid_for(nc)), dim3(NT), 0, st, static_cast<const double*>(fine), ptr, idx, nc, static_cast<double*>(coarse));
  LGPU_HIP_CHECK(hipGetLastError());
}

void prolong_add(const void* x, const void* xc, bool f32, const uint32_t* agg, int64_t n,
                 double inv_relax, void* out, hipStream_t st) {
  if (f32) hipLaunchKernelGGL(prolong_kernel<float>, dim3(grid_for(n)), dim3(NT), 0, st, static_cast<const float*>(x), static_cast<const float*>(xc), agg, n, float(inv_relax), static_cast<float*>(out));
  else hipLaunchKernelGGL(prolong_kernel<double>, dim3(grid_for(n)), dim3(NT), 0, st, static_cast<const double*>(x), static_cast<const double*>(xc), agg, n, inv_relax, static_cast<double*>(out));
  LGPU_HIP_CHECK(hipGetLastError());
}

void dense_cholesky_solve(const double* L, int64_t n, const void* b, void* x, bool f32,
                          hipStream_t st) {
  LGPU_REQUIRE(n <= 4096, "coarsest level too large for the dense direct solve");
  const size_t sh = size_t(n) * sizeof(double);
  if (f32) hipLaunchKernelGGL(chol_solve_kernel<float>, dim3(1), dim3(NT), sh, st, L, n, static_cast<const float*>(b), static_cast<float*>(x));
  else hipLaunchKernelGGL(chol_solve_kernel<double>, dim3(1), dim3(NT), sh, st, L, n, static_cast<const double*>(b), static_cast<double*>(x));
  LGPU_HIP_CHECK(hipGetLastError());
}

}  // namespace lgpu
