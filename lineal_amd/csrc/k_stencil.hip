// Matrix-free stencil passes: SpMV, residual, fused SpMV+dot, fused
// residual+norm, Jacobi, and the Dirichlet right-hand side.
//
// Layout: cell (x, y, z) at index (x*ny + y)*nz + z (z contiguous).  A block of
// 64 (z) x 4 (y) threads marches along x over XS planes, keeping the -x, centre
// and +x values of its own (y, z) column in registers, so every coefficient and
// vector element is fetched from HBM once (the +-y / +-z neighbours of the
// current plane are L1/L2 hits of data the same or an adjacent tile loads).
#include "device.hpp"

namespace lgpu {

namespace {

constexpr int BZ = 64;
constexpr int BY = 4;
constexpr int NT = BZ * BY;

int64_t march_len(const StencilDesc& s) {
  // enough planes per block to amortise the two extra boundary planes, enough
  // blocks to fill 256 CUs
  const int64_t tiles = ((s.nz + BZ - 1) / BZ) * ((s.ny + BY - 1) / BY);
  int64_t xs = 32;
  while (xs > 4 && tiles * ((s.nx + xs - 1) / xs) < 4096) xs /= 2;
  return xs;
}

template <typename T, int K, int MODE, bool HAS_X>
__global__ __launch_bounds__(NT) void stencil_kernel(StencilDesc s, const T* __restrict__ xv,
                                                     const T* __restrict__ bv, T* __restrict__ yv,
                                                     T omega, int64_t xs, int64_t tz, int64_t ty,
                                                     int64_t tx, double* __restrict__ partials) {
  using CI = CellInfo<T, K>;
  using Raw = typename CI::Raw;
  const int64_t hw = blockIdx.x;
  const int64_t total = tz * ty * tx;
  const int64_t tile = xcd_tile(hw, total);
  const int64_t bz = tile % tz, by = (tile / tz) % ty, bx = tile / (tz * ty);
  const int64_t z = bz * BZ + threadIdx.x % BZ;
  const int64_t y = by * BY + threadIdx.x / BZ;
  const int64_t x0 = bx * xs;
  const int64_t x1 = min(x0 + xs, s.nx);
  const int64_t nyz = s.ny * s.nz;
  const bool active = z < s.nz && y < s.ny;
  const T m1w = T(1) - omega;
  double part = 0.0;
  if (active) {
    const int64_t col = y * s.nz + z;
    // rolling registers along x
    Raw cm{}, cc, cp{};
    T xm = T(0), xc = T(0), xp = T(0);
    cc = CI::load(s, x0 * nyz + col);
    if (x0 > 0) cm = CI::load(s, (x0 - 1) * nyz + col);
    if (HAS_X) {
      xc = xv[x0 * nyz + col];
      if (x0 > 0) xm = xv[(x0 - 1) * nyz + col];
    }
    for (int64_t x = x0; x < x1; ++x) {
      const int64_t i = x * nyz + col;
      if (x + 1 < s.nx) {
        cp = CI::load(s, i + nyz);
        if (HAS_X) xp = xv[i + nyz];
      }
      Raw cmy{}, cpy{}, cmz{}, cpz{};
      T xmy = T(0), xpy = T(0), xmz = T(0), xpz = T(0);
      if (y > 0) {
        cmy = CI::load(s, i - s.nz);
        if (HAS_X) xmy = xv[i - s.nz];
      }
      if (y + 1 < s.ny) {
        cpy = CI::load(s, i + s.nz);
        if (HAS_X) xpy = xv[i + s.nz];
      }
      if (z > 0) {
        cmz = CI::load(s, i - 1);
        if (HAS_X) xmz = xv[i - 1];
      }
      if (z + 1 < s.nz) {
        cpz = CI::load(s, i + 1);
        if (HAS_X) xpz = xv[i + 1];
      }
      const Row<T> r = eval_row<T, K>(s, x, y, z, cc, cm, cmy, cmz, cpz, cpy, cp);
      if constexpr (MODE == SM_JACOBI) {
        // jacobi/sweep.hpp:66-87: accum = b - sum a_ij x_j (fnmadd with a_ij = -w)
        T acc = bv[i];
        acc = fma(r.wmx, xm, acc);
        acc = fma(r.wmy, xmy, acc);
        acc = fma(r.wmz, xmz, acc);
        acc = fma(r.wpz, xpz, acc);
        acc = fma(r.wpy, xpy, acc);
        acc = fma(r.wpx, xp, acc);
        const T quot = acc / r.d;
        yv[i] = fma(omega, quot, m1w * xc);
      } else {
        // matrix-vector-product.hpp:71-82: sum = fmadd(a_ij, x_j, sum) in iterate order
        T sum = T(0);
        sum = fma(-r.wmx, xm, sum);
        sum = fma(-r.wmy, xmy, sum);
        sum = fma(-r.wmz, xmz, sum);
        sum = fma(-r.wpz, xpz, sum);
        sum = fma(-r.wpy, xpy, sum);
        sum = fma(-r.wpx, xp, sum);
        sum = fma(r.d, xc, sum);
        if constexpr (MODE == SM_SPMV) {
          yv[i] = sum;
        } else if constexpr (MODE == SM_SPMV_DOT) {
          yv[i] = sum;
          part += double(xc) * double(sum);
        } else if constexpr (MODE == SM_RESID) {
          yv[i] = bv[i] - sum;
        } else {
          const T rr = bv[i] - sum;
          yv[i] = rr;
          part += double(rr) * double(rr);
        }
      }
      cm = cc;
      cc = cp;
      xm = xc;
      xc = xp;
    }
  }
  if constexpr (MODE == SM_SPMV_DOT || MODE == SM_RESID_NORM) {
    const double t = block_sum<NT>(part);
    if (threadIdx.x == 0) partials[hw] = t;
  }
}

template <typename T>
__global__ void rhs_kernel(StencilDesc s, T* b) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int64_t n = s.nx * s.ny * s.nz;
  if (i >= n) return;
  const int64_t x = i / (s.ny * s.nz);
  const T q0 = T(s.q[0]);
  T ci;
  if (s.kind == 0) ci = coeff_ptr<T>(s)[i];
  else ci = T(s.mc[s.mat[i]]);
  // tensor/dynamic/stencil/vector/shared.hpp:181-205
  T sum = T(0);
  if (x == 0) sum += ((T(2) * q0) * ci) * T(s.p_start);
  if (x + 1 == s.nx) sum += ((T(2) * q0) * ci) * T(s.p_end);
  b[i] = sum;
}

template <typename T, int K, int MODE>
void launch_mode(const StencilDesc& s, const void* x, const void* b, void* y, double omega,
                 double* partials, hipStream_t st) {
  const int64_t xs = march_len(s);
  const int64_t tz = (s.nz + BZ - 1) / BZ, ty = (s.ny + BY - 1) / BY, tx = (s.nx + xs - 1) / xs;
  const int64_t blocks = tz * ty * tx;
  if (x) {
    hipLaunchKernelGGL((stencil_kernel<T, K, MODE, true>), dim3(blocks), dim3(NT), 0, st, s,
                       static_cast<const T*>(x), static_cast<const T*>(b), static_cast<T*>(y),
                       T(omega), xs, tz, ty, tx, partials);
  } else {
    hipLaunchKernelGGL((stencil_kernel<T, K, MODE, false>), dim3(blocks), dim3(NT), 0, st, s,
                       static_cast<const T*>(x), static_cast<const T*>(b), static_cast<T*>(y),
                       T(omega), xs, tz, ty, tx, partials);
  }
  LGPU_HIP_CHECK(hipGetLastError());
}

template <typename T, int K>
void dispatch_mode(const StencilDesc& s, int mode, const void* x, const void* b, void* y,
                   double omega, double* partials, hipStream_t st) {
  switch (mode) {
    case SM_SPMV: launch_mode<T, K, SM_SPMV>(s, x, b, y, omega, partials, st); break;
    case SM_RESID: launch_mode<T, K, SM_RESID>(s, x, b, y, omega, partials, st); break;
    case SM_SPMV_DOT: launch_mode<T, K, SM_SPMV_DOT>(s, x, b, y, omega, partials, st); break;
    case SM_RESID_NORM: launch_mode<T, K, SM_RESID_NORM>(s, x, b, y, omega, partials, st); break;
    case SM_JACOBI: launch_mode<T, K, SM_JACOBI>(s, x, b, y, omega, partials, st); break;
    default: throw Error{LGPU_INVALID_ARG, "unknown stencil mode"};
  }
}

}  // namespace

int64_t stencil_partials(const StencilDesc& s) {
  const int64_t xs = march_len(s);
  return ((s.nz + BZ - 1) / BZ) * ((s.ny + BY - 1) / BY) * ((s.nx + xs - 1) / xs);
}

void stencil_apply(const StencilDesc& s, int mode, bool f32, const void* x, const void* b,
                   void* y, double omega, double* partials, hipStream_t st) {
  if (mode != SM_JACOBI) LGPU_REQUIRE(x != nullptr, "stencil pass needs an input vector");
  if (f32) {
    if (s.kind == 0) dispatch_mode<float, 0>(s, mode, x, b, y, omega, partials, st);
    else dispatch_mode<float, 1>(s, mode, x, b, y, omega, partials, st);
  } else {
    if (s.kind == 0) dispatch_mode<double, 0>(s, mode, x, b, y, omega, partials, st);
    else dispatch_mode<double, 1>(s, mode, x, b, y, omega, partials, st);
  }
}

void stencil_rhs(const StencilDesc& s, bool f32, void* b, hipStream_t st) {
  const int64_t n = s.nx * s.ny * s.nz;
  const int64_t blocks = (n + 255) / 256;
  if (f32) hipLaunchKernelGGL(rhs_kernel<float>, dim3(blocks), dim3(256), 0, st, s, static_cast<float*>(b));
  else hipLaunchKernelGGL(rhs_kernel<double>, dim3(blocks), dim3(256), 0, st, s, static_cast<double*>(b));
  LGPU_HIP_CHECK(hipGetLastError());
}

}  // namespace lgpu
