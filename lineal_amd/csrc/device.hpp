// Device-side helpers: stencil row evaluation, reductions, XCD-aware tile mapping.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace lgpu {

// DPP whole-wave lane shifts (gfx9 wave_shr:1 / wave_shl:1): lane l receives
// the value of lane l-1 / l+1; the lane without a source (0 / 63) keeps `old`.
constexpr int kWaveShr1 = 0x138;
constexpr int kWaveShl1 = 0x130;

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ int shift(int old, int v) { return dpp_i32<CTRL>(old, v); }
template <int CTRL>
__device__ __forceinline__ uint8_t shift(uint8_t old, uint8_t v) {
  return uint8_t(dpp_i32<CTRL>(int(old), int(v)));
}
template <int CTRL>
__device__ __forceinline__ float shift(float old, float v) {
  return __int_as_float(dpp_i32<CTRL>(__float_as_int(old), __float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double shift(double old, double v) {
  const long long o = __double_as_longlong(old), b = __double_as_longlong(v);
  const int lo = dpp_i32<CTRL>(static_cast<int>(o), static_cast<int>(b));
  const int hi = dpp_i32<CTRL>(static_cast<int>(o >> 32), static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// Harmonic mean H(a,b) = (ab != 0) ? (ab + ab) / (a + b) : 0,
// math/harmonic-mean.hpp:14-25.
template <typename T>
__device__ __forceinline__ T hmean(T a, T b) {
  const T p = a * b;
  return p != T(0) ? (p + p) / (a + b) : T(0);
}

// n / d rounded to nearest, for operands that need no range scaling: a
// reciprocal seed refined by Newton-Raphson steps, then Markstein's remainder
// correction(s) (P. Markstein, IBM J. Res. Dev. 34(1), 1990).  These are the
// operations of gfx9's IEEE division sequence once its v_div_scale steps leave
// the operands unchanged (exponent gap below 768 / 96, no denormal operand,
// quotient or reciprocal) and its v_div_fixup has no special case to patch,
// so in that range the result equals IEEE division bit for bit.  Callers
// guarantee the range: f64 numerators and denominators within [2^-300, 2^300],
// f32 within [2^-50, 2^50].
__device__ __forceinline__ double div_rn_scaled_free(double n, double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  const double q = n * r;
  const double rem = fma(-d, q, n);
  return fma(rem, r, q);
}
__device__ __forceinline__ float div_rn_scaled_free(float n, float d) {
  const float r0 = __builtin_amdgcn_rcpf(d);
  const float e = fmaf(-d, r0, 1.0f);
  const float r = fmaf(e, r0, r0);
  const float q0 = n * r;
  const float rem0 = fmaf(-d, q0, n);
  const float q1 = fmaf(rem0, r, q0);
  const float rem1 = fmaf(-d, q1, n);
  return fmaf(rem1, r, q1);
}

// Harmonic mean evaluated without a branch.  FAST: both coefficients are +0
// or lie in [2^-100, 2^100] (f64) / [2^-24, 2^24] (f32), see
// StencilDesc::fast_div64 / fast_div32.  Then ab = 0 exactly when a or b is 0,
// a + b is 0 only when both are, and clamping the denominator to the range's
// lower end changes nothing else: 0 / max(a + b, lo) is the +0 the reference
// returns for ab = 0, and every other quotient is the IEEE one.
template <typename T, bool FAST>
__device__ __forceinline__ T hmean_sel(T a, T b) {
  const T p = a * b;
  if constexpr (FAST) {
    constexpr T lo = sizeof(T) == 8 ? T(0x1p-100) : T(0x1p-24);
    return div_rn_scaled_free(p + p, fmax(a + b, lo));
  } else {
    const T q = (p + p) / (a + b);
    return p != T(0) ? q : T(0);
  }
}

template <typename T>
__device__ __forceinline__ const T* coeff_ptr(const StencilDesc& s);
template <>
__device__ __forceinline__ const double* coeff_ptr<double>(const StencilDesc& s) { return s.c64; }
template <>
__device__ __forceinline__ const float* coeff_ptr<float>(const StencilDesc& s) { return s.c32; }

// Cell information of the two valuators: a coefficient (values valuator) or a
// material id (lookup valuator).
template <typename T, int K>
struct CellInfo;
template <typename T>
struct CellInfo<T, 0> {
  using Raw = T;
  __device__ __forceinline__ static Raw load(const StencilDesc& s, int64_t i) {
    return coeff_ptr<T>(s)[i];
  }
  __device__ __forceinline__ static T value(const StencilDesc&, Raw r) { return r; }
  // connection value H(c_i, c_j) * q_d
  // (tensor/dynamic/stencil/valuator/stationary-diffusion-values.hpp:131-135)
  __device__ __forceinline__ static T conn(const StencilDesc&, Raw a, Raw b, T qd, int) {
    return hmean(a, b) * qd;
  }
};
template <typename T>
struct CellInfo<T, 1> {
  using Raw = uint8_t;
  __device__ __forceinline__ static Raw load(const StencilDesc& s, int64_t i) { return s.mat[i]; }
  __device__ __forceinline__ static T value(const StencilDesc& s, Raw r) { return T(s.mc[r]); }
  // table gather T_d[m_i * 4 + m_j]
  // (tensor/dynamic/stencil/valuator/stationary-diffusion-lookup.hpp:129-136)
  __device__ __forceinline__ static T conn(const StencilDesc& s, Raw a, Raw b, T, int d);
};
template <>
__device__ __forceinline__ double CellInfo<double, 1>::conn(const StencilDesc& s, uint8_t a,
                                                           uint8_t b, double, int d) {
  return s.tab64[d][a * 4 + b];
}
template <>
__device__ __forceinline__ float CellInfo<float, 1>::conn(const StencilDesc& s, uint8_t a,
                                                         uint8_t b, float, int d) {
  return s.tab32[d][a * 4 + b];
}

// Wave-level sum of a double (64 lanes).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum; every thread passes its value, thread 0 receives the total.
template <int NT>
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[NT / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) t += red[w];
  }
  return t;
}

// Block sum for any block of up to 1024 threads (whole wavefronts); thread 0
// receives the total.
__device__ __forceinline__ double block_sum_dyn(double v) {
  __shared__ double red[16];
  v = wave_sum(v);
  const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0) {
    const int nw = int(blockDim.x / 64);
    for (int w = 0; w < nw; ++w) t += red[w];
  }
  return t;
}

// Map the hardware dispatch index to a logical tile index such that each of
// the 8 XCDs (which receive workgroups round robin) processes a contiguous
// range of logical tiles and so shares halo data through its own L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t hw, int64_t total) {
  constexpr int64_t kXcd = 8;
  const int64_t full = total - total % kXcd;
  if (hw >= full) return hw;
  const int64_t per = full / kXcd;
  return (hw % kXcd) * per + hw / kXcd;
}

// Dense direct solves of the coarsest level by one block (any number of
// threads): `s` is n doubles of LDS.  Cholesky with the dense lower factor,
// forward substitution column by column (for every row i the updates
// s_i -= L_ij aux_j happen for j = 0, 1, ... as in the row-oriented
// substitution of direct/substitution.hpp:27-52; zero entries leave s_i
// unchanged) and the transposed backward substitution, column oriented in
// the reference as well (direct/substitution.hpp:82-113).
template <typename T>
__device__ __forceinline__ void block_chol_solve(const double* L, int64_t n, const T* b, T* x,
                                                 double* s) {
  const int64_t tid = threadIdx.x, nt = blockDim.x;
  for (int64_t i = tid; i < n; i += nt) s[i] = double(b[i]);
  __syncthreads();
  for (int64_t j = 0; j < n; ++j) {
    const double aj = s[j] / L[j * n + j];
    __syncthreads();
    if (tid == 0) s[j] = aj;
    for (int64_t i = j + 1 + tid; i < n; i += nt) s[i] = s[i] - L[i * n + j] * aj;
    __syncthreads();
  }
  // s holds aux; backward: sol = aux; for j = n-1..0: sol_j /= L_jj; sol_i -= L_ji sol_j (i < j)
  for (int64_t j = n - 1; j >= 0; --j) {
    const double sj = s[j] / L[j * n + j];
    __syncthreads();
    if (tid == 0) s[j] = sj;
    for (int64_t i = tid; i < j; i += nt) s[i] = s[i] - L[j * n + i] * sj;
    __syncthreads();
  }
  for (int64_t i = tid; i < n; i += nt) x[i] = T(s[i]);
}

// Dense LU solve with a unit lower factor L and an upper factor U (f64, row
// major), column by column: forward, x_j is final once reached and is
// subtracted from the rows below; backward, x_j = s_j / u_jj is subtracted
// from the rows above (the oracle's order).
template <typename T>
__device__ __forceinline__ void block_lu_solve(const double* L, const double* U, int64_t n, const T* b,
                                               T* x, double* s) {
  const int64_t tid = threadIdx.x, nt = blockDim.x;
  for (int64_t i = tid; i < n; i += nt) s[i] = double(b[i]);
  __syncthreads();
  for (int64_t j = 0; j < n; ++j) {
    const double xj = s[j];
    for (int64_t i = j + 1 + tid; i < n; i += nt) s[i] = s[i] - L[i * n + j] * xj;
    __syncthreads();
  }
  for (int64_t j = n - 1; j >= 0; --j) {
    const double xj = s[j] / U[j * n + j];
    __syncthreads();
    if (tid == 0) s[j] = xj;
    for (int64_t i = tid; i < j; i += nt) s[i] = s[i] - U[i * n + j] * xj;
    __syncthreads();
  }
  for (int64_t i = tid; i < n; i += nt) x[i] = T(s[i]);
}

}  // namespace lgpu
