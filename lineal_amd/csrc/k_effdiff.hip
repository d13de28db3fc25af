// Effective diffusivity through the mid plane of a stencil system
// (component-wise/sink/effective-diffusion.hpp:49-335) and the index gather that
// compresses a full-length vector onto the rows a ZeroToZero CSR system keeps
// (tensor/dynamic/csr/create-from.hpp:197-229).
//
// The mid plane is x = nx / 2: its ny*nz cells i and their +x neighbours
// i + ny*nz.  One thread per plane cell; block sums go to partials that
// reduce_partials adds in a fixed order, so every call returns the same bits.
//
//   filled count  sum_i filled(i) + filled(i + ny nz)            (AreaPartSink)
//   flux sum      sum_i D(i, i + ny nz) (sol[i + ny nz] - sol[i]) (EffDiffSink)
//
// D is the valuator's diffusion coefficient without the axis quotient: the
// harmonic mean of the two per-cell coefficients, or the material factor
// table H(C[m], C[m']) of the lookup valuator.
#include "device.hpp"

namespace lgpu {
namespace {

constexpr int kThreads = 256;

int64_t plane_cells(const StencilDesc& s) { return s.ny * s.nz; }

__global__ __launch_bounds__(kThreads) void filled_count_kernel(StencilDesc s, EffDiffTables t,
                                                                double* __restrict__ partials) {
  const int64_t plane = s.ny * s.nz, base = (s.nx / 2) * plane;
  const int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  double v = 0.0;
  if (k < plane) {
    const int64_t i = base + k;
    if (s.kind == 1) v = t.filled[s.mat[i]] + t.filled[s.mat[i + plane]];
    else v = 2.0;
  }
  const double sum = block_sum<kThreads>(v);
  if (threadIdx.x == 0) partials[blockIdx.x] = sum;
}

template <typename T, bool MAPPED>
__global__ __launch_bounds__(kThreads) void flux_sum_kernel(StencilDesc s, EffDiffTables t,
                                                            const T* __restrict__ sol,
                                                            const int64_t* __restrict__ cmap,
                                                            double* __restrict__ partials) {
  const int64_t plane = s.ny * s.nz, base = (s.nx / 2) * plane;
  const int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  double v = 0.0;
  if (k < plane) {
    const int64_t i1 = base + k, i2 = i1 + plane;
    int64_t k1 = i1, k2 = i2;
    bool present = true;
    if constexpr (MAPPED) {
      k1 = cmap[i1];
      k2 = cmap[i2];
      present = k1 >= 0 && k2 >= 0;
    }
    if (present) {
      const double d = s.kind == 1 ? t.factor[s.mat[i1] * 4 + s.mat[i2]] : hmean(s.c64[i1], s.c64[i2]);
      const T diff = sol[k2] - sol[k1];
      v = d * double(diff);
    }
  }
  const double sum = block_sum<kThreads>(v);
  if (threadIdx.x == 0) partials[blockIdx.x] = sum;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void gather_kernel(const T* __restrict__ src,
                                                          const int64_t* __restrict__ idx,
                                                          int64_t n, T* __restrict__ dst) {
  for (int64_t r = int64_t(blockIdx.x) * kThreads + threadIdx.x; r < n;
       r += int64_t(gridDim.x) * kThreads)
    dst[r] = src[idx[r]];
}

}  // namespace

int64_t eff_diff_partials(const StencilDesc& s) {
  return (plane_cells(s) + kThreads - 1) / kThreads;
}

void eff_diff_filled(const StencilDesc& s, const EffDiffTables& t, double* partials,
                     hipStream_t st) {
  hipLaunchKernelGGL(filled_count_kernel, dim3(unsigned(eff_diff_partials(s))), dim3(kThreads), 0,
                     st, s, t, partials);
  LGPU_HIP_CHECK(hipGetLastError());
}

void eff_diff_flux(const StencilDesc& s, const EffDiffTables& t, const void* sol, bool f32,
                   const int64_t* cmap, double* partials, hipStream_t st) {
  const dim3 g(unsigned(eff_diff_partials(s))), b(kThreads);
  if (f32) {
    const float* x = static_cast<const float*>(sol);
    if (cmap) hipLaunchKernelGGL((flux_sum_kernel<float, true>), g, b, 0, st, s, t, x, cmap, partials);
    else hipLaunchKernelGGL((flux_sum_kernel<float, false>), g, b, 0, st, s, t, x, cmap, partials);
  } else {
    const double* x = static_cast<const double*>(sol);
    if (cmap) hipLaunchKernelGGL((flux_sum_kernel<double, true>), g, b, 0, st, s, t, x, cmap, partials);
    else hipLaunchKernelGGL((flux_sum_kernel<double, false>), g, b, 0, st, s, t, x, cmap, partials);
  }
  LGPU_HIP_CHECK(hipGetLastError());
}

void vec_gather(const void* src, bool f32, const int64_t* idx, int64_t n, void* dst,
                hipStream_t st) {
  if (n == 0) return;
  const dim3 g(unsigned(std::min<int64_t>((n + kThreads - 1) / kThreads, 8192))), b(kThreads);
  if (f32)
    hipLaunchKernelGGL(gather_kernel<float>, g, b, 0, st, static_cast<const float*>(src), idx, n,
                       static_cast<float*>(dst));
  else
    hipLaunchKernelGGL(gather_kernel<double>, g, b, 0, st, static_cast<const double*>(src), idx, n,
                       static_cast<double*>(dst));
  LGPU_HIP_CHECK(hipGetLastError());
}

}  // namespace lgpu
