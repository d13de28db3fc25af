// SpMV of a CSR matrix with B x B block values (CsrMatrix over
// fix::DenseMatrix<Real, B, B>, tensor/dynamic/csr/shared.hpp:37-1103,
// tensor/fixed/matrix.hpp:151).  MatrixVectorProductExpr's row path
// (component-wise/matrix-vector-product.hpp:71-82) folds the row's blocks in
// column order with compat::fmadd(A_ij, x_j, sum) = A_ij * x_j + sum
// (tensor/fixed/compat.hpp:24-33), the block product component r being the
// left fold 0 + a_r0 x_0 + a_r1 x_1 + ... (tensor/fixed/expressions.hpp:22-44),
// all in the work type f64; the sums are stored in the vector's type.
#include "internal.hpp"

namespace lgpu {

namespace {

constexpr int NT = 256;

template <typename V, int B>
__global__ __launch_bounds__(NT) void bcsr_spmv_kernel(int64_t nb, const int64_t* __restrict__ rp,
                                                       const int32_t* __restrict__ cols,
                                                       const V* __restrict__ vals,
                                                       const V* __restrict__ x, V* __restrict__ y) {
  const int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x;
  if (i >= nb) return;
  double sum[B];
#pragma unroll
  for (int r = 0; r < B; ++r) sum[r] = 0.0;
  for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
    const V* a = vals + p * (B * B);
    const V* xj = x + int64_t(cols[p]) * B;
    double xv[B];
#pragma unroll
    for (int c = 0; c < B; ++c) xv[c] = double(xj[c]);
#pragma unroll
    for (int r = 0; r < B; ++r) {
      double t = 0.0;
#pragma unroll
      for (int c = 0; c < B; ++c) t = t + double(a[r * B + c]) * xv[c];
      sum[r] = t + sum[r];
    }
  }
#pragma unroll
  for (int r = 0; r < B; ++r) y[i * B + r] = V(sum[r]);
}

template <typename V, int B>
void launch(const BcsrDesc& a, const void* x, void* y, hipStream_t st) {
  hipLaunchKernelGGL((bcsr_spmv_kernel<V, B>), dim3(unsigned((a.nb + NT - 1) / NT)), dim3(NT), 0, st,
                     a.nb, a.rowptr, a.cols, static_cast<const V*>(a.vals),
                     static_cast<const V*>(x), static_cast<V*>(y));
  LGPU_HIP_CHECK(hipGetLastError());
}

template <typename V>
void dispatch(const BcsrDesc& a, const void* x, void* y, hipStream_t st) {
  switch (a.block) {
    case 1: launch<V, 1>(a, x, y, st); break;
    case 2: launch<V, 2>(a, x, y, st); break;
    case 3: launch<V, 3>(a, x, y, st); break;
    case 4: launch<V, 4>(a, x, y, st); break;
    default: throw Error{LGPU_INVALID_ARG, "block size must lie in 1..4"};
  }
}

}  // namespace

void bcsr_spmv(const BcsrDesc& a, const void* x, void* y, hipStream_t st) {
  if (a.nb == 0) return;
  if (a.f32) dispatch<float>(a, x, y, st);
  else dispatch<double>(a, x, y, st);
}

}  // namespace lgpu
