// Outlet flux of a per-cell coefficient stencil, the right-hand side of its
// adjoint system, and its gradient with respect to the per-cell log
// coefficients (not in the reference).
//
//   J = sum over the x = nx-1 cells of (2 q_0 c_i)(p_i - p_end)
//   dJ/dlog c_k = c_k ( [x = nx-1] 2 q_0 (p_k - p_end)
//                     + lambda_k 2 q_0 ([x = 0](p_start - p_k) + [x = nx-1](p_end - p_k)) )
//               - sum over faces (k, j): H(c_k, c_j) q_d c_j / (c_k + c_j)
//                                        (lambda_k - lambda_j)(p_k - p_j)
// with A p = b and A lambda = dJ/dp (A symmetric without advection).  The
// border summands 2 q_0 c_i are the reference's
// (tensor/dynamic/stencil/valuator/stationary-diffusion-values.hpp:124-138).
// Every operation is the oracle's, in the oracle's order, so the gradient is
// bit-identical to it for the same p and lambda.
#include "device.hpp"

namespace lgpu {
namespace {

constexpr int kThreads = 256;
constexpr int kGradXChunk = 16;  // cells each thread marches along x

__global__ __launch_bounds__(kThreads) void outlet_flux_kernel(StencilDesc s,
                                                               const double* __restrict__ p,
                                                               double* __restrict__ partials) {
  const int64_t plane = s.ny * s.nz, base = (s.nx - 1) * plane;
  const int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const double b2 = 2.0 * s.q[0];
  double v = 0.0;
  if (t < plane) {
    const int64_t i = base + t;
    v = (b2 * s.c64[i]) * (p[i] - s.p_end);
  }
  const double sum = block_sum<kThreads>(v);
  if (threadIdx.x == 0) partials[blockIdx.x] = sum;
}

__global__ __launch_bounds__(kThreads) void adjoint_rhs_kernel(StencilDesc s,
                                                               double* __restrict__ g) {
  const int64_t plane = s.ny * s.nz, n = s.nx * plane, base = n - plane;
  const double b2 = 2.0 * s.q[0];
  for (int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * kThreads)
    g[i] = i >= base ? b2 * s.c64[i] : 0.0;
}

// One face's contribution, subtracted from the cell's accumulator.
__device__ __forceinline__ double face(double acc, double ck, double pk, double lk, double cj,
                                       double pj, double lj, double qd) {
  const double w = hmean(ck, cj) * qd;
  if (w == 0.0) return acc;
  const double t = w * (cj / (ck + cj));
  return acc - t * ((lk - lj) * (pk - pj));
}

// Each thread owns one (y, z) line and marches kGradXChunk cells along x,
// carrying the x-1 / x / x+1 values of c, p and lambda in registers; the y
// and z neighbours come from lines that adjacent lanes load as well.
__global__ __launch_bounds__(kThreads) void outlet_grad_kernel(StencilDesc s,
                                                               const double* __restrict__ p,
                                                               const double* __restrict__ lam,
                                                               double* __restrict__ grad) {
  const int64_t nyz = s.ny * s.nz;
  const int64_t line_blocks = (nyz + kThreads - 1) / kThreads;
  const int64_t cb = blockIdx.x % line_blocks, xc = blockIdx.x / line_blocks;
  const int64_t yz = cb * kThreads + threadIdx.x;
  if (yz >= nyz) return;
  const int64_t y = yz / s.nz, z = yz % s.nz;
  const int64_t x0 = xc * kGradXChunk;
  const int64_t x1 = x0 + kGradXChunk < s.nx ? x0 + kGradXChunk : s.nx;
  const double* c = s.c64;
  const double b2 = 2.0 * s.q[0];
  int64_t k = x0 * nyz + yz;
  double cm = 0, pm = 0, lm = 0;
  if (x0 > 0) {
    cm = c[k - nyz];
    pm = p[k - nyz];
    lm = lam[k - nyz];
  }
  double ck = c[k], pk = p[k], lk = lam[k];
  for (int64_t x = x0; x < x1; ++x, k += nyz) {
    const bool has_px = x + 1 < s.nx;
    double cp = 0, pp = 0, lp = 0;
    if (has_px) {
      cp = c[k + nyz];
      pp = p[k + nyz];
      lp = lam[k + nyz];
    }
    double acc = 0.0;
    if (x + 1 == s.nx) acc = b2 * (pk - s.p_end);
    if (x == 0) acc += lk * (b2 * (s.p_start - pk));
    if (x + 1 == s.nx) acc += lk * (b2 * (s.p_end - pk));
    acc = ck * acc;
    // faces in the row order -x, -y, -z, +z, +y, +x
    if (x > 0) acc = face(acc, ck, pk, lk, cm, pm, lm, s.q[0]);
    if (y > 0) acc = face(acc, ck, pk, lk, c[k - s.nz], p[k - s.nz], lam[k - s.nz], s.q[1]);
    if (z > 0) acc = face(acc, ck, pk, lk, c[k - 1], p[k - 1], lam[k - 1], s.q[2]);
    if (z + 1 < s.nz) acc = face(acc, ck, pk, lk, c[k + 1], p[k + 1], lam[k + 1], s.q[2]);
    if (y + 1 < s.ny) acc = face(acc, ck, pk, lk, c[k + s.nz], p[k + s.nz], lam[k + s.nz], s.q[1]);
    if (has_px) acc = face(acc, ck, pk, lk, cp, pp, lp, s.q[0]);
    grad[k] = acc;
    cm = ck, pm = pk, lm = lk;
    ck = cp, pk = pp, lk = lp;
  }
}

}  // namespace

int64_t outlet_flux_partials(const StencilDesc& s) {
  return (s.ny * s.nz + kThreads - 1) / kThreads;
}

void outlet_flux(const StencilDesc& s, const double* p, double* partials, hipStream_t st) {
  hipLaunchKernelGGL(outlet_flux_kernel, dim3(unsigned(outlet_flux_partials(s))), dim3(kThreads),
                     0, st, s, p, partials);
}

void outlet_flux_adjoint_rhs(const StencilDesc& s, double* g, hipStream_t st) {
  const int64_t n = s.nx * s.ny * s.nz;
  const int64_t blocks = std::min<int64_t>((n + kThreads - 1) / kThreads, 8192);
  hipLaunchKernelGGL(adjoint_rhs_kernel, dim3(unsigned(blocks)), dim3(kThreads), 0, st, s, g);
}

void outlet_flux_gradient(const StencilDesc& s, const double* p, const double* lam, double* grad,
                          hipStream_t st) {
  const int64_t line_blocks = (s.ny * s.nz + kThreads - 1) / kThreads;
  const int64_t chunks = (s.nx + kGradXChunk - 1) / kGradXChunk;
  hipLaunchKernelGGL(outlet_grad_kernel, dim3(unsigned(line_blocks * chunks)), dim3(kThreads), 0,
                     st, s, p, lam, grad);
}

}  // namespace lgpu
