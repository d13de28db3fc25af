// Lexicographic SOR on the matrix-free stencil as a tile-column wavefront.
//
// The reference's one-segment SOR (gauss-seidel-variant/sweep.hpp:134-200)
// updates cell i from the new values of its -x, -y, -z neighbours and the old
// values of its +x, +y, +z neighbours.  Here the (x, y) plane is cut into
// columns of 8 x 8 z-lines; one wavefront owns a column, lane (a, b) owns line
// (8I + a, 8J + b) and at step t computes z = t - a - b.  Its -x and -y
// neighbours were computed by lanes (a-1, b) and (a, b-1) one step earlier and
// arrive by cross-lane shuffles; its -z neighbour is its own previous value.
// Lines on the column border read the new values of the upstream columns
// (I-1, J) and (I, J-1) from memory after waiting on their progress counters.
// Columns are claimed through an atomic ticket in diagonal (I + J) order, so a
// column only ever waits on columns already claimed by running wavefronts:
// every wait ends.  The backward sweep is the same schedule in reflected
// coordinates.  Each cell performs exactly the oracle's operations in the
// oracle's order, so the result equals one-segment lexicographic SOR.
#include "device.hpp"

namespace lgpu {

namespace {

constexpr int CB = 8;        // lines per column side
constexpr int KCH = 16;      // steps between progress publications
constexpr int SPIN_LIMIT = 1 << 24;

__device__ __forceinline__ void wait_progress(int* state, int64_t col, int need) {
  int* progress = state + 2;
  int it = 0;
  while (__hip_atomic_load(progress + col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
    if (__hip_atomic_load(state + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    __builtin_amdgcn_s_sleep(1);
    if (++it > SPIN_LIMIT) {
      __hip_atomic_store(state + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
  }
}

template <typename T, int K, bool FWD, bool ULTRA, bool HAS_X>
__global__ __launch_bounds__(64) void sor_wavefront(StencilDesc s, const T* __restrict__ xin,
                                                    T* __restrict__ xout,
                                                    const T* __restrict__ bv, T omega,
                                                    int* state, const uint32_t* order,
                                                    int64_t cy) {
  using CI = CellInfo<T, K>;
  using Raw = typename CI::Raw;
  const int lane = threadIdx.x;
  const int a = lane / CB, b = lane % CB;
  __shared__ int s_ticket;
  if (lane == 0) s_ticket = atomicAdd(state, 1);
  __syncthreads();
  const int64_t col = order[s_ticket];
  const int64_t I = col / cy, J = col % cy;
  const int64_t lx = I * CB + a, ly = J * CB + b;
  const bool line = lx < s.nx && ly < s.ny;
  const int64_t x = FWD ? lx : s.nx - 1 - lx;
  const int64_t y = FWD ? ly : s.ny - 1 - ly;
  const int64_t nyz = s.ny * s.nz;
  const int64_t nsteps = s.nz + 2 * (CB - 1);
  const T m1w = T(1) - omega;
  int* progress = state + 2;
  T prev = T(0);
  for (int64_t t0 = 0; t0 < nsteps; t0 += KCH) {
    const int64_t tend = min(t0 + KCH, nsteps);
    const int need = int(min(tend + CB - 1, nsteps));
    if (lane == 0) {
      if (I > 0) wait_progress(state, (I - 1) * cy + J, need);
      if (J > 0) wait_progress(state, I * cy + J - 1, need);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __builtin_amdgcn_wave_barrier();
    for (int64_t t = t0; t < tend; ++t) {
      // new values of the logical -x / -y predecessors, computed one step ago
      const T nbx = __shfl(prev, lane >= CB ? lane - CB : lane, 64);
      const T nby = __shfl(prev, lane >= 1 ? lane - 1 : lane, 64);
      const int64_t lz = t - a - b;
      if (line && lz >= 0 && lz < s.nz) {
        const int64_t z = FWD ? lz : s.nz - 1 - lz;
        const int64_t i = (x * s.ny + y) * s.nz + z;
        const bool hmx = x > 0, hpx = x + 1 < s.nx;
        const bool hmy = y > 0, hpy = y + 1 < s.ny;
        const bool hmz = z > 0, hpz = z + 1 < s.nz;
        const Raw cc = CI::load(s, i);
        const Raw cmx = hmx ? CI::load(s, i - nyz) : Raw{};
        const Raw cpx = hpx ? CI::load(s, i + nyz) : Raw{};
        const Raw cmy = hmy ? CI::load(s, i - s.nz) : Raw{};
        const Raw cpy = hpy ? CI::load(s, i + s.nz) : Raw{};
        const Raw cmz = hmz ? CI::load(s, i - 1) : Raw{};
        const Raw cpz = hpz ? CI::load(s, i + 1) : Raw{};
        const Row<T> r = eval_row<T, K>(s, x, y, z, cc, cmx, cmy, cmz, cpz, cpy, cpx);
        // predecessor (new) values in logical coordinates
        T newx, newy;
        if (a > 0) newx = nbx;
        else newx = lx > 0 ? xout[FWD ? i - nyz : i + nyz] : T(0);
        if (b > 0) newy = nby;
        else newy = ly > 0 ? xout[FWD ? i - s.nz : i + s.nz] : T(0);
        const T newz = lz > 0 ? prev : T(0);
        // successor (old) values
        T oldx = T(0), oldy = T(0), oldz = T(0), xi = T(0);
        if (HAS_X) {
          xi = xin[i];
          if (FWD) {
            if (hpx) oldx = xin[i + nyz];
            if (hpy) oldy = xin[i + s.nz];
            if (hpz) oldz = xin[i + 1];
          } else {
            if (hmx) oldx = xin[i - nyz];
            if (hmy) oldy = xin[i - s.nz];
            if (hmz) oldz = xin[i - 1];
          }
        }
        const T vmx = FWD ? newx : oldx, vpx = FWD ? oldx : newx;
        const T vmy = FWD ? newy : oldy, vpy = FWD ? oldy : newy;
        const T vmz = FWD ? newz : oldz, vpz = FWD ? oldz : newz;
        // accum = b_i - sum_{j != i} a_ij x~_j, fnmadd in iterate order (a_ij = -w)
        T acc = bv[i];
        acc = fma(r.wmx, vmx, acc);
        acc = fma(r.wmy, vmy, acc);
        acc = fma(r.wmz, vmz, acc);
        acc = fma(r.wpz, vpz, acc);
        acc = fma(r.wpy, vpy, acc);
        acc = fma(r.wpx, vpx, acc);
        T xn;
        if (ULTRA) {
          acc = fma(-r.d, xi, acc);
          const T ad = r.d < T(0) ? -r.d : r.d;
          xn = xi + (omega * acc) / (ad + T(0.5) * T(0));
        } else {
          const T quot = acc / r.d;
          xn = fma(omega, quot, m1w * xi);
        }
        xout[i] = xn;
        prev = xn;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0)
      __hip_atomic_store(progress + col, int(tend), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename T, int K, bool FWD, bool ULTRA>
void launch4(const StencilDesc& s, const void* x_in, void* x_out, const void* b, double omega,
             const SorSched& sched, hipStream_t st) {
  const int64_t cy = (s.ny + CB - 1) / CB;
  const dim3 grid(unsigned(sched.ncols)), block(64);
  if (x_in) {
    hipLaunchKernelGGL((sor_wavefront<T, K, FWD, ULTRA, true>), grid, block, 0, st, s,
                       static_cast<const T*>(x_in), static_cast<T*>(x_out),
                       static_cast<const T*>(b), T(omega), sched.state, sched.order, cy);
  } else {
    hipLaunchKernelGGL((sor_wavefront<T, K, FWD, ULTRA, false>), grid, block, 0, st, s,
                       static_cast<const T*>(x_in), static_cast<T*>(x_out),
                       static_cast<const T*>(b), T(omega), sched.state, sched.order, cy);
  }
  LGPU_HIP_CHECK(hipGetLastError());
}

template <typename T, int K>
void launch2(const StencilDesc& s, const void* x_in, void* x_out, const void* b, double omega,
             bool fwd, bool ultra, const SorSched& sched, hipStream_t st) {
  if (fwd) {
    if (ultra) launch4<T, K, true, true>(s, x_in, x_out, b, omega, sched, st);
    else launch4<T, K, true, false>(s, x_in, x_out, b, omega, sched, st);
  } else {
    if (ultra) launch4<T, K, false, true>(s, x_in, x_out, b, omega, sched, st);
    else launch4<T, K, false, false>(s, x_in, x_out, b, omega, sched, st);
  }
}

}  // namespace

size_t sor_sched_ints(const StencilDesc& s) {
  const int64_t cx = (s.nx + CB - 1) / CB, cy = (s.ny + CB - 1) / CB;
  return size_t(2 + cx * cy);
}

// tickets in diagonal order: columns sorted by I + J, then by I
void sor_sched_order(const StencilDesc& s, std::vector<uint32_t>& order) {
  const int64_t cx = (s.nx + CB - 1) / CB, cy = (s.ny + CB - 1) / CB;
  order.clear();
  order.reserve(size_t(cx * cy));
  for (int64_t d = 0; d <= cx + cy - 2; ++d)
    for (int64_t I = 0; I < cx; ++I) {
      const int64_t J = d - I;
      if (J >= 0 && J < cy) order.push_back(uint32_t(I * cy + J));
    }
}

void stencil_sor(const StencilDesc& s, bool f32, const void* x_in, void* x_out, const void* b,
                 double omega, bool forward, bool ultra, const SorSched& sched, hipStream_t st) {
  LGPU_REQUIRE(x_in != x_out, "SOR sweep needs distinct input and output vectors");
  // reset the ticket and the progress counters; state[1] (the error word) is
  // sticky until the host clears it
  vec_fill_int(sched.state, 1, 0, st);
  vec_fill_int(sched.state + 2, sched.ncols, 0, st);
  if (f32) {
    if (s.kind == 0) launch2<float, 0>(s, x_in, x_out, b, omega, forward, ultra, sched, st);
    else launch2<float, 1>(s, x_in, x_out, b, omega, forward, ultra, sched, st);
  } else {
    if (s.kind == 0) launch2<double, 0>(s, x_in, x_out, b, omega, forward, ultra, sched, st);
    else launch2<double, 1>(s, x_in, x_out, b, omega, forward, ultra, sched, st);
  }
}

}  // namespace lgpu
