// CSR passes (coarse AMG levels and the assembled finest operator): SpMV,
// residual, fused reductions, Jacobi, and lexicographic SOR.
//
// One thread per row accumulates the row in column order with fma, exactly as
// MatrixVectorProductExpr's row path (component-wise/matrix-vector-product.hpp:71-82).
//
// SOR: a sync-free wavefront over the row dependency graph.  Blocks of 256
// consecutive rows are claimed through an atomic ticket in sweep order; a row
// waits until every lower (forward) or higher (backward) neighbour row carries
// the current sweep's epoch in its flag, then computes, publishes its value and
// raises its own flag.  Dependencies only point to rows claimed earlier, so
// every wait ends; waits are bounded and report a timeout on the error word.
#include "device.hpp"

namespace lgpu {

namespace {

constexpr int NT = 256;
constexpr int SPIN_LIMIT = 1 << 22;

template <typename T, typename V, int MODE>
__global__ __launch_bounds__(NT) void csr_kernel(int64_t n, const int64_t* __restrict__ rp,
                                                 const int32_t* __restrict__ ci,
                                                 const V* __restrict__ va, const T* __restrict__ x,
                                                 const T* __restrict__ b, T* __restrict__ y,
                                                 T omega, double* __restrict__ partials) {
  const int64_t i = blockIdx.x * int64_t(NT) + threadIdx.x;
  double part = 0.0;
  if (i < n) {
    const int64_t k0 = rp[i], k1 = rp[i + 1];
    if constexpr (MODE == SM_JACOBI) {
      // jacobi/sweep.hpp:66-87
      T acc = b[i], d = T(1);
      for (int64_t k = k0; k < k1; ++k) {
        const int64_t j = ci[k];
        const T aij = T(va[k]);
        if (j == i) d = aij;
        else acc = fma(-aij, x ? x[j] : T(0), acc);
      }
      const T quot = acc / d;
      y[i] = fma(omega, quot, (T(1) - omega) * (x ? x[i] : T(0)));
    } else {
      // entries in groups of 8: the index, value and vector loads of a group
      // are independent, so they are in flight together; the sum still runs
      // in column order
      T sum = T(0);
      for (int64_t k = k0; k < k1; k += 8) {
        T av[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const bool ok = k + u < k1;
          const int32_t j = ok ? ci[k + u] : 0;
          av[u] = ok ? T(va[k + u]) : T(0);
          xv[u] = ok ? x[j] : T(0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (k + u < k1) sum = fma(av[u], xv[u], sum);
      }
      if constexpr (MODE == SM_SPMV) {
        y[i] = sum;
      } else if constexpr (MODE == SM_SPMV_DOT) {
        y[i] = sum;
        part = double(x[i]) * double(sum);
      } else if constexpr (MODE == SM_RESID) {
        y[i] = b[i] - sum;
      } else {
        const T r = b[i] - sum;
        y[i] = r;
        part = double(r) * double(r);
      }
    }
  }
  if constexpr (MODE == SM_SPMV_DOT || MODE == SM_RESID_NORM) {
    const double t = block_sum<NT>(part);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
  }
}

// The same passes with the block's entries staged: the entry range of the
// block's rows is read with coalesced loads into LDS, together with x at each
// entry's column, and every thread then runs its row from LDS.  One thread
// per row strides through memory by the row length, so its loads touch many
// more cache lines per instruction than they use.  The row arithmetic is
// csr_kernel's, in the same order.
template <typename T, typename V, int MODE>
__host__ __device__ constexpr size_t staged_entry_bytes() {
  return sizeof(T) + sizeof(V) + (MODE == SM_JACOBI ? sizeof(int32_t) : 0);
}

template <typename T, typename V, int MODE>
__global__ __launch_bounds__(NT) void csr_staged_kernel(int64_t n, const int64_t* __restrict__ rp,
                                                        const int32_t* __restrict__ ci,
                                                        const V* __restrict__ va,
                                                        const T* __restrict__ x,
                                                        const T* __restrict__ b,
                                                        T* __restrict__ y, T omega, int64_t cap,
                                                        double* __restrict__ partials) {
  extern __shared__ double staged_lds[];
  T* const sx = reinterpret_cast<T*>(staged_lds);
  V* const sv = reinterpret_cast<V*>(sx + cap);
  int32_t* const sc = reinterpret_cast<int32_t*>(sv + cap);
  const int64_t r0 = blockIdx.x * int64_t(NT);
  const int64_t r1 = min(r0 + NT, n);
  const int64_t e0 = rp[r0], e1 = rp[r1];
  for (int64_t k = e0 + threadIdx.x; k < e1; k += NT) {
    const int32_t j = ci[k];
    sv[k - e0] = va[k];
    sx[k - e0] = x ? x[j] : T(0);
    if constexpr (MODE == SM_JACOBI) sc[k - e0] = j;
  }
  __syncthreads();
  const int64_t i = r0 + threadIdx.x;
  double part = 0.0;
  if (i < n) {
    const int64_t k0 = rp[i] - e0, k1 = rp[i + 1] - e0;
    if constexpr (MODE == SM_JACOBI) {
      // jacobi/sweep.hpp:66-87
      T acc = b[i], d = T(1);
      for (int64_t k = k0; k < k1; ++k) {
        const T aij = T(sv[k]);
        if (sc[k] == i) d = aij;
        else acc = fma(-aij, sx[k], acc);
      }
      const T quot = acc / d;
      y[i] = fma(omega, quot, (T(1) - omega) * (x ? x[i] : T(0)));
    } else {
      T sum = T(0);
#pragma unroll 4
      for (int64_t k = k0; k < k1; ++k) sum = fma(T(sv[k]), sx[k], sum);
      if constexpr (MODE == SM_SPMV) {
        y[i] = sum;
      } else if constexpr (MODE == SM_SPMV_DOT) {
        y[i] = sum;
        part = double(x[i]) * double(sum);
      } else if constexpr (MODE == SM_RESID) {
        y[i] = b[i] - sum;
      } else {
        const T r = b[i] - sum;
        y[i] = r;
        part = double(r) * double(r);
      }
    }
  }
  if constexpr (MODE == SM_SPMV_DOT || MODE == SM_RESID_NORM) {
    const double t = block_sum<NT>(part);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
  }
}

// gauss-seidel-variant/sweep.hpp:134-200 with one segment covering all rows
template <typename T, typename V, bool FWD, bool ULTRA, bool HAS_X>
__global__ __launch_bounds__(NT) void csr_sor_kernel(int64_t n, const int64_t* __restrict__ rp,
                                                     const int32_t* __restrict__ ci,
                                                     const V* __restrict__ va,
                                                     const T* __restrict__ xin, T* xout,
                                                     const T* __restrict__ b, T omega, int* sched,
                                                     int* err, int* flags, int epoch) {
  __shared__ int s_ticket;
  if (threadIdx.x == 0) s_ticket = atomicAdd(sched, 1);
  __syncthreads();
  const int64_t nblk = (n + NT - 1) / NT;
  const int64_t blk = FWD ? s_ticket : nblk - 1 - s_ticket;
  // rows inside the block in sweep order
  const int64_t base = blk * NT;
  const int64_t i = FWD ? base + threadIdx.x : base + (NT - 1 - threadIdx.x);
  bool done = !(i >= 0 && i < n);
  const T m1w = T(1) - omega;
  int spins = 0;
  while (true) {
    if (!done) {
      const int64_t k0 = rp[i], k1 = rp[i + 1];
      bool ready = true;
      for (int64_t k = k0; k < k1 && ready; ++k) {
        const int64_t j = ci[k];
        if (FWD ? (j < i) : (j > i))
          ready = __hip_atomic_load(flags + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch;
      }
      if (!ready) {
        // after any expired wait every row proceeds, so the sweep still ends
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          ready = true;
        } else if (++spins > SPIN_LIMIT) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ready = true;
        }
      }
      if (ready) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const T xi = HAS_X ? xin[i] : T(0);
        T acc = b[i], d = T(1);
        for (int64_t k = k0; k < k1; ++k) {
          const int64_t j = ci[k];
          const T aij = T(va[k]);
          if (j == i) {
            d = aij;
            if (ULTRA) acc = fma(-aij, xi, acc);
            continue;
          }
          const bool use_new = FWD ? (j < i) : (j > i);
          T xj;
          if (use_new) xj = __hip_atomic_load(xout + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          else xj = HAS_X ? xin[j] : T(0);
          acc = fma(-aij, xj, acc);
        }
        T xn;
        if (ULTRA) {
          const T ad = d < T(0) ? -d : d;
          xn = xi + (omega * acc) / (ad + T(0.5) * T(0));
        } else {
          const T quot = acc / d;
          xn = fma(omega, quot, m1w * xi);
        }
        __hip_atomic_store(xout + i, xn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(flags + i, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        done = true;
      } else {
        __builtin_amdgcn_s_sleep(1);
      }
    }
    // the whole wave leaves together once all its rows are done
    if (__all(done)) break;
  }
}

// SOR with the reference's thread-segment semantics
// (gauss-seidel-variant/sweep.hpp:134-200): rows are cut into segments of
// `seg` consecutive rows; one thread sweeps one segment in order, using new
// values only for columns inside its own segment (kept in LDS) and old values
// elsewhere; the ultra variant adds |a_ij| of columns outside the segment to
// the divisor.
constexpr int SEG_NT = 128;

template <typename T, typename V, bool FWD, bool ULTRA, bool HAS_X>
__global__ __launch_bounds__(SEG_NT) void csr_sor_segment_kernel(
    int64_t n, const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
    const V* __restrict__ va, const T* __restrict__ xin, T* __restrict__ xout,
    const T* __restrict__ b, T omega, int seg) {
  extern __shared__ unsigned char smem[];
  T* fresh = reinterpret_cast<T*>(smem) + int64_t(threadIdx.x) * seg;
  const int64_t s = blockIdx.x * int64_t(SEG_NT) + threadIdx.x;
  const int64_t s0 = s * seg;
  if (s0 >= n) return;
  const int64_t s1 = min(s0 + seg, n);
  const T m1w = T(1) - omega;
  for (int64_t t = 0; t < s1 - s0; ++t) {
    const int64_t i = FWD ? s0 + t : s1 - 1 - t;
    const T xi = HAS_X ? xin[i] : T(0);
    T acc = b[i], d = T(1), div = T(0);
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
      const int64_t j = ci[k];
      const T aij = T(va[k]);
      if (j == i) {
        d = aij;
        if (ULTRA) acc = fma(-aij, xi, acc);
        continue;
      }
      const bool inside = j >= s0 && j < s1;
      if (ULTRA) div += inside ? T(0) : (aij < T(0) ? -aij : aij);
      const bool use_new = inside && (FWD ? j < i : j > i);
      const T xj = use_new ? fresh[j - s0] : (HAS_X ? xin[j] : T(0));
      acc = fma(-aij, xj, acc);
    }
    T xn;
    if (ULTRA) {
      const T ad = d < T(0) ? -d : d;
      xn = xi + (omega * acc) / (ad + T(0.5) * div);
    } else {
      const T quot = acc / d;
      xn = fma(omega, quot, m1w * xi);
    }
    fresh[i - s0] = xn;
    xout[i] = xn;
  }
}

// The same thread-segment SOR with segments of exactly 64 rows, one wavefront
// per segment and one lane per row.  The segment's entries (columns, values
// and the old values they refer to) are staged in LDS with coalesced loads;
// then every lane walks its row in column order and stops at the first entry
// that needs a new value from a row of its own segment that is not finished
// yet.  A ballot publishes the finished rows after every pass, so the passes
// resolve the segment's dependency chain while each row still accumulates in
// exactly the sequential order of the reference.
constexpr int WSEG = 64;
static_assert(WSEG == kMaxSegmentRows, "wave segments are the longest segments");
// LDS budget of one wave-segment block: what the device grants a block
size_t wave_lds_budget() {
  static const size_t budget = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      return size_t(64 * 1024);
    return size_t(v);
  }();
  return budget;
}

// One entry of a staged segment: column, value and the column's old value.
template <typename T, typename V>
struct Staged {
  int32_t col;
  T old;
  V val;
};

// The staged entries in LDS as three arrays, so no entry carries padding (a
// structure would take 16 bytes per f32 entry and 32 per f64 one): LDS is
// what limits how many segment wavefronts share a CU on the large levels.
// Layout: values[cap], old values[cap], the segment's new values[64],
// columns[cap], with cap even so that every array stays aligned.
template <typename T, typename V>
struct StageView {
  V* vals;
  T* olds;
  int32_t* cols;
  __device__ __forceinline__ Staged<T, V> operator[](int64_t k) const {
    return Staged<T, V>{cols[k], olds[k], vals[k]};
  }
  __device__ __forceinline__ void put(int64_t k, const Staged<T, V>& e) const {
    cols[k] = e.col;
    olds[k] = e.old;
    vals[k] = e.val;
  }
};

__host__ __device__ constexpr int64_t stage_cap(int64_t entries) { return (entries + 1) & ~int64_t(1); }

template <typename T, typename V>
size_t wave_lds_bytes(int64_t entries) {
  return size_t(stage_cap(entries)) * (sizeof(V) + sizeof(T) + sizeof(int32_t)) + WSEG * sizeof(T);
}

// One wavefront per segment of 64 rows.  The segment's entries (and the old
// values of their columns) are staged in LDS with coalesced loads; lane r owns
// row s0 + r.  In column order a row's entries split into
//   forward:  [other segments' rows before s0: old] [s0..i-1: new] [i] [after i: old]
//   backward: [before i: old] [i] [i+1..s1-1: new] [s1 and after: old]
// so each lane first applies its leading old-value entries, then the rounds of
// the wave loop let a lane apply its new-value entries once the rows they name
// are finished, and a lane whose new-value entries are all applied runs
// through its trailing old-value entries without further checks.  Every row
// performs the oracle's operations in its order (sweep.hpp:134-200 with the
// segment's own new values).
// Register path of the wave segment sweep.  In the prologue every lane takes
// its row apart in parallel: the leading old-value entries go into the
// accumulator at once, the new-value entries (at most kNewMax, one contiguous
// column range) become (value, segment index) pairs plus a bit mask of the
// rows they wait for, and the trailing old-value entries (at most kTrailMax)
// become (value, old value) pairs.  A round is then a mask test, independent
// LDS reads of the finished rows' values and the FMA chain, in the row's
// column order, so the results are those of the LDS-walking rounds below.
// Returns false (nothing written) when some row of the wave has longer lists.
// list sizes (multiples of 4) of the two variants: a level whose rows all fit
// the short lists runs a kernel with fewer registers, so more wavefronts
// share a SIMD
constexpr int kNewShort = 8, kTrailShort = 16;
constexpr int kNewLong = 16, kTrailLong = 24;

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
  return __builtin_amdgcn_readfirstlane(v);
}

template <typename T, typename V, bool FWD, bool ULTRA, bool HAS_X, int NM, int TM>
__device__ __forceinline__ bool wave_segment_registers(const StageView<T, V>& st, T* s_new,
                                                       int64_t s0, int64_t s1, int64_t e0,
                                                       const int64_t* __restrict__ rp,
                                                       const T* __restrict__ xin,
                                                       T* __restrict__ xout,
                                                       const T* __restrict__ b, T omega) {
  using S = Staged<T, V>;
  static_assert(NM % 4 == 0 && TM % 4 == 0, "lists run in groups of four");
  const int lane = threadIdx.x;
  const int64_t i = s0 + lane;
  const bool live = i < s1;
  T acc = T(0), d = T(1), div = T(0), xi = T(0);
  bool has_diag = false;
  V nval[NM], tval[TM];
  int nidx[NM];
  T told[TM];
  uint64_t deps = 0;
  int nn = 0, nt = 0;
  auto outside_div = [&](const S& e) {
    if constexpr (ULTRA) {
      if (e.col < s0 || e.col >= s1) {
        const T a = T(e.val);
        div += a < T(0) ? -a : a;
      }
    }
  };
  if (live) {
    int64_t k = rp[i] - e0;
    const int64_t kend = rp[i + 1] - e0;
    acc = b[i];
    xi = HAS_X ? xin[i] : T(0);
    // leading old-value entries: forward the columns below s0, backward below i
    const int64_t lead = FWD ? s0 : i;
    while (k < kend) {
      const S e = st[k];
      if (e.col >= lead) break;
      outside_div(e);
      acc = fma(-T(e.val), e.old, acc);
      ++k;
    }
    int64_t kn0 = k, kn1 = k;
    if constexpr (FWD) {
      while (k < kend && st[k].col < i) ++k;
      kn1 = k;
      if (k < kend && st[k].col == i) {
        d = T(st[k].val);
        has_diag = true;
        ++k;
      }
    } else {
      if (k < kend && st[k].col == i) {
        d = T(st[k].val);
        if constexpr (ULTRA) acc = fma(-d, xi, acc);
        ++k;
      }
      kn0 = k;
      while (k < kend && st[k].col < s1) ++k;
      kn1 = k;
    }
    nn = int(kn1 - kn0);
    nt = int(kend - k);
    if (nn <= NM && nt <= TM) {
#pragma unroll
      for (int u = 0; u < NM; ++u) {
        nval[u] = V(0);
        nidx[u] = lane;
        if (u < nn) {
          const S e = st[kn0 + u];
          nval[u] = e.val;
          nidx[u] = int(e.col - s0);
          deps |= uint64_t(1) << (e.col - s0);
        }
      }
#pragma unroll
      for (int u = 0; u < TM; ++u) {
        tval[u] = V(0);
        told[u] = T(0);
        if (u < nt) {
          const S e = st[k + u];
          tval[u] = e.val;
          told[u] = e.old;
          outside_div(e);
        }
      }
    }
  }
  if (__ballot(nn > NM || nt > TM)) return false;
  // Unused list slots hold value 0 with old value 0 (trailing) or point at the
  // lane's own row, which reads 0 until the lane finishes (new values), and
  // fma(-0, 0, a) == a for every a: padded slots leave the accumulator's bits
  // unchanged, so every lane runs the same chain.  Its length is the wave's
  // longest list, rounded up to groups of four.
  s_new[lane] = T(0);
  const int nn_w = wave_max_int(nn), nt_w = wave_max_int(nt);
  const T m1w = T(1) - omega;
  bool finished = !live;
  uint64_t done = __ballot(finished);
  while (done != ~uint64_t(0)) {
    if (!finished && (deps & ~done) == 0) {
      T nv[NM];
#pragma unroll
      for (int c = 0; c < NM; c += 4) {
        if (c < nn_w) {
#pragma unroll
          for (int u = c; u < c + 4; ++u) nv[u] = s_new[nidx[u]];
        }
      }
      T a = acc;
#pragma unroll
      for (int c = 0; c < NM; c += 4) {
        if (c < nn_w) {
#pragma unroll
          for (int u = c; u < c + 4; ++u) a = fma(-T(nval[u]), nv[u], a);
        }
      }
      if constexpr (FWD && ULTRA) {
        if (has_diag) a = fma(-d, xi, a);
      }
#pragma unroll
      for (int c = 0; c < TM; c += 4) {
        if (c < nt_w) {
#pragma unroll
          for (int u = c; u < c + 4; ++u) a = fma(-T(tval[u]), told[u], a);
        }
      }
      T xn;
      if (ULTRA) {
        const T ad = d < T(0) ? -d : d;
        xn = xi + (omega * a) / (ad + T(0.5) * div);
      } else {
        const T quot = a / d;
        xn = fma(omega, quot, m1w * xi);
      }
      s_new[lane] = xn;
      xout[i] = xn;
      finished = true;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    done = __ballot(finished);
  }
  return true;
}

template <typename T, typename V, bool FWD, bool ULTRA, bool HAS_X, bool STAGED, bool SHORT>
__global__ __launch_bounds__(WSEG) void csr_sor_wave_kernel(
    int64_t n, const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
    const V* __restrict__ va, const T* __restrict__ xin, T* __restrict__ xout,
    const T* __restrict__ b, T omega, int64_t cap) {
  using S = Staged<T, V>;
  extern __shared__ double wave_lds[];
  const int64_t scap = STAGED ? stage_cap(cap) : 0;
  V* const svals = reinterpret_cast<V*>(wave_lds);
  T* const solds = reinterpret_cast<T*>(svals + scap);
  T* const s_new = solds + scap;
  const StageView<T, V> st{svals, solds, reinterpret_cast<int32_t*>(s_new + WSEG)};
  const int lane = threadIdx.x;
  const int64_t s0 = blockIdx.x * int64_t(WSEG);
  const int64_t s1 = min(s0 + WSEG, n);
  const int64_t e0 = rp[s0], e1 = rp[s1];
  if constexpr (STAGED) {
    for (int64_t k = e0 + lane; k < e1; k += WSEG) {
      S e;
      e.col = ci[k];
      e.val = va[k];
      e.old = HAS_X ? xin[e.col] : T(0);
      st.put(k - e0, e);
    }
  }
  __syncthreads();
  if constexpr (STAGED) {
    constexpr int NM = SHORT ? kNewShort : kNewLong, TM = SHORT ? kTrailShort : kTrailLong;
    if (wave_segment_registers<T, V, FWD, ULTRA, HAS_X, NM, TM>(st, s_new, s0, s1, e0, rp, xin,
                                                                xout, b, omega))
      return;
  }
  auto entry = [&](int64_t k) {
    if constexpr (STAGED) {
      return st[k];
    } else {
      S e;
      e.col = ci[e0 + k];
      e.val = va[e0 + k];
      e.old = HAS_X ? xin[e.col] : T(0);
      return e;
    }
  };
  const int64_t i = s0 + lane;
  bool finished = i >= s1;
  int64_t k = 0, kend = 0;
  T acc = T(0), d = T(1), div = T(0), xi = T(0);
  // one old-value entry (ULTRA: entries outside the segment also add |a_ij|)
  auto apply_old = [&](const S& e) {
    const T a = T(e.val);
    if constexpr (ULTRA) {
      if (e.col < s0 || e.col >= s1) div += a < T(0) ? -a : a;
    }
    acc = fma(-a, e.old, acc);
  };
  auto apply_diag = [&](const S& e) {
    d = T(e.val);
    if constexpr (ULTRA) acc = fma(-d, xi, acc);
  };
  // old-value entries while the column stays below `lim`
  auto run_old = [&](int64_t lim) {
    while (k < kend) {
      const S e = entry(k);
      if (e.col >= lim) break;
      apply_old(e);
      ++k;
    }
  };
  if (!finished) {
    k = rp[i] - e0;
    kend = rp[i + 1] - e0;
    acc = b[i];
    xi = HAS_X ? xin[i] : T(0);
    if constexpr (FWD) {
      run_old(s0);
    } else {
      run_old(i);
      if (k < kend) {
        const S e = entry(k);
        if (e.col == i) {
          apply_diag(e);
          ++k;
        }
      }
    }
  }
  const T m1w = T(1) - omega;
  uint64_t done = __ballot(finished);
  while (done != ~uint64_t(0)) {
    if (!finished) {
      // new-value entries: columns in [s0, i) forward, (i, s1) backward
      const int64_t lim = FWD ? i : s1;
      bool ready = true;
      while (k < kend) {
        const S e = entry(k);
        if (e.col >= lim) break;
        if (!((done >> (e.col - s0)) & 1u)) {
          ready = false;  // wait for that row
          break;
        }
        acc = fma(-T(e.val), s_new[e.col - s0], acc);
        ++k;
      }
      if (ready) {
        if constexpr (FWD) {
          if (k < kend) {
            const S e = entry(k);
            if (e.col == i) {
              apply_diag(e);
              ++k;
            }
          }
        }
        // trailing old-value entries, four reads in flight
        while (k + 4 <= kend) {
          const S ea = entry(k), eb = entry(k + 1), ec = entry(k + 2), ed = entry(k + 3);
          apply_old(ea);
          apply_old(eb);
          apply_old(ec);
          apply_old(ed);
          k += 4;
        }
        while (k < kend) {
          apply_old(entry(k));
          ++k;
        }
        T xn;
        if (ULTRA) {
          const T ad = d < T(0) ? -d : d;
          xn = xi + (omega * acc) / (ad + T(0.5) * div);
        } else {
          const T quot = acc / d;
          xn = fma(omega, quot, m1w * xi);
        }
        s_new[lane] = xn;
        xout[i] = xn;
        finished = true;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    done = __ballot(finished);
  }
}

template <typename T, typename V, bool FWD, bool ULTRA, bool HAS_X>
void launch_wave_segment2(const CsrDesc& a, const void* x_in, void* x_out, const void* b,
                          double omega, hipStream_t st) {
  const int64_t nseg = (a.n + WSEG - 1) / WSEG;
  const size_t staged = wave_lds_bytes<T, V>(a.seg_entries);
  const auto xi = static_cast<const T*>(x_in);
  const auto xo = static_cast<T*>(x_out);
  const auto bp = static_cast<const T*>(b);
  const auto vp = static_cast<const V*>(a.vals);
  if (staged <= wave_lds_budget()) {
    if (a.seg_new_max <= kNewShort && a.seg_trail_max <= kTrailShort)
      hipLaunchKernelGGL((csr_sor_wave_kernel<T, V, FWD, ULTRA, HAS_X, true, true>), dim3(nseg),
                         dim3(WSEG), staged, st, a.n, a.rowptr, a.cols, vp, xi, xo, bp, T(omega),
                         a.seg_entries);
    else
      hipLaunchKernelGGL((csr_sor_wave_kernel<T, V, FWD, ULTRA, HAS_X, true, false>), dim3(nseg),
                         dim3(WSEG), staged, st, a.n, a.rowptr, a.cols, vp, xi, xo, bp, T(omega),
                         a.seg_entries);
  } else {
    const size_t direct = wave_lds_bytes<T, V>(0);
    hipLaunchKernelGGL((csr_sor_wave_kernel<T, V, FWD, ULTRA, HAS_X, false, false>), dim3(nseg),
                       dim3(WSEG), direct, st, a.n, a.rowptr, a.cols, vp, xi, xo, bp, T(omega),
                       int64_t(0));
  }
  LGPU_HIP_CHECK(hipGetLastError());
}

template <typename T, typename V, bool FWD, bool ULTRA>
void launch_wave_segment(const CsrDesc& a, const void* x_in, void* x_out, const void* b,
                         double omega, hipStream_t st) {
  if (x_in) launch_wave_segment2<T, V, FWD, ULTRA, true>(a, x_in, x_out, b, omega, st);
  else launch_wave_segment2<T, V, FWD, ULTRA, false>(a, x_in, x_out, b, omega, st);
}

template <typename T, typename V, bool FWD, bool ULTRA>
void launch_segment(const CsrDesc& a, const void* x_in, void* x_out, const void* b, double omega,
                    int seg, hipStream_t st) {
  const int64_t nseg = (a.n + seg - 1) / seg;
  const int64_t blocks = (nseg + SEG_NT - 1) / SEG_NT;
  const size_t sh = size_t(SEG_NT) * size_t(seg) * sizeof(T);
  if (x_in) {
    hipLaunchKernelGGL((csr_sor_segment_kernel<T, V, FWD, ULTRA, true>), dim3(blocks),
                       dim3(SEG_NT), sh, st, a.n, a.rowptr, a.cols, static_cast<const V*>(a.vals),
                       static_cast<const T*>(x_in), static_cast<T*>(x_out),
                       static_cast<const T*>(b), T(omega), seg);
  } else {
    hipLaunchKernelGGL((csr_sor_segment_kernel<T, V, FWD, ULTRA, false>), dim3(blocks),
                       dim3(SEG_NT), sh, st, a.n, a.rowptr, a.cols, static_cast<const V*>(a.vals),
                       static_cast<const T*>(x_in), static_cast<T*>(x_out),
                       static_cast<const T*>(b), T(omega), seg);
  }
  LGPU_HIP_CHECK(hipGetLastError());
}

template <typename T, typename V>
void dispatch_segment(const CsrDesc& a, const void* x_in, void* x_out, const void* b,
                      double omega, bool fwd, bool ultra, int seg, hipStream_t st) {
  if (seg == WSEG) {
    if (fwd) {
      if (ultra) launch_wave_segment<T, V, true, true>(a, x_in, x_out, b, omega, st);
      else launch_wave_segment<T, V, true, false>(a, x_in, x_out, b, omega, st);
    } else {
      if (ultra) launch_wave_segment<T, V, false, true>(a, x_in, x_out, b, omega, st);
      else launch_wave_segment<T, V, false, false>(a, x_in, x_out, b, omega, st);
    }
    return;
  }
  if (fwd) {
    if (ultra) launch_segment<T, V, true, true>(a, x_in, x_out, b, omega, seg, st);
    else launch_segment<T, V, true, false>(a, x_in, x_out, b, omega, seg, st);
  } else {
    if (ultra) launch_segment<T, V, false, true>(a, x_in, x_out, b, omega, seg, st);
    else launch_segment<T, V, false, false>(a, x_in, x_out, b, omega, seg, st);
  }
}

template <typename T, typename V, int MODE>
void launch_apply(const CsrDesc& a, const void* x, const void* b, void* y, double omega,
                  double* partials, hipStream_t st) {
  static_assert(NT == kCsrBlockRows, "blk_entries counts blocks of the pass");
  const int64_t blocks = (a.n + NT - 1) / NT;
  // an even capacity keeps the LDS arrays (T, then V, then i32) aligned
  const int64_t cap = (a.blk_entries + 1) & ~int64_t(1);
  const size_t lds = size_t(cap) * staged_entry_bytes<T, V, MODE>();
  if (a.blk_entries > 0 && lds <= wave_lds_budget()) {
    hipLaunchKernelGGL((csr_staged_kernel<T, V, MODE>), dim3(blocks), dim3(NT), lds, st, a.n,
                       a.rowptr, a.cols, static_cast<const V*>(a.vals), static_cast<const T*>(x),
                       static_cast<const T*>(b), static_cast<T*>(y), T(omega), cap, partials);
  } else {
    hipLaunchKernelGGL((csr_kernel<T, V, MODE>), dim3(blocks), dim3(NT), 0, st, a.n, a.rowptr,
                       a.cols, static_cast<const V*>(a.vals), static_cast<const T*>(x),
                       static_cast<const T*>(b), static_cast<T*>(y), T(omega), partials);
  }
  LGPU_HIP_CHECK(hipGetLastError());
}

template <typename T, typename V>
void dispatch_apply(const CsrDesc& a, int mode, const void* x, const void* b, void* y,
                    double omega, double* partials, hipStream_t st) {
  switch (mode) {
    case SM_SPMV: launch_apply<T, V, SM_SPMV>(a, x, b, y, omega, partials, st); break;
    case SM_RESID: launch_apply<T, V, SM_RESID>(a, x, b, y, omega, partials, st); break;
    case SM_SPMV_DOT: launch_apply<T, V, SM_SPMV_DOT>(a, x, b, y, omega, partials, st); break;
    case SM_RESID_NORM: launch_apply<T, V, SM_RESID_NORM>(a, x, b, y, omega, partials, st); break;
    case SM_JACOBI: launch_apply<T, V, SM_JACOBI>(a, x, b, y, omega, partials, st); break;
    default: throw Error{LGPU_INVALID_ARG, "unknown CSR mode"};
  }
}

template <typename T, typename V, bool FWD, bool ULTRA>
void launch_sor(const CsrDesc& a, const void* x_in, void* x_out, const void* b, double omega,
                int* sched, int* err, int* flags, int epoch, hipStream_t st) {
  const int64_t blocks = (a.n + NT - 1) / NT;
  if (x_in) {
    hipLaunchKernelGGL((csr_sor_kernel<T, V, FWD, ULTRA, true>), dim3(blocks), dim3(NT), 0, st,
                       a.n, a.rowptr, a.cols, static_cast<const V*>(a.vals),
                       static_cast<const T*>(x_in), static_cast<T*>(x_out),
                       static_cast<const T*>(b), T(omega), sched, err, flags, epoch);
  } else {
    hipLaunchKernelGGL((csr_sor_kernel<T, V, FWD, ULTRA, false>), dim3(blocks), dim3(NT), 0, st,
                       a.n, a.rowptr, a.cols, static_cast<const V*>(a.vals),
                       static_cast<const T*>(x_in), static_cast<T*>(x_out),
                       static_cast<const T*>(b), T(omega), sched, err, flags, epoch);
  }
  LGPU_HIP_CHECK(hipGetLastError());
}

template <typename T, typename V>
void dispatch_sor(const CsrDesc& a, const void* x_in, void* x_out, const void* b, double omega,
                  bool fwd, bool ultra, int* sched, int* err, int* flags, int epoch, hipStream_t st) {
  if (fwd) {
    if (ultra) launch_sor<T, V, true, true>(a, x_in, x_out, b, omega, sched, err, flags, epoch, st);
    else launch_sor<T, V, true, false>(a, x_in, x_out, b, omega, sched, err, flags, epoch, st);
  } else {
    if (ultra) launch_sor<T, V, false, true>(a, x_in, x_out, b, omega, sched, err, flags, epoch, st);
    else launch_sor<T, V, false, false>(a, x_in, x_out, b, omega, sched, err, flags, epoch, st);
  }
}

}  // namespace

int64_t csr_partials(const CsrDesc& a) { return (a.n + NT - 1) / NT; }

// Arithmetic precision follows the vectors: f64 vectors with f64 or f32
// values, f32 vectors with f32 values.
void csr_apply(const CsrDesc& a, int mode, const void* x, const void* b, void* y, double omega,
               double* partials, hipStream_t st) {
  if (a.n == 0) return;  // an empty system (every row compressed away): nothing to launch
  if (a.f32) dispatch_apply<float, float>(a, mode, x, b, y, omega, partials, st);
  else dispatch_apply<double, double>(a, mode, x, b, y, omega, partials, st);
}

void csr_sor_segmented(const CsrDesc& a, const void* x_in, void* x_out, const void* b,
                       double omega, bool forward, bool ultra, int seg, hipStream_t st) {
  LGPU_REQUIRE(x_in != x_out, "SOR sweep needs distinct input and output vectors");
  LGPU_REQUIRE(seg >= 1 && seg <= kMaxSegmentRows, "segment length out of range");
  if (a.n == 0) return;
  if (a.f32) dispatch_segment<float, float>(a, x_in, x_out, b, omega, forward, ultra, seg, st);
  else dispatch_segment<double, double>(a, x_in, x_out, b, omega, forward, ultra, seg, st);
}

void csr_sor(const CsrDesc& a, const void* x_in, void* x_out, const void* b, double omega,
             bool forward, bool ultra, int* sched, int* err, int* flags, int epoch, hipStream_t st) {
  LGPU_REQUIRE(x_in != x_out, "SOR sweep needs distinct input and output vectors");
  if (a.n == 0) return;
  vec_fill_int(sched, 1, 0, st);
  if (a.f32) dispatch_sor<float, float>(a, x_in, x_out, b, omega, forward, ultra, sched, err, flags, epoch, st);
  else dispatch_sor<double, double>(a, x_in, x_out, b, omega, forward, ultra, sched, err, flags, epoch, st);
}

}  // namespace lgpu
