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
#include <algorithm>
#include <type_traits>

#include "device.hpp"

namespace lgpu {

namespace {

constexpr int NT = 256;
constexpr int SPIN_LIMIT = 1 << 22;
// entries a thread stages per batch into LDS: a batch's loads are issued
// together, so staging costs one memory round trip (two with the gathers)
// per batch rather than per entry
constexpr int kStageBatch = 8;
constexpr int kWaveStageBatch = 16;

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
  // the thread's row bounds and vector entries, requested before the staging
  // so that their memory round trip overlaps it (rows past n read row r0)
  const int64_t i = r0 + threadIdx.x;
  const int64_t ir = i < n ? i : r0;
  const int64_t rk0 = rp[ir], rk1 = rp[ir + 1];
  constexpr bool NEED_B = MODE == SM_RESID || MODE == SM_RESID_NORM || MODE == SM_JACOBI;
  constexpr bool NEED_X = MODE == SM_SPMV_DOT || MODE == SM_JACOBI;
  const T bi = NEED_B ? b[ir] : T(0);
  const T xi = NEED_X && x ? x[ir] : T(0);
  // batches of kStageBatch entries per thread: the column and value loads of a
  // batch, then its gathers, are all in flight together
  for (int64_t kb = e0 + threadIdx.x; kb < e1; kb += kStageBatch * NT) {
    int32_t j[kStageBatch];
    V v[kStageBatch];
    T xv[kStageBatch];
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int64_t k = kb + u * NT < e1 ? kb + u * NT : e1 - 1;
      j[u] = ci[k];
      v[u] = va[k];
    }
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) xv[u] = x ? x[j[u]] : T(0);
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int64_t k = kb + u * NT;
      if (k < e1) {
        sv[k - e0] = v[u];
        sx[k - e0] = xv[u];
        if constexpr (MODE == SM_JACOBI) sc[k - e0] = j[u];
      }
    }
  }
  __syncthreads();
  double part = 0.0;
  if (i < n) {
    const int64_t k0 = rk0 - e0, k1 = rk1 - e0;
    if constexpr (MODE == SM_JACOBI) {
      // jacobi/sweep.hpp:66-87
      T acc = bi, d = T(1);
      for (int64_t k = k0; k < k1; ++k) {
        const T aij = T(sv[k]);
        if (sc[k] == i) d = aij;
        else acc = fma(-aij, sx[k], acc);
      }
      const T quot = acc / d;
      y[i] = fma(omega, quot, (T(1) - omega) * xi);
    } else {
      T sum = T(0);
#pragma unroll 4
      for (int64_t k = k0; k < k1; ++k) sum = fma(T(sv[k]), sx[k], sum);
      if constexpr (MODE == SM_SPMV) {
        y[i] = sum;
      } else if constexpr (MODE == SM_SPMV_DOT) {
        y[i] = sum;
        part = double(xi) * double(sum);
      } else if constexpr (MODE == SM_RESID) {
        y[i] = bi - sum;
      } else {
        const T r = bi - sum;
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
constexpr int kNewShort = kSegNewShort, kTrailShort = kSegTrailShort;
constexpr int kNewLong = 16, kTrailLong = 24;

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
  return __builtin_amdgcn_readfirstlane(v);
}

// The lane's row bounds, right-hand side and old value, loaded before the
// segment's staging so that their memory round trip overlaps it.
template <typename T>
struct RowHead {
  int64_t k0, k1;
  T b, x;
};

template <typename T, typename V, bool FWD, bool ULTRA, bool HAS_X, int NM, int TM>
__device__ __forceinline__ bool wave_segment_registers(const StageView<T, V>& st, T* s_new,
                                                       int64_t s0, int64_t s1, int64_t e0,
                                                       const RowHead<T>& rh,
                                                       T* __restrict__ xout, T omega) {
  using S = Staged<T, V>;
  static_assert(NM % 4 == 0 && TM % 4 == 0, "lists run in groups of four");
  const int lane = int(threadIdx.x % WSEG);
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
    int64_t k = rh.k0 - e0;
    const int64_t kend = rh.k1 - e0;
    acc = rh.b;
    xi = HAS_X ? rh.x : T(0);
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
  // unchanged, so every lane runs the same chain.  The wave runs the shortest
  // of the chain lengths (4, 8), (8, 16), (NM, TM) its longest lists fit: a
  // round is a fixed FMA chain with no per-group test (tested groups came out
  // as every group evaluated plus a select each).
  s_new[lane] = T(0);
  const int nn_w = wave_max_int(nn), nt_w = wave_max_int(nt);
  const T m1w = T(1) - omega;
  auto rounds = [&](auto NMC, auto TMC) {
    constexpr int CN = decltype(NMC)::value, CT = decltype(TMC)::value;
    bool finished = !live;
    uint64_t done = __ballot(finished);
    while (done != ~uint64_t(0)) {
      if (!finished && (deps & ~done) == 0) {
        T nv[CN];
#pragma unroll
        for (int u = 0; u < CN; ++u) nv[u] = s_new[nidx[u]];
        T a = acc;
#pragma unroll
        for (int u = 0; u < CN; ++u) a = fma(-T(nval[u]), nv[u], a);
        if constexpr (FWD && ULTRA) {
          if (has_diag) a = fma(-d, xi, a);
        }
#pragma unroll
        for (int u = 0; u < CT; ++u) a = fma(-T(tval[u]), told[u], a);
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
  };
  static_assert(NM >= 8 && TM >= 16, "the chain tiers need lists of at least 8 and 16 slots");
  if (nn_w <= 4 && nt_w <= 8) rounds(std::integral_constant<int, 4>{}, std::integral_constant<int, 8>{});
  else if (nn_w <= 8 && nt_w <= 16) rounds(std::integral_constant<int, 8>{}, std::integral_constant<int, 16>{});
  else rounds(std::integral_constant<int, NM>{}, std::integral_constant<int, TM>{});
  return true;
}

// The sweep of one 64-row segment [s0, s0 + 64) by one wavefront, with its
// staging area at `lds` (wave_lds_bytes(cap)); the wavefronts of a block may
// sweep different segments, each in its own area.
template <typename T, typename V, bool FWD, bool ULTRA, bool HAS_X, bool STAGED, bool SHORT>
__device__ __forceinline__ void wave_sweep_segment(int64_t s0, int64_t n, const int64_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ci,
                                                   const V* __restrict__ va, const T* __restrict__ xin,
                                                   T* __restrict__ xout, const T* __restrict__ b, T omega,
                                                   int64_t cap, double* lds) {
  using S = Staged<T, V>;
  const int64_t scap = STAGED ? stage_cap(cap) : 0;
  V* const svals = reinterpret_cast<V*>(lds);
  T* const solds = reinterpret_cast<T*>(svals + scap);
  T* const s_new = solds + scap;
  const StageView<T, V> st{svals, solds, reinterpret_cast<int32_t*>(s_new + WSEG)};
  const int lane = int(threadIdx.x % WSEG);
  const int64_t s1 = min(s0 + WSEG, n);
  const int64_t e0 = rp[s0], e1 = rp[s1];
  const int64_t i = s0 + lane;
  const int64_t ir = i < s1 ? i : s0;
  const RowHead<T> rh{rp[ir], rp[ir + 1], b[ir], HAS_X ? xin[ir] : T(0)};
  if constexpr (STAGED) {
    // batches of kWaveStageBatch entries per lane (see csr_staged_kernel): a
    // segment's ~1000 entries in one or two batches
    constexpr int B = kWaveStageBatch;
    for (int64_t kb = e0 + lane; kb < e1; kb += B * WSEG) {
      S e[B];
#pragma unroll
      for (int u = 0; u < B; ++u) {
        const int64_t k = kb + u * WSEG < e1 ? kb + u * WSEG : e1 - 1;
        e[u].col = ci[k];
        e[u].val = va[k];
      }
#pragma unroll
      for (int u = 0; u < B; ++u) e[u].old = HAS_X ? xin[e[u].col] : T(0);
#pragma unroll
      for (int u = 0; u < B; ++u)
        if (kb + u * WSEG < e1) st.put(kb + u * WSEG - e0, e[u]);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if constexpr (STAGED) {
    constexpr int NM = SHORT ? kNewShort : kNewLong, TM = SHORT ? kTrailShort : kTrailLong;
    if (wave_segment_registers<T, V, FWD, ULTRA, HAS_X, NM, TM>(st, s_new, s0, s1, e0, rh, xout,
                                                                omega))
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
    k = rh.k0 - e0;
    kend = rh.k1 - e0;
    acc = rh.b;
    xi = rh.x;
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

// Scheduled segment sweep: the thread-segment SOR of any segment length (one
// segment for the whole level included), one workgroup per segment.  The host
// sorts each segment's rows by their depth in the segment's dependency graph
// (level 1: no new-value entry; a row's level is one more than the deepest
// row whose new value it reads), `pos` holding {row, level} per position.
// Round r computes the rows of level r, and a barrier closes every round, so
// each row finds the new values it reads in LDS.  Thread t takes positions
// s0 + t, s0 + t + SCHED_NT, ... in order and loads a position's row (its
// entries classified as old value, new value or diagonal, the old values
// gathered) as soon as it has finished the previous one, so the loads run
// while the rounds in between pass.  Each row performs the oracle's
// operations in its column order (gauss-seidel-variant/sweep.hpp:134-200
// with the segment's own new values): padded slots hold value 0 and old value
// 0, and fma(-0, 0, a) == a bit for bit.
constexpr int SCHED_NT = 256;

template <typename T>
struct BitsOf;
template <>
struct BitsOf<float> {
  using type = uint32_t;
};
template <>
struct BitsOf<double> {
  using type = uint64_t;
};

// One row of a scheduled sweep in registers: SLOTS entry slots (rows with more
// entries read them from memory when they run).
template <typename T, typename V, bool FWD, bool ULTRA, bool HAS_X, int SLOTS>
struct SchedRow {
  int64_t i = -1;   // row (-1: none)
  int lev = 0;
  int len = 0;
  T acc = T(0), d = T(1), div = T(0), xi = T(0);
  T val[SLOTS];
  T src[SLOTS];     // old value, or the LDS slot of a new value (bits)
  uint64_t newmask = 0;

  // the row's loads are issued together (no load waits on a branch), then
  // the entries are classified and the old values gathered
  __device__ __forceinline__ void load(int64_t ii, int lv, int64_t s0, int64_t s1,
                                       const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                       const V* __restrict__ va, const T* __restrict__ xin,
                                       const T* __restrict__ b) {
    using Bits = typename BitsOf<T>::type;
    i = ii;
    lev = lv;
    const int64_t k0 = rp[i];
    len = int(rp[i + 1] - k0);
    acc = b[i];
    xi = HAS_X ? xin[i] : T(0);
    d = T(1);
    div = T(0);
    newmask = 0;
    const int m = len <= SLOTS ? len : 0;
    // first the entries (src holds the column's bits until classified)
#pragma unroll
    for (int u = 0; u < SLOTS; ++u) {
      const int64_t k = k0 + (u < m ? u : 0);
      src[u] = __builtin_bit_cast(T, static_cast<Bits>(m > 0 ? ci[k] : int32_t(ii)));
      val[u] = m > 0 ? T(va[k]) : T(0);
    }
#pragma unroll
    for (int u = 0; u < SLOTS; ++u) {
      const int64_t j = int64_t(static_cast<int32_t>(__builtin_bit_cast(Bits, src[u])));
      const T a = val[u];
      const bool valid = u < m;
      const bool inside = j >= s0 && j < s1;
      const bool is_diag = valid && j == i;
      const bool is_new = valid && !is_diag && inside && (FWD ? j < i : j > i);
      const T old = HAS_X ? xin[j] : T(0);
      d = is_diag ? a : d;
      if (ULTRA) div += (valid && !is_diag && !inside) ? (a < T(0) ? -a : a) : T(0);
      // the diagonal's slot: for ultra acc - a_ii x_i at its place, else nothing
      val[u] = valid && (!is_diag || ULTRA) ? a : T(0);
      src[u] = is_new ? __builtin_bit_cast(T, static_cast<Bits>(j - s0))
                      : (is_diag ? (ULTRA ? xi : T(0)) : (valid ? old : T(0)));
      newmask |= uint64_t(is_new) << u;
    }
  }
};

template <typename T, typename V, bool FWD, bool ULTRA, bool HAS_X, int SLOTS>
__global__ __launch_bounds__(SCHED_NT) void csr_sor_sched_kernel(
    int64_t n, const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
    const V* __restrict__ va, const T* __restrict__ xin, T* __restrict__ xout,
    const T* __restrict__ b, T omega, int64_t seg, const int2* __restrict__ pos) {
  using Bits = typename BitsOf<T>::type;
  extern __shared__ double sched_lds[];
  T* const xnew = reinterpret_cast<T*>(sched_lds);
  const int64_t s0 = blockIdx.x * seg;
  const int64_t s1 = min(s0 + seg, n);
  // positions are sorted by level: the last one holds the segment's depth
  const int depth = pos[s1 - 1].y;
  const T m1w = T(1) - omega;
  SchedRow<T, V, FWD, ULTRA, HAS_X, SLOTS> row;
  int64_t p = s0 + threadIdx.x;
  if (p < s1) {
    const int2 q = pos[p];
    row.load(q.x, q.y, s0, s1, rp, ci, va, xin, b);
  }
  for (int r = 1; r <= depth; ++r) {
    while (row.i >= 0 && row.lev == r) {
      T a = row.acc, d = row.d, div = row.div;
      const int self = int(row.i - s0);
      if (row.len <= SLOTS) {
#pragma unroll
        for (int u = 0; u < SLOTS; u += 4) {
          if (u < row.len) {
#pragma unroll
            for (int w = u; w < u + 4; ++w) {
              const bool is_new = (row.newmask >> w) & 1u;
              const T nv = xnew[is_new ? int(__builtin_bit_cast(Bits, row.src[w])) : self];
              a = fma(-row.val[w], is_new ? nv : row.src[w], a);
            }
          }
        }
      } else {
        // a row longer than the register slots: its entries in column order
        // from memory
        const int64_t k0 = rp[row.i];
        for (int u = 0; u < row.len; ++u) {
          const int64_t j = ci[k0 + u];
          const T aij = T(va[k0 + u]);
          if (j == row.i) {
            d = aij;
            if (ULTRA) a = fma(-aij, row.xi, a);
            continue;
          }
          const bool inside = j >= s0 && j < s1;
          if (ULTRA && !inside) div += aij < T(0) ? -aij : aij;
          const bool use_new = inside && (FWD ? j < row.i : j > row.i);
          const T xj = use_new ? xnew[j - s0] : (HAS_X ? xin[j] : T(0));
          a = fma(-aij, xj, a);
        }
      }
      T xn;
      if (ULTRA) {
        const T ad = d < T(0) ? -d : d;
        xn = row.xi + (omega * a) / (ad + T(0.5) * div);
      } else {
        const T quot = a / d;
        xn = fma(omega, quot, m1w * row.xi);
      }
      xnew[self] = xn;
      xout[row.i] = xn;
      p += SCHED_NT;
      if (p < s1) {
        const int2 q = pos[p];
        row.load(q.x, q.y, s0, s1, rp, ci, va, xin, b);
      } else {
        row.i = -1;
      }
    }
    // a barrier that orders only the LDS: the next rows' loads stay in flight
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  }
}

// One segment for the whole level, swept by one wavefront with the level
// staged in LDS: the matrix (values, 16-bit columns), b, and the vector,
// which the sweep updates in place.  The schedule's levels order every row
// after the rows whose new values it reads and before the rows whose old
// values it reads, so x[j] holds exactly the value the sequential sweep
// reads.  Positions carry {row | level << 16, first entry | length << 24}.
// Lane t takes positions t, t + 64, ... and loads a position's entries into
// registers as soon as it has finished the previous one; a round is then
// the reads of x at those columns, the FMA chain in column order and the
// division.  Rows longer than the register slots read their entries from LDS.
constexpr int LDS_NT = 64;
constexpr int kLdsSlots = 32;

__host__ __device__ inline size_t lds_sweep_bytes(int64_t n, int64_t nnz, size_t es) {
  // (16 bytes of slack: a padded slot of an empty last row reads one entry
  // past the columns)
  return size_t(2 * n + 1 + nnz) * es + size_t(n) * sizeof(int2) + size_t(nnz) * sizeof(uint16_t) + 16;
}

template <typename T, bool FWD, bool ULTRA, bool HAS_X>
__global__ __launch_bounds__(LDS_NT) void csr_sor_lds_kernel(int64_t n, const int64_t* __restrict__ rp,
                                                           const int32_t* __restrict__ ci,
                                                           const T* __restrict__ va, const T* __restrict__ xin,
                                                           T* __restrict__ xout, const T* __restrict__ b,
                                                           T omega, const int2* __restrict__ pos) {
  extern __shared__ double lds_sweep[];
  const int lane = threadIdx.x;
  const int64_t e0 = rp[0], nnz = rp[n] - e0;
  T* const x = reinterpret_cast<T*>(lds_sweep);  // n + 1
  T* const bs = x + n + 1;
  T* const vals = bs + n;
  int2* const info = reinterpret_cast<int2*>(vals + nnz);
  uint16_t* const cols = reinterpret_cast<uint16_t*>(info + n);
  // staged in batches of kStage loads per lane, all issued before the first
  // store, so that each batch costs one memory round trip
  constexpr int kStage = 16;
  for (int64_t k0 = lane; k0 < nnz; k0 += kStage * LDS_NT) {
    T v[kStage];
    int32_t c[kStage];
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      const int64_t k = k0 + u * LDS_NT < nnz ? k0 + u * LDS_NT : nnz - 1;
      v[u] = va[e0 + k];
      c[u] = ci[e0 + k];
    }
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      const int64_t k = k0 + u * LDS_NT;
      if (k < nnz) {
        vals[k] = v[u];
        cols[k] = uint16_t(c[u]);
      }
    }
  }
  for (int64_t p0 = lane; p0 < n; p0 += kStage * LDS_NT) {
    T xv[kStage], bv[kStage];
    int2 q[kStage];
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      const int64_t p = p0 + u * LDS_NT < n ? p0 + u * LDS_NT : n - 1;
      xv[u] = HAS_X ? xin[p] : T(0);
      bv[u] = b[p];
      q[u] = pos[p];
    }
    int64_t r0[kStage], r1[kStage];
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      r0[u] = rp[q[u].x];
      r1[u] = rp[q[u].x + 1];
    }
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      const int64_t p = p0 + u * LDS_NT;
      if (p < n) {
        x[p] = xv[u];
        bs[p] = bv[u];
        info[p] = int2{q[u].x | (q[u].y << 16), int((r0[u] - e0) | ((r1[u] - r0[u]) << 24))};
      }
    }
  }
  if (lane == 0) x[n] = T(0);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  const int depth = info[n - 1].x >> 16;
  const T m1w = T(1) - omega;
  // the lane's next row: position, row, level, entries
  int64_t p = lane;
  int i = 0, lev = 0x7FFFFFFF, k0 = 0, len = 0;
  int col[kLdsSlots];
  T val[kLdsSlots];
  auto fetch = [&]() {
    if (p < n) {
      const int2 q = info[p];
      i = q.x & 0xFFFF;
      lev = q.x >> 16;
      k0 = q.y & 0xFFFFFF;
      len = int(uint32_t(q.y) >> 24);
    } else {
      lev = 0x7FFFFFFF;
      len = 0;
    }
    // every slot reads (the entry itself, or the row's first one), then
    // selects: no branch, so the reads are all in flight together
#pragma unroll
    for (int u = 0; u < kLdsSlots; ++u) {
      const int k = k0 + (u < len ? u : 0);
      const int c = int(cols[k]);
      const T v = vals[k];
      col[u] = u < len ? c : int(n);  // padding reads x[n] = 0 with value 0
      val[u] = u < len ? v : T(0);
    }
  };
  fetch();
  for (int r = 1; r <= depth; ++r) {
    while (lev == r) {
      T acc = bs[i], d = T(1);
      if (len <= kLdsSlots) {
        T xv[kLdsSlots];
#pragma unroll
        for (int u = 0; u < kLdsSlots; ++u) xv[u] = x[col[u]];
#pragma unroll
        for (int u = 0; u < kLdsSlots; u += 4) {
          if (u < len) {
#pragma unroll
            for (int w = u; w < u + 4; ++w) {
              // the diagonal: d, and for ultra acc - a_ii x_i at its place
              // (x[i] still holds the old value)
              const bool diag = col[w] == i;
              d = diag ? val[w] : d;
              const T t = fma(-val[w], xv[w], acc);
              acc = (diag && !ULTRA) ? acc : t;
            }
          }
        }
      } else {
        for (int k = k0; k < k0 + len; ++k) {
          const int j = cols[k];
          const T a = vals[k];
          if (j == i) {
            d = a;
            if (!ULTRA) continue;
          }
          acc = fma(-a, x[j], acc);
        }
      }
      const T xi = x[i];
      T xn;
      if (ULTRA) {
        const T ad = d < T(0) ? -d : d;
        xn = xi + (omega * acc) / (ad + T(0.5) * T(0));
      } else {
        const T quot = acc / d;
        xn = fma(omega, quot, m1w * xi);
      }
      x[i] = xn;
      p += LDS_NT;
      fetch();
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  for (int64_t q = lane; q < n; q += LDS_NT) xout[q] = x[q];
}

template <typename T, bool FWD, bool ULTRA, bool HAS_X>
void launch_lds3(const CsrDesc& a, const void* x_in, void* x_out, const void* b, double omega,
                 const int2* pos, hipStream_t st) {
  const size_t lds = lds_sweep_bytes(a.n, a.nnz, sizeof(T));
  hipLaunchKernelGGL((csr_sor_lds_kernel<T, FWD, ULTRA, HAS_X>), dim3(1), dim3(LDS_NT), lds, st, a.n,
                     a.rowptr, a.cols, static_cast<const T*>(a.vals), static_cast<const T*>(x_in),
                     static_cast<T*>(x_out), static_cast<const T*>(b), T(omega), pos);
  LGPU_HIP_CHECK(hipGetLastError());
}

template <typename T>
void dispatch_lds(const CsrDesc& a, const void* x_in, void* x_out, const void* b, double omega,
                  bool fwd, bool ultra, hipStream_t st) {
  const int2* pos = fwd ? a.sched_fwd : a.sched_bwd;
  if (fwd) {
    if (ultra) {
      if (x_in) launch_lds3<T, true, true, true>(a, x_in, x_out, b, omega, pos, st);
      else launch_lds3<T, true, true, false>(a, x_in, x_out, b, omega, pos, st);
    } else {
      if (x_in) launch_lds3<T, true, false, true>(a, x_in, x_out, b, omega, pos, st);
      else launch_lds3<T, true, false, false>(a, x_in, x_out, b, omega, pos, st);
    }
  } else {
    if (ultra) {
      if (x_in) launch_lds3<T, false, true, true>(a, x_in, x_out, b, omega, pos, st);
      else launch_lds3<T, false, true, false>(a, x_in, x_out, b, omega, pos, st);
    } else {
      if (x_in) launch_lds3<T, false, false, true>(a, x_in, x_out, b, omega, pos, st);
      else launch_lds3<T, false, false, false>(a, x_in, x_out, b, omega, pos, st);
    }
  }
}

template <typename T, typename V, bool FWD, bool ULTRA, bool HAS_X>
void launch_sched3(const CsrDesc& a, const void* x_in, void* x_out, const void* b, double omega,
                   const int2* pos, hipStream_t st) {
  const int64_t seg = a.sched_seg;
  const int64_t nseg = (a.n + seg - 1) / seg;
  const size_t lds = size_t(std::min(seg, a.n)) * sizeof(T);
  const auto xi = static_cast<const T*>(x_in);
  const auto xo = static_cast<T*>(x_out);
  const auto bp = static_cast<const T*>(b);
  const auto vp = static_cast<const V*>(a.vals);
  // rows longer than the slots read their entries when they run
  if (a.sched_max_len <= 16)
    hipLaunchKernelGGL((csr_sor_sched_kernel<T, V, FWD, ULTRA, HAS_X, 16>), dim3(nseg), dim3(SCHED_NT),
                       lds, st, a.n, a.rowptr, a.cols, vp, xi, xo, bp, T(omega), seg, pos);
  else
    hipLaunchKernelGGL((csr_sor_sched_kernel<T, V, FWD, ULTRA, HAS_X, 32>), dim3(nseg), dim3(SCHED_NT),
                       lds, st, a.n, a.rowptr, a.cols, vp, xi, xo, bp, T(omega), seg, pos);
  LGPU_HIP_CHECK(hipGetLastError());
}

template <typename T, typename V>
void dispatch_sched(const CsrDesc& a, const void* x_in, void* x_out, const void* b, double omega,
                    bool fwd, bool ultra, hipStream_t st) {
  const int2* pos = fwd ? a.sched_fwd : a.sched_bwd;
  if (fwd) {
    if (ultra) {
      if (x_in) launch_sched3<T, V, true, true, true>(a, x_in, x_out, b, omega, pos, st);
      else launch_sched3<T, V, true, true, false>(a, x_in, x_out, b, omega, pos, st);
    } else {
      if (x_in) launch_sched3<T, V, true, false, true>(a, x_in, x_out, b, omega, pos, st);
      else launch_sched3<T, V, true, false, false>(a, x_in, x_out, b, omega, pos, st);
    }
  } else {
    if (ultra) {
      if (x_in) launch_sched3<T, V, false, true, true>(a, x_in, x_out, b, omega, pos, st);
      else launch_sched3<T, V, false, true, false>(a, x_in, x_out, b, omega, pos, st);
    } else {
      if (x_in) launch_sched3<T, V, false, false, true>(a, x_in, x_out, b, omega, pos, st);
      else launch_sched3<T, V, false, false, false>(a, x_in, x_out, b, omega, pos, st);
    }
  }
}

template <typename T, typename V, bool FWD, bool ULTRA, bool HAS_X, bool STAGED, bool SHORT>
__global__ __launch_bounds__(WSEG) void csr_sor_wave_kernel(
    int64_t n, const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
    const V* __restrict__ va, const T* __restrict__ xin, T* __restrict__ xout,
    const T* __restrict__ b, T omega, int64_t cap) {
  extern __shared__ double wave_lds[];
  wave_sweep_segment<T, V, FWD, ULTRA, HAS_X, STAGED, SHORT>(blockIdx.x * int64_t(WSEG), n, rp, ci, va,
                                                           xin, xout, b, omega, cap, wave_lds);
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

// Round programs (RoundSweep): sections 16-byte aligned.
__host__ __device__ constexpr int64_t round16(int64_t bytes) { return (bytes + 15) & ~int64_t(15); }

// Block-wide copy of `bytes` (a multiple of 16) with 16-byte loads, eight
// per thread in flight before their stores.
__device__ __forceinline__ void block_copy16(void* dst, const void* src, int64_t bytes) {
  using V4 = unsigned int __attribute__((ext_vector_type(4)));
  const V4* s = static_cast<const V4*>(src);
  V4* d = static_cast<V4*>(dst);
  const int64_t nch = bytes / 16, nt = blockDim.x;
  constexpr int B = 8;
  for (int64_t k0 = threadIdx.x; k0 < nch; k0 += B * nt) {
    V4 v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int64_t k = k0 + u * nt;
      v[u] = s[k < nch ? k : nch - 1];
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int64_t k = k0 + u * nt;
      if (k < nch) d[k] = v[u];
    }
  }
}

// One-segment SOR of a small level by its round program (RoundSweep), by
// the whole workgroup: x (x_in or 0, and x[n] = 0), b and the program's
// staging region are copied to LDS; wavefront 0 then runs the rounds, lane j
// of a round computing the round's j-th row in place in x: b_i, then
// fma(-a_ij, x_j, acc) over the row's entries in column order (the diagonal
// at its column for ultra, else reading x[n] = 0 with its value), then the
// regular or ultra update of the oracle's segment sweep (one segment, so no
// off-segment entries).  The rows of a round read only values the rounds
// before finished or that no round before changed, which is what the
// sequential sweep reads.  Last, the workgroup writes x out.
//
// A round's critical path is its gathers of x, the FMA chain and the
// division.  Everything else a round needs does not depend on x (header,
// row, diagonal, right-hand side, the first kRegSlots columns and values),
// so it is read one round ahead into the other of two register buffers,
// after the current round's gathers have been issued.
template <typename T, int SLOTS>
struct RoundRegs {
  static constexpr int kSlots = SLOTS;  // entries held in registers
  using U4 = unsigned int __attribute__((ext_vector_type(4)));
  using TV = T __attribute__((ext_vector_type(16 / sizeof(T))));
  static constexpr int kPer = 16 / int(sizeof(T));
  int4 h;
  int i, e;
  bool act;
  T d, b;
  U4 c[kSlots / 4];
  TV v[kSlots / kPer];
  __device__ __forceinline__ unsigned col(int u) const { return c[u / 4][u % 4]; }
  __device__ __forceinline__ T val(int u) const { return v[u / kPer][u % kPer]; }
};

template <typename T, bool ULTRA, bool HAS_X, int SLOTS>
__device__ void round_sweep_slots(const RoundSweep& g, int64_t n64, const T* __restrict__ xin,
                                  T* __restrict__ xout, const T* __restrict__ b, T omega, double* lds) {
  using R = RoundRegs<T, SLOTS>;
  constexpr int RS = R::kSlots;
  const int n = int(n64);
  const int64_t E = g.entries + kRoundPad;
  char* const base = reinterpret_cast<char*>(lds);
  T* const x = reinterpret_cast<T*>(base);  // at byte 0: column byte offsets address it directly
  T* const bs = reinterpret_cast<T*>(base + round16((n + 1) * int64_t(sizeof(T))));
  char* const stage = base + round16((n + 1) * int64_t(sizeof(T))) + round16(n * int64_t(sizeof(T)));
  const int32_t* const rows = reinterpret_cast<const int32_t*>(stage);
  const T* const dg = reinterpret_cast<const T*>(stage + round16(n * 4));
  const T* const vals = reinterpret_cast<const T*>(stage + round16(n * 4) + round16(n * int64_t(sizeof(T))));
  const int4* const hdr = reinterpret_cast<const int4*>(vals + E);
  const int64_t stage_bytes = round_stage_bytes(n, g.entries, g.nrounds, sizeof(T));
  uint32_t* const cols = reinterpret_cast<uint32_t*>(stage + stage_bytes);
  // the staging region follows the rows pointer in the program buffer
  block_copy16(stage, g.rows, stage_bytes);
  block_copy16(cols, ULTRA ? g.cols_ultra : g.cols_plain, 4 * E);
  for (int k = threadIdx.x; k <= n; k += blockDim.x) {
    x[k] = HAS_X && k < n ? xin[k] : T(0);
    if (k < n) bs[k] = b[k];
  }
  __syncthreads();
  if (threadIdx.x < WSEG) {
    const int lane = int(threadIdx.x);
    const T m1w = T(1) - omega;
    auto xat = [&](unsigned off) { return *reinterpret_cast<const T*>(base + off); };
    // what round r needs besides x (its right-hand side follows in rhs())
    auto fetch = [&](int r, R& q) {
      q.h = hdr[r];
      q.act = lane < q.h.z;
      const int p = q.h.y + (q.act ? lane : 0);
      q.e = q.h.x + (q.act ? lane : 0) * q.h.w;
      q.i = rows[p];
      q.d = dg[p];
#pragma unroll
      for (int k = 0; k < RS / 4; ++k) q.c[k] = *reinterpret_cast<const typename R::U4*>(cols + q.e + 4 * k);
#pragma unroll
      for (int k = 0; k < RS / R::kPer; ++k)
        q.v[k] = *reinterpret_cast<const typename R::TV*>(vals + q.e + R::kPer * k);
    };
    auto rhs = [&](R& q) { q.b = bs[q.i]; };
    // round r from cur; the next round's data into nxt
    auto round = [&](R& cur, R& nxt, int r) {
      const T xi = x[cur.i];
      T acc = cur.b;
      auto run = [&](auto CTc) {
        constexpr int CT = decltype(CTc)::value;  // slots taken from registers
        T xv[CT];
#pragma unroll
        for (int u = 0; u < CT; ++u) xv[u] = xat(cur.col(u));
        if (r + 1 < g.nrounds) fetch(r + 1, nxt);
#pragma unroll
        for (int u = 0; u < CT; ++u) acc = fma(-cur.val(u), xv[u], acc);
        // entries beyond the register slots, eight at a time from LDS
        for (int k = CT; k < cur.h.w; k += 8) {
          unsigned c[8];
          T v[8], xw[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            c[u] = cols[cur.e + k + u];
            v[u] = vals[cur.e + k + u];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) xw[u] = xat(c[u]);
#pragma unroll
          for (int u = 0; u < 8; ++u) acc = fma(-v[u], xw[u], acc);
        }
      };
      // every chain holds at least the register slots (the host pads rows to
      // them), so one path serves every round
      run(std::integral_constant<int, RS>{});
      T xn;
      if constexpr (ULTRA) {
        const T ad = cur.d < T(0) ? -cur.d : cur.d;
        xn = xi + (omega * acc) / (ad + T(0.5) * T(0));
      } else {
        const T quot = acc / cur.d;
        xn = fma(omega, quot, m1w * xi);
      }
      if (cur.act) x[cur.i] = xn;
      if (r + 1 < g.nrounds) rhs(nxt);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
  R A, B;
    if (g.nrounds > 0) {
      fetch(0, A);
      rhs(A);
    }
    for (int r = 0; r < g.nrounds; r += 2) {
      round(A, B, r);
      if (r + 1 >= g.nrounds) break;
      round(B, A, r + 1);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += blockDim.x) xout[k] = x[k];
}

// the register-slot count the host chose for the level (f64: always 16)
template <typename T, bool ULTRA, bool HAS_X>
__device__ void round_sweep(const RoundSweep& g, int64_t n, const T* __restrict__ xin, T* __restrict__ xout,
                            const T* __restrict__ b, T omega, double* lds) {
  if constexpr (sizeof(T) == 8) {
    round_sweep_slots<T, ULTRA, HAS_X, kRoundSlots64>(g, n, xin, xout, b, omega, lds);
  } else {
    if (g.slots <= 16) round_sweep_slots<T, ULTRA, HAS_X, 16>(g, n, xin, xout, b, omega, lds);
    else if (g.slots <= 24) round_sweep_slots<T, ULTRA, HAS_X, 24>(g, n, xin, xout, b, omega, lds);
    else round_sweep_slots<T, ULTRA, HAS_X, 32>(g, n, xin, xout, b, omega, lds);
  }
}

template <typename T, bool ULTRA, bool HAS_X>
__global__ __launch_bounds__(512) void csr_round_kernel(RoundSweep g, int64_t n, const T* __restrict__ xin,
                                                        T* __restrict__ xout, const T* __restrict__ b,
                                                        T omega) {
  extern __shared__ double round_lds[];
  round_sweep<T, ULTRA, HAS_X>(g, n, xin, xout, b, omega, round_lds);
}

// The cycle's tail in one workgroup (amg_tail): every phase of the small
// levels' V-cycle in turn, with a barrier between phases, so the tail costs
// one launch instead of five per level.  Each phase repeats the standalone
// kernels' arithmetic: the 64-row segment sweeps (wave_sweep_segment, one
// segment per wavefront at a time, each wavefront in its own LDS area), the
// residual (csr_staged_kernel's row order), the restriction sums in list
// order, the dense coarsest solve and the prolongation.
template <typename T, bool FWD, bool ULTRA, bool HAS_X>
__device__ __forceinline__ void tail_sweep(const TailLevel& a, const T* xin, T* xout, const T* b,
                                           double* my_lds, double* block_lds) {
  const RoundSweep& rs = FWD ? a.rounds_fwd : a.rounds_bwd;
  if (rs.nrounds > 0) {
    round_sweep<T, ULTRA, HAS_X>(rs, a.n, xin, xout, b, T(1), block_lds);
    return;
  }
  const int wid = int(threadIdx.x / WSEG), nw = int(blockDim.x / WSEG);
  const int64_t nseg = (a.n + WSEG - 1) / WSEG;
  const auto va = static_cast<const T*>(a.vals);
  for (int64_t sg = wid; sg < nseg; sg += nw) {
    if (a.shortlists)
      wave_sweep_segment<T, T, FWD, ULTRA, HAS_X, true, true>(sg * WSEG, a.n, a.rowptr, a.cols, va, xin,
                                                             xout, b, T(1), a.cap, my_lds);
    else
      wave_sweep_segment<T, T, FWD, ULTRA, HAS_X, true, false>(sg * WSEG, a.n, a.rowptr, a.cols, va, xin,
                                                              xout, b, T(1), a.cap, my_lds);
  }
}

template <typename T, bool ULTRA>
__global__ __launch_bounds__(512) void amg_tail_kernel(const TailLevel* __restrict__ lv, int count,
                                                       int64_t coarsest_n, const double* L,
                                                       const double* U, T inv_relax, int64_t wave_lds) {
  extern __shared__ double tail_lds[];
  double* const my_lds = tail_lds + int64_t(threadIdx.x / WSEG) * (wave_lds / int64_t(sizeof(double)));
  const int64_t tid = threadIdx.x, nt = blockDim.x;
  for (int t = 0; t + 1 < count; ++t) {
    const TailLevel a = lv[t];
    T* const x = static_cast<T*>(a.sol);
    const T* const b = static_cast<const T*>(a.rhs);
    tail_sweep<T, true, ULTRA, false>(a, nullptr, x, b, my_lds, tail_lds);
    __syncthreads();
    // residual (sm_resid: the row's products summed with fma in column
    // order), the entries in batches whose loads are all in flight together
    T* const r = static_cast<T*>(a.tmp);
    const T* const va = static_cast<const T*>(a.vals);
    for (int64_t i = tid; i < a.n; i += nt) {
      const int64_t k0 = a.rowptr[i], k1 = a.rowptr[i + 1];
      T sum = T(0);
      for (int64_t kb = k0; kb < k1; kb += kStageBatch) {
        int32_t j[kStageBatch];
        T v[kStageBatch], xv[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
          const int64_t k = kb + u < k1 ? kb + u : k1 - 1;
          j[u] = a.cols[k];
          v[u] = va[k];
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) xv[u] = x[j[u]];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u)
          if (kb + u < k1) sum = fma(v[u], xv[u], sum);
      }
      r[i] = b[i] - sum;
    }
    __syncthreads();
    // restriction into the next level: each aggregate in list order
    const TailLevel nx = lv[t + 1];
    T* const bc = static_cast<T*>(nx.rhs);
    for (int64_t c = tid; c < nx.n; c += nt) {
      const int64_t k0 = a.c2f_ptr[c], k1 = a.c2f_ptr[c + 1];
      T sum = T(0);
      for (int64_t kb = k0; kb < k1; kb += kStageBatch) {
        T v[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) v[u] = r[a.c2f_idx[kb + u < k1 ? kb + u : k1 - 1]];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u)
          if (kb + u < k1) sum += v[u];
      }
      bc[c] = sum;
    }
    __syncthreads();
  }
  {
    const TailLevel c = lv[count - 1];
    if (U) block_lu_solve(L, U, coarsest_n, static_cast<const T*>(c.rhs), static_cast<T*>(c.sol), tail_lds);
    else block_chol_solve(L, coarsest_n, static_cast<const T*>(c.rhs), static_cast<T*>(c.sol), tail_lds);
    __syncthreads();
  }
  for (int t = count - 2; t >= 0; --t) {
    const TailLevel a = lv[t];
    const T* const xc = static_cast<const T*>(lv[t + 1].sol);
    T* const x = static_cast<T*>(a.sol);
    T* const aux = static_cast<T*>(a.aux);
    for (int64_t i = tid; i < a.n; i += nt) {
      const uint32_t g = a.agg[i];
      const T refined = g < 0xFFFFFFFEu ? xc[g] : T(0);
      aux[i] = x[i] + refined * inv_relax;
    }
    __syncthreads();
    tail_sweep<T, false, ULTRA, true>(a, aux, x, static_cast<const T*>(a.rhs), my_lds, tail_lds);
    __syncthreads();
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

size_t tail_wave_lds_bytes(int64_t cap, bool f32) {
  const size_t b = f32 ? wave_lds_bytes<float, float>(cap) : wave_lds_bytes<double, double>(cap);
  return (b + 15) & ~size_t(15);
}

size_t tail_lds_budget() { return wave_lds_budget(); }

size_t round_lds_bytes(int64_t n, int64_t entries, int nrounds, bool f32) {
  const int64_t es = f32 ? 4 : 8;
  return size_t(round16((n + 1) * es) + round16(n * es) + round_stage_bytes(n, entries, nrounds, es) +
                4 * (entries + kRoundPad));
}

void csr_sor_rounds(const CsrDesc& a, const RoundSweep& rs, const void* x_in, void* x_out,
                    const void* b, double omega, bool ultra, hipStream_t st) {
  LGPU_REQUIRE(rs.nrounds > 0 && a.n <= kRoundMaxRows, "no round program for this level");
  const size_t lds = round_lds_bytes(a.n, rs.entries, rs.nrounds, a.f32 != 0);
  LGPU_REQUIRE(lds <= wave_lds_budget(), "round program exceeds the LDS budget");
  const dim3 grid(1), block(512);
  auto go = [&](auto tag) {
    using T = decltype(tag);
    const auto xi = static_cast<const T*>(x_in);
    const auto xo = static_cast<T*>(x_out);
    const auto bb = static_cast<const T*>(b);
    if (ultra) {
      if (x_in) hipLaunchKernelGGL((csr_round_kernel<T, true, true>), grid, block, lds, st, rs, a.n, xi, xo, bb, T(omega));
      else hipLaunchKernelGGL((csr_round_kernel<T, true, false>), grid, block, lds, st, rs, a.n, xi, xo, bb, T(omega));
    } else {
      if (x_in) hipLaunchKernelGGL((csr_round_kernel<T, false, true>), grid, block, lds, st, rs, a.n, xi, xo, bb, T(omega));
      else hipLaunchKernelGGL((csr_round_kernel<T, false, false>), grid, block, lds, st, rs, a.n, xi, xo, bb, T(omega));
    }
  };
  if (a.f32) go(float{});
  else go(double{});
  LGPU_HIP_CHECK(hipGetLastError());
}

void amg_tail(const TailLevel* levels, int count, int64_t coarsest_n, const double* L,
              const double* U, bool f32, bool ultra, double inv_relax, int waves, size_t wave_lds,
              hipStream_t st) {
  LGPU_REQUIRE(count >= 2 && waves >= 1 && waves <= kTailMaxWaves, "invalid cycle tail");
  const size_t lds = std::max(size_t(waves) * wave_lds, size_t(coarsest_n) * sizeof(double));
  const dim3 grid(1), block(unsigned(waves * WSEG));
  if (f32) {
    if (ultra) hipLaunchKernelGGL((amg_tail_kernel<float, true>), grid, block, lds, st, levels, count, coarsest_n, L, U, float(inv_relax), int64_t(wave_lds));
    else hipLaunchKernelGGL((amg_tail_kernel<float, false>), grid, block, lds, st, levels, count, coarsest_n, L, U, float(inv_relax), int64_t(wave_lds));
  } else {
    if (ultra) hipLaunchKernelGGL((amg_tail_kernel<double, true>), grid, block, lds, st, levels, count, coarsest_n, L, U, inv_relax, int64_t(wave_lds));
    else hipLaunchKernelGGL((amg_tail_kernel<double, false>), grid, block, lds, st, levels, count, coarsest_n, L, U, inv_relax, int64_t(wave_lds));
  }
  LGPU_HIP_CHECK(hipGetLastError());
}

int sched_positions(int64_t n, const int64_t* rowptr, const int32_t* cols, int64_t seg,
                    std::vector<int2>& fwd, std::vector<int2>& bwd) {
  LGPU_REQUIRE(seg > 0, "segment length must be positive");
  LGPU_REQUIRE(n < (int64_t(1) << 31), "too many rows for a sweep schedule");
  std::vector<int32_t> lev(static_cast<size_t>(n), 0);
  int longest = 0;
  for (int64_t i = 0; i < n; ++i) longest = std::max(longest, int(rowptr[i + 1] - rowptr[i]));
  auto order = [&](std::vector<int2>& out) {
    out.resize(size_t(n));
    for (int64_t s0 = 0; s0 < n; s0 += seg) {
      const int64_t s1 = std::min(s0 + seg, n);
      int depth = 0;
      for (int64_t i = s0; i < s1; ++i) depth = std::max(depth, lev[size_t(i)]);
      // counting sort by level, rows ascending within a level
      std::vector<int64_t> start(size_t(depth) + 2, 0);
      for (int64_t i = s0; i < s1; ++i) ++start[size_t(lev[size_t(i)]) + 1];
      for (size_t k = 1; k < start.size(); ++k) start[k] += start[k - 1];
      for (int64_t i = s0; i < s1; ++i) {
        const int l = lev[size_t(i)];
        out[size_t(s0 + start[size_t(l)]++)] = int2{int(i), l};
      }
    }
  };
  // Levels also keep every row that reads the old value of a row of its
  // segment before that row (the sweep order's write-after-read), so a sweep
  // may update one vector in place.  after[i]: the deepest level of a row
  // that reads i's old value.
  std::vector<int32_t> after(static_cast<size_t>(n), 0);
  // forward: new values below the row, old values above it
  for (int64_t i = 0; i < n; ++i) {
    const int64_t s0 = i / seg * seg, s1 = std::min(s0 + seg, n);
    int l = after[size_t(i)] + 1;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      const int64_t j = cols[k];
      if (j >= s0 && j < i) l = std::max(l, lev[size_t(j)] + 1);
    }
    lev[size_t(i)] = l;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      const int64_t j = cols[k];
      if (j > i && j < s1) after[size_t(j)] = std::max(after[size_t(j)], l);
    }
  }
  order(fwd);
  // backward: new values above the row, old values below it
  std::fill(after.begin(), after.end(), 0);
  for (int64_t i = n - 1; i >= 0; --i) {
    const int64_t s0 = i / seg * seg, s1 = std::min(s0 + seg, n);
    int l = after[size_t(i)] + 1;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      const int64_t j = cols[k];
      if (j > i && j < s1) l = std::max(l, lev[size_t(j)] + 1);
    }
    lev[size_t(i)] = l;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      const int64_t j = cols[k];
      if (j >= s0 && j < i) after[size_t(j)] = std::max(after[size_t(j)], l);
    }
  }
  order(bwd);
  return longest;
}

// rows and levels fit 15 bits of the packed positions, entry offsets 24 bits
static bool csr_lds_sweep_fits(int64_t n, int64_t nnz, bool f32) {
  return n <= 0x7FFF && nnz < (int64_t(1) << 24) &&
         lds_sweep_bytes(n, nnz, f32 ? sizeof(float) : sizeof(double)) <= wave_lds_budget();
}

void csr_sor_scheduled(const CsrDesc& a, const void* x_in, void* x_out, const void* b,
                       double omega, bool forward, bool ultra, hipStream_t st) {
  LGPU_REQUIRE(x_in != x_out, "SOR sweep needs distinct input and output vectors");
  LGPU_REQUIRE(a.sched_seg > 0 && a.sched_fwd && a.sched_bwd, "no sweep schedule for this matrix");
  if (a.n == 0) return;
  // one segment that fits one workgroup's LDS: one wavefront over the LDS copy
  if (a.sched_seg >= a.n && a.sched_max_len < 256 && csr_lds_sweep_fits(a.n, a.nnz, a.f32)) {
    if (a.f32) dispatch_lds<float>(a, x_in, x_out, b, omega, forward, ultra, st);
    else dispatch_lds<double>(a, x_in, x_out, b, omega, forward, ultra, st);
    return;
  }
  if (a.f32) dispatch_sched<float, float>(a, x_in, x_out, b, omega, forward, ultra, st);
  else dispatch_sched<double, double>(a, x_in, x_out, b, omega, forward, ultra, st);
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
