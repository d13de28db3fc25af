// Lexicographic SOR on the matrix-free stencil as a tile-column wavefront.
//
// The reference's one-segment SOR (gauss-seidel-variant/sweep.hpp:134-200)
// updates cell i from the new values of its -x, -y, -z neighbours and the old
// values of its +x, +y, +z neighbours.  Here the (x, y) plane is cut into
// columns of 8 x 8 z-lines; one wavefront owns a column, lane (a, b) owns line
// (8I + a, 8J + b) and at step t computes z = t - a - b.  Its -x and -y
// neighbours were computed by lanes (a-1, b) and (a, b-1) one step earlier and
// arrive by cross-lane shuffles; its -z neighbour is its own previous value.
// The face weights are shared the same way: the -x / -y / -z weight of a cell
// is the +x / +y / +z weight its predecessor computed one step earlier (the
// harmonic mean is symmetric bit for bit), so each cell evaluates three faces.
//
// Steps run in chunks of K.  Every lane loads the coefficient,
// right-hand-side and old-value windows of its own line for a chunk's K cells
// (lanes on the column border also load the neighbouring column's lines) as
// 16-byte runs, one chunk ahead, so the K steps run from registers and
// shuffles while the next chunk's windows are in flight.  The lines on a column's downstream borders
// publish every new value into a mailbox (a word initialised to a sentinel NaN
// and written once with a single-copy-atomic store); the border lines of the
// downstream columns (I+1, J) and (I, J+1) poll those words, so no memory
// fence and no separate progress flag is needed.  Columns are claimed through
// an atomic ticket in diagonal (I + J) order, so a column only waits on
// columns already claimed by running wavefronts: every wait ends (waits are
// also bounded and report a timeout).
// The backward sweep is the same schedule in reflected coordinates.  Each cell
// performs exactly the oracle's operations in the oracle's order, so the
// result equals one-segment lexicographic SOR bit for bit.
#include "device.hpp"

namespace lgpu {

namespace {

constexpr int CB = 8;        // lines per column side
// steps per chunk (one window load and one mailbox prefetch each): shorter
// chunks let a column follow its upstream columns more closely, longer ones
// amortise the loads.  With 16-byte window runs, 4 measured best at 256^3 for
// f32 (the forward sweep from zero 7% faster than 8; 16 slower) and 8 for f64
// (4 and 16 slower)
template <typename T>
constexpr int kChunk = 8;
template <>
constexpr int kChunk<float> = 4;
constexpr int SPIN_LIMIT = 1 << 24;

// Mailbox words start as a NaN with a payload that no arithmetic produces
// (arithmetic NaNs are the canonical quiet NaN or carry an input's payload);
// a word is published with one single-copy-atomic store, so a reader that
// sees anything else sees the final value and needs no fence.  The f64
// sentinel is the f32 one twice, so a set cleared in either precision reads
// as cleared in f32.
constexpr uint32_t kSentinel32 = 0xFFFABADBu;
template <typename T>
struct Sentinel;
template <>
struct Sentinel<float> {
  using Bits = uint32_t;
  static constexpr Bits bits = kSentinel32;
};
template <>
struct Sentinel<double> {
  using Bits = uint64_t;
  static constexpr Bits bits = (uint64_t(kSentinel32) << 32) | kSentinel32;
};

template <typename T>
__device__ __forceinline__ bool is_sentinel(T v) {
  return __builtin_bit_cast(typename Sentinel<T>::Bits, v) == Sentinel<T>::bits;
}

// bounded wait for a mailbox word; an expired wait sets the sticky error word
// and every later wait returns at once, so the sweep always ends
template <typename T>
__device__ __forceinline__ T mailbox_poll(const T* p, int* err) {
  T v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int it = 0;
#pragma clang loop unroll(disable)
  while (is_sentinel(v)) {
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
    __builtin_amdgcn_s_sleep(1);
    if (++it > SPIN_LIMIT) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return v;
}

// the whole of one mailbox set (n dwords) and its ticket, when no sweep has
// cleared them for the coming one
__global__ void mailbox_reset(uint32_t* box, int64_t n, int* ticket) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ticket = 0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    box[i] = kSentinel32;
}

template <typename R>
__device__ __forceinline__ R shfl_raw(R v, int src) {
  if constexpr (sizeof(R) == 1) return R(__shfl(int(v), src, 64));
  else return __shfl(v, src, 64);
}

// Buffer resources over the planes a column touches: base in scalar
// registers, 32-bit per-lane byte offsets.  Offsets are clamped into the range
// (values of cells outside a line are never used by a valid step); masked-off
// lanes use kOff, which reads 0 and drops a store, so loads and stores need no
// branches.
using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr uint32_t kOff = 0xFFFFFFF0u;
template <typename E>
__device__ __forceinline__ Rsrc col_rsrc(const E* base, int64_t elem0, int64_t elems) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<E*>(base + elem0), short(0),
                                           int(uint32_t(elems * int64_t(sizeof(E)))), 0x00020000);
}
template <typename E>
__device__ __forceinline__ E rload(Rsrc r, uint32_t off) {
  if constexpr (sizeof(E) == 1) return E(__builtin_amdgcn_raw_buffer_load_b8(r, off, 0u, 0));
  else if constexpr (sizeof(E) == 4) return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0u, 0));
  else return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0u, 0));
}
// POL: cache policy bits (16: sc1, the write-through that relaxed agent-scope
// atomic stores use; a naturally aligned dword or dwordx2 store is
// single-copy atomic)
template <typename T, int POL = 0>
__device__ __forceinline__ void rstore(T v, Rsrc r, uint32_t off) {
  if constexpr (sizeof(T) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, 0u, POL);
  else {
    using V2 = unsigned int __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, off, 0u, POL);
  }
}

// Two mailbox sets (and two tickets) alternate between sweeps: a sweep uses
// one and leaves the other cleared for the next, so no reset kernel runs
// between sweeps.  The wave that draws ticket 0 zeroes the other ticket; every
// column wave, once its steps are done, clears its 16 nz words of the other
// set (any column layout of either sweep kernel lies in [0, 16 nz ncols)
// words).  Nothing reads the other set during this sweep, and the next sweep
// starts after this one ends.
__device__ __forceinline__ int claim_column(int* ticket, int* ticket_next) {
  __shared__ int s_ticket;
  if (threadIdx.x == 0) {
    const int t = atomicAdd(ticket, 1);
    if (t == 0) *ticket_next = 0;
    s_ticket = t;
  }
  __syncthreads();
  return s_ticket;
}
template <typename T>
__device__ __forceinline__ void clear_next_box(T* next, int64_t col, int64_t nz, int lane) {
  using V4 = unsigned int __attribute__((ext_vector_type(4)));
  // 16 nz words of T are nz * sizeof(T) 16-byte units
  const int32_t units = int32_t(nz * int64_t(sizeof(T)));
  const Rsrc r = col_rsrc(next, col * 16 * nz, 16 * nz);
  const V4 v = {kSentinel32, kSentinel32, kSentinel32, kSentinel32};
  for (int32_t i = lane; i < units; i += 64)
    __builtin_amdgcn_raw_buffer_store_b128(v, r, uint32_t(i) * 16u, 0u, 0);
}

// Runs of consecutive elements with 16-byte loads and stores: a lane's K-step
// window of one line is contiguous, so one instruction moves four dwords of
// it.  Element loads touch one cache line per lane and instruction, and these
// sweeps are bound by that line rate of the vector memory pipeline, not by
// bytes.  The lowest element's byte offset must be valid (the run may reach
// Callers keep every run they use inside its resource; kRunOff moves nothing
// (2^31 and the run's length stay beyond any column resource, which holds
// fewer than 2^31 bytes, with or without 32-bit wrap-around).
constexpr uint32_t kRunOff = 0x80000000u;
// POL: cache policy bits of the loads (16: sc1, the coherence that relaxed
// agent-scope atomic loads use)
template <int NDW, int POL = 0>
__device__ __forceinline__ void load_dwords(Rsrc r, uint32_t off, uint32_t (&d)[NDW]) {
  using V4 = unsigned int __attribute__((ext_vector_type(4)));
  using V2 = unsigned int __attribute__((ext_vector_type(2)));
  constexpr int n4 = NDW / 4, rem = NDW % 4;
#pragma unroll
  for (int i = 0; i < n4; ++i) {
    const V4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16u * i, 0u, POL);
    d[4 * i] = v[0];
    d[4 * i + 1] = v[1];
    d[4 * i + 2] = v[2];
    d[4 * i + 3] = v[3];
  }
  if constexpr (rem >= 2) {
    const V2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off + 16u * n4, 0u, POL);
    d[4 * n4] = v[0];
    d[4 * n4 + 1] = v[1];
  }
  if constexpr (rem % 2) d[NDW - 1] = __builtin_amdgcn_raw_buffer_load_b32(r, off + 4u * (NDW - 1), 0u, POL);
}

// out[j] = element lo + j (REV: lo + N - 1 - j)
template <typename E, int N, bool REV, int POL = 0>
__device__ __forceinline__ void load_run(Rsrc r, uint32_t lo_off, E (&out)[N]) {
  static_assert(sizeof(E) == 4 || sizeof(E) == 8, "runs of dword elements");
  constexpr int W = int(sizeof(E) / 4);
  uint32_t d[N * W];
  load_dwords<N * W, POL>(r, lo_off, d);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    E v;
    if constexpr (W == 1) v = __builtin_bit_cast(E, d[j]);
    else v = __builtin_bit_cast(E, uint64_t(d[2 * j]) | (uint64_t(d[2 * j + 1]) << 32));
    out[REV ? N - 1 - j : j] = v;
  }
}

// element lo + j <- v[j] (REV: lo + N - 1 - j <- v[j])
template <typename T, int N, bool REV>
__device__ __forceinline__ void store_run(const T (&v)[N], Rsrc r, uint32_t lo_off) {
  using V4 = unsigned int __attribute__((ext_vector_type(4)));
  constexpr int W = int(sizeof(T) / 4), NDW = N * W;
  static_assert(NDW % 4 == 0, "whole 16-byte stores");
  uint32_t d[NDW];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const T x = v[REV ? N - 1 - j : j];
    if constexpr (W == 1) {
      d[j] = __builtin_bit_cast(uint32_t, x);
    } else {
      const uint64_t u = __builtin_bit_cast(uint64_t, x);
      d[2 * j] = uint32_t(u);
      d[2 * j + 1] = uint32_t(u >> 32);
    }
  }
#pragma unroll
  for (int i = 0; i < NDW / 4; ++i) {
    const V4 q = {d[4 * i], d[4 * i + 1], d[4 * i + 2], d[4 * i + 3]};
    __builtin_amdgcn_raw_buffer_store_b128(q, r, lo_off + 16u * i, 0u, 0);
  }
}

// The per-cell array a valuator reads: coefficients or material ids.
template <typename T, int KIND>
__device__ __forceinline__ const typename CellInfo<T, KIND>::Raw* cell_ptr(const StencilDesc& s) {
  if constexpr (KIND == 0) return coeff_ptr<T>(s);
  else return s.mat;
}

// Connection value of two cells: harmonic mean of per-cell coefficients
// (branch-free with FD, whose range the host checked) or the material table.
template <typename T, int KIND, bool FD>
__device__ __forceinline__ T conn_of(const StencilDesc& s, typename CellInfo<T, KIND>::Raw a,
                                     typename CellInfo<T, KIND>::Raw b, T qd, int d) {
  if constexpr (KIND == 0) return hmean_sel<T, FD>(a, b) * qd;
  else return CellInfo<T, KIND>::conn(s, a, b, qd, d);
}

template <typename T, int KIND, bool FWD, bool ULTRA, bool HAS_X, bool FD>
__global__ __launch_bounds__(64) void sor_wavefront(StencilDesc s, const T* __restrict__ xin,
                                                    T* __restrict__ xout,
                                                    const T* __restrict__ bv, T omega,
                                                    int* ticket, int* ticket_next, int* err,
                                                    const uint32_t* order, int64_t cy, T* mbox,
                                                    T* mbox_next) {
  using CI = CellInfo<T, KIND>;
  using Raw = typename CI::Raw;
  constexpr int K = kChunk<T>;
  constexpr uint32_t SR = sizeof(Raw), ST = sizeof(T);
  const int lane = threadIdx.x;
  const int a = lane / CB, b = lane % CB;
  const int64_t col = order[claim_column(ticket, ticket_next)];
  const int64_t I = col / cy, J = col % cy;
  const int64_t nx = s.nx, ny = s.ny, nz = s.nz, nyz = ny * nz;
  const int64_t lx = I * CB + a, ly = J * CB + b;  // logical line coordinates
  const bool line = lx < nx && ly < ny;
  // physical x of the line and the presence of the logical neighbour lines
  const int64_t x = FWD ? lx : nx - 1 - lx;
  const bool has_sx = lx + 1 < nx, has_px = lx > 0;
  const bool has_sy = ly + 1 < ny, has_py = ly > 0;
  // the physical planes of logical x in [I*CB - 1, I*CB + CB]
  const int64_t l0 = I * CB > 0 ? I * CB - 1 : 0;
  const int64_t l1 = I * CB + CB < nx - 1 ? I * CB + CB : nx - 1;
  const int64_t p0 = FWD ? l0 : nx - 1 - l1, p1 = FWD ? l1 : nx - 1 - l0;
  const int64_t relems = (p1 - p0 + 1) * nyz;
  const Rsrc rc = col_rsrc(cell_ptr<T, KIND>(s), p0 * nyz, relems);
  const Rsrc rxi = col_rsrc(xin, p0 * nyz, HAS_X ? relems : 0);
  const Rsrc rbv = col_rsrc(bv, p0 * nyz, relems);
  const Rsrc rxo = col_rsrc(static_cast<const T*>(xout), p0 * nyz, relems);
  // element offset (in the resources) of logical z = 0 on this line (lines
  // outside the box are clamped into it), and the element steps of logical
  // +z / +x / +y
  const int64_t cxl = lx < nx ? lx : nx - 1, cyl = ly < ny ? ly : ny - 1;
  const int64_t pxl = FWD ? cxl : nx - 1 - cxl, pyl = FWD ? cyl : ny - 1 - cyl;
  const int32_t dz = FWD ? 1 : -1;
  const int32_t zorg = int32_t((pxl - p0) * nyz + pyl * nz + (FWD ? 0 : nz - 1));
  const int32_t hsx = int32_t(FWD ? nyz : -nyz), hsy = int32_t(FWD ? nz : -nz);
  // halo lines: the +x line for a = 7, the -x line for a = 0 (+y / -y for b)
  const bool hx_on = a == CB - 1 || a == 0, hy_on = b == CB - 1 || b == 0;
  const int32_t hxo = a == CB - 1 ? hsx : -hsx, hyo = b == CB - 1 ? hsy : -hsy;
  // the halo lines a lane's steps read
  const bool hx_need = (a == 0 && has_px) || (a == CB - 1 && has_sx);
  const bool hy_need = (b == 0 && has_py) || (b == CB - 1 && has_sy);
  // 16-byte runs for dword elements (the u8 lookup ids keep element loads)
  constexpr bool RUNS = sizeof(Raw) >= 4;
  // byte offset of element e, clamped into the resource: the clamp also keeps
  // the compiler from splitting a constant out of the offset into the
  // instruction's immediate, which the hardware would add without wrapping
  const int32_t emax = int32_t(relems - 1);
  auto boff = [emax](int32_t e, uint32_t sz) {
    return uint32_t(e < 0 ? 0 : (e > emax ? emax : e)) * sz;
  };
  const int64_t nsteps = nz + 2 * (CB - 1);
  const T m1w = T(1) - omega;
  const T qx = T(s.q[0]), qy = T(s.q[1]), qz = T(s.q[2]);
  const T two_q0 = T(2) * qx;
  const T adv = T(s.adv);  // upwind advection: -x coupling and diagonal
  // mailboxes: [X part: column, b, z] the a = 7 lines; [Y part: column, a, z]
  // the b = 7 lines
  const int64_t ncols = ((nx + CB - 1) / CB) * cy;
  T* const ybox = mbox + ncols * CB * nz;
  const int64_t col_x = (I > 0 ? I - 1 : 0) * cy + J, col_y = I * cy + (J > 0 ? J - 1 : 0);
  const T* const in_x = mbox + (col_x * CB + b) * nz;
  const T* const in_y = ybox + (col_y * CB + a) * nz;
  const bool take_x = line && a == 0 && has_px;
  const bool take_y = line && b == 0 && has_py;
  const bool give_x = a == CB - 1, give_y = b == CB - 1;
  // the four mailbox lines a column reads or writes, as resources: its own
  // X and Y parts and those of its upstream columns (the lines of lane b / a
  // start at b * nz / a * nz)
  const Rsrc rox = col_rsrc(static_cast<const T*>(mbox), col * CB * nz, CB * nz);
  const Rsrc roy = col_rsrc(static_cast<const T*>(ybox), col * CB * nz, CB * nz);
  const Rsrc rix = col_rsrc(static_cast<const T*>(mbox), col_x * CB * nz, CB * nz);
  const Rsrc riy = col_rsrc(static_cast<const T*>(ybox), col_y * CB * nz, CB * nz);
  const int32_t lx_off = int32_t(b * nz), ly_off = int32_t(a * nz);
  T prev = T(0), wpx_prev = T(0), wpy_prev = T(0), wpz_prev = T(0);
  // a chunk's windows; the next chunk's are issued while the current one's
  // steps run (after its mailbox prefetch, so waiting for those words does
  // not wait for the windows).  PF = false loads them at the chunk start.
  constexpr bool PF = true;
  // A chunk's windows hold exactly the elements its K steps read (a window
  // element loaded but never read would let the compiler reuse its register
  // while the load is still in flight, and wait for it there):
  //   cw  [K + 1]  coefficients of the own line, logical z0 .. z0 + K
  //   hx/hy [K]    coefficients of the x / y halo line, z0 .. z0 + K - 1
  //   xw  [K + 1]  old values of the own line, z0 .. z0 + K
  //   exx/exy [K]  old values of the +x / +y line beyond the column, z0 ..
  //   bw  [K]      right-hand side, z0 .. z0 + K - 1
  // element loads of a lane's windows (clamped into the resource)
  auto load_elements = [&](Raw(&w_cw)[K + 1], Raw(&w_hx)[K], Raw(&w_hy)[K], T(&w_xw)[K + 1],
                           T(&w_exx)[K], T(&w_exy)[K], T(&w_bw)[K], int32_t z0) {
#pragma unroll
    for (int j = 0; j < K + 1; ++j) {
      const int32_t e = zorg + dz * (z0 + j);
      w_cw[j] = rload<Raw>(rc, boff(e, SR));
      w_xw[j] = HAS_X ? rload<T>(rxi, boff(e, ST)) : T(0);
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int32_t e = zorg + dz * (z0 + j);
      w_hx[j] = rload<Raw>(rc, hx_on ? boff(e + hxo, SR) : kOff);
      w_hy[j] = rload<Raw>(rc, hy_on ? boff(e + hyo, SR) : kOff);
      w_exx[j] = HAS_X ? rload<T>(rxi, give_x ? boff(e + hsx, ST) : kOff) : T(0);
      w_exy[j] = HAS_X ? rload<T>(rxi, give_y ? boff(e + hsy, ST) : kOff) : T(0);
      w_bw[j] = rload<T>(rbv, boff(e, ST));
    }
  };
  // a chunk's windows; the next chunk's are issued while the current one's
  // steps run (after its mailbox prefetch, so waiting for those words does
  // not wait for the windows).  Every lane issues the same 16-byte runs; the
  // few lanes whose windows reach past their column's resource (the first
  // and last chunks of a line) take zeros from those and reload their
  // windows element by element in a masked block.  No path skips the runs,
  // so the waits for data loaded before them can count them.
  auto load_windows = [&](Raw(&w_cw)[K + 1], Raw(&w_hx)[K], Raw(&w_hy)[K], T(&w_xw)[K + 1],
                          T(&w_exx)[K], T(&w_exy)[K], T(&w_bw)[K], int64_t tc) {
    const int32_t z0 = int32_t(tc) - a - b;
    if constexpr (RUNS) {
      // the lowest element of each window (forward: the first step's;
      // backward: the last step's)
      const int32_t z1 = FWD ? zorg + z0 : zorg - z0 - K;        // K + 1 elements
      const int32_t zk = FWD ? zorg + z0 : zorg - z0 - (K - 1);  // K elements
      auto out = [emax](int32_t lo, int n) { return lo < 0 || lo + n - 1 > emax; };
      const bool bad = line && (out(z1, K + 1) || out(zk, K) ||
                                (hx_need && out(zk + hxo, K)) || (hy_need && out(zk + hyo, K)) ||
                                (HAS_X && give_x && has_sx && out(zk + hsx, K)) ||
                                (HAS_X && give_y && has_sy && out(zk + hsy, K)));
      const bool ok = line && !bad;
      load_run<Raw, K + 1, !FWD>(rc, ok ? uint32_t(z1) * SR : kRunOff, w_cw);
      load_run<Raw, K, !FWD>(rc, ok && hx_need ? uint32_t(zk + hxo) * SR : kRunOff, w_hx);
      load_run<Raw, K, !FWD>(rc, ok && hy_need ? uint32_t(zk + hyo) * SR : kRunOff, w_hy);
      load_run<T, K, !FWD>(rbv, ok ? uint32_t(zk) * ST : kRunOff, w_bw);
      if constexpr (HAS_X) {
        load_run<T, K + 1, !FWD>(rxi, ok ? uint32_t(z1) * ST : kRunOff, w_xw);
        load_run<T, K, !FWD>(rxi, ok && give_x && has_sx ? uint32_t(zk + hsx) * ST : kRunOff, w_exx);
        load_run<T, K, !FWD>(rxi, ok && give_y && has_sy ? uint32_t(zk + hsy) * ST : kRunOff, w_exy);
      } else {
#pragma unroll
        for (int j = 0; j < K + 1; ++j) w_xw[j] = T(0);
#pragma unroll
        for (int j = 0; j < K; ++j) w_exx[j] = w_exy[j] = T(0);
      }
      if (bad) load_elements(w_cw, w_hx, w_hy, w_xw, w_exx, w_exy, w_bw, z0);
    } else {
      load_elements(w_cw, w_hx, w_hy, w_xw, w_exx, w_exy, w_bw, z0);
    }
  };
  // the next chunk's windows (PF)
  Raw pcw[K + 1], phx[K], phy[K];
  T pxw[K + 1], pexx[K], pexy[K], pbw[K];
  if constexpr (PF) load_windows(pcw, phx, phy, pxw, pexx, pexy, pbw, 0);
  for (int64_t t0 = 0; t0 < nsteps; t0 += K) {
    Raw cw[K + 1], hx[K], hy[K];
    T xw[K + 1], exx[K], exy[K], bw[K];
    if constexpr (PF) {
#pragma unroll
      for (int j = 0; j < K + 1; ++j) {
        cw[j] = pcw[j];
        xw[j] = pxw[j];
      }
#pragma unroll
      for (int j = 0; j < K; ++j) {
        hx[j] = phx[j];
        hy[j] = phy[j];
        exx[j] = pexx[j];
        exy[j] = pexy[j];
        bw[j] = pbw[j];
      }
    } else {
      load_windows(cw, hx, hy, xw, exx, exy, bw, t0);
    }
    const int32_t z0 = int32_t(t0) - a - b;
    const int32_t zb = FWD ? zorg + z0 : zorg - z0 - (K - 1);  // lowest result element
    // the chunk's results leave as one run when every line's K steps are valid
    const bool run_store = RUNS && __builtin_amdgcn_ballot_w64(line && (z0 < 0 || z0 + K > nz)) == 0;
    T xo[K];
    // ---- new border values of the upstream columns: prefetched from their
    // mailboxes now, polled again at the step that needs them if still
    // unset ----
    // Words inside the taking lines' ranges come as 16-byte sc1 runs (each
    // aligned 4- or 8-byte word is read whole; a word still unset reads as
    // the sentinel and is polled at its step).  The lanes of a chunk that
    // reaches past the line's ends read zeros from the runs and their words
    // one by one in a masked block, so no path skips the runs.
    T nwx[K], nwy[K];
    const bool m_in = z0 >= 0 && z0 + K <= nz;
    load_run<T, K, false, 16>(rix, take_x && m_in ? uint32_t(lx_off + z0) * ST : kRunOff, nwx);
    load_run<T, K, false, 16>(riy, take_y && m_in ? uint32_t(ly_off + z0) * ST : kRunOff, nwy);
    if ((take_x || take_y) && !m_in) {
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int32_t lz = z0 + j;
        const bool ok = lz >= 0 && lz < nz;
        const int32_t cz = lz < 0 ? 0 : (lz >= nz ? int32_t(nz - 1) : lz);
        const T vx = __hip_atomic_load(in_x + cz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const T vy = __hip_atomic_load(in_y + cz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        nwx[j] = take_x && ok ? vx : T(0);
        nwy[j] = take_y && ok ? vy : T(0);
      }
    }
    // issued on every chunk, the last one included (its windows lie beyond
    // the line and load clamped elements that no step uses): every chunk then
    // issues the same sequence of vector memory instructions, so the waits
    // for the mailbox words below count the window loads behind them and do
    // not wait for those
    if constexpr (PF) load_windows(pcw, phx, phy, pxw, pexx, pexy, pbw, t0 + K);
    // ---- K steps from registers ----
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int32_t lz = z0 + k;
      const bool valid = line && lz >= 0 && lz < nz && t0 + k < nsteps;
      if (valid && ((take_x && is_sentinel(nwx[k])) || (take_y && is_sentinel(nwy[k])))) {
        // one round trip refreshes this step's word and every later word of
        // the chunk still unset, so the following steps find theirs present
#pragma unroll
        for (int j = k + 1; j < K; ++j) {
          const int32_t lzj = z0 + j;
          const bool okj = lzj >= 0 && lzj < nz;
          const int32_t czj = lzj < 0 ? 0 : (lzj >= nz ? int32_t(nz - 1) : lzj);
          const T vx = __hip_atomic_load(in_x + czj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const T vy = __hip_atomic_load(in_y + czj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (take_x && okj && is_sentinel(nwx[j])) nwx[j] = vx;
          if (take_y && okj && is_sentinel(nwy[j])) nwy[j] = vy;
        }
        if (take_x) nwx[k] = mailbox_poll(in_x + lz, err);
        if (take_y) nwy[k] = mailbox_poll(in_y + lz, err);
        // leave with no load of this block in flight: otherwise the registers
        // it loaded would count as pending on the path that skips it too, and
        // their later reuse would wait for the next chunk's windows there
        __builtin_amdgcn_s_waitcnt(0);
      }
      // successor data of the neighbouring lanes (same logical z)
      // (lane +-1 by DPP whole-wave shifts, lane +-8 by permutes)
      Raw csx = shfl_raw(cw[k + 1], lane + CB < 64 ? lane + CB : lane);
      Raw csy = shift<kWaveShl1>(cw[k + 1], cw[k + 1]);
      T xsx = __shfl(xw[k + 1], lane + CB < 64 ? lane + CB : lane, 64);
      T xsy = shift<kWaveShl1>(xw[k + 1], xw[k + 1]);
      // predecessor results of the neighbouring lanes (one step ago)
      const T npx = __shfl(prev, lane >= CB ? lane - CB : lane, 64);
      const T npy = shift<kWaveShr1>(prev, prev);
      const T wpx_in = __shfl(wpx_prev, lane >= CB ? lane - CB : lane, 64);
      const T wpy_in = shift<kWaveShr1>(wpy_prev, wpy_prev);
      if (give_x) {
        csx = hx[k];
        xsx = exx[k];
      }
      if (give_y) {
        csy = hy[k];
        xsy = exy[k];
      }
      const Raw cc = cw[k];
      // face weights: successor faces computed here, predecessor faces shared
      // (computed unconditionally, absent faces selected to 0)
      const bool has_sz = lz + 1 < nz, has_pz = lz > 0;
      const T c_sx = conn_of<T, KIND, FD>(s, cc, csx, qx, 0);
      const T c_sy = conn_of<T, KIND, FD>(s, cc, csy, qy, 1);
      const T c_sz = conn_of<T, KIND, FD>(s, cc, cw[k + 1], qz, 2);
      const T c_px = conn_of<T, KIND, FD>(s, cc, hx[k], qx, 0);
      const T c_py = conn_of<T, KIND, FD>(s, cc, hy[k], qy, 1);
      const T w_sx = has_sx ? c_sx : T(0);
      const T w_sy = has_sy ? c_sy : T(0);
      const T w_sz = has_sz ? c_sz : T(0);
      const T w_px = a == 0 ? (has_px ? c_px : T(0)) : wpx_in;
      const T w_py = b == 0 ? (has_py ? c_py : T(0)) : wpy_in;
      const T w_pz = has_pz ? wpz_prev : T(0);
      // predecessor (new) and successor (old) values
      const T n_px = a == 0 ? nwx[k] : npx;
      const T n_py = b == 0 ? nwy[k] : npy;
      const T n_pz = has_pz ? prev : T(0);
      const T o_sx = has_sx ? xsx : T(0);
      const T o_sy = has_sy ? xsy : T(0);
      const T o_sz = has_sz ? xw[k + 1] : T(0);
      // physical orientation: forward, predecessors are -x/-y/-z
      const T wmx = FWD ? w_px : w_sx, wpx = FWD ? w_sx : w_px;
      const T wmy = FWD ? w_py : w_sy, wpy = FWD ? w_sy : w_py;
      const T wmz = FWD ? w_pz : w_sz, wpz = FWD ? w_sz : w_pz;
      const T vmx = FWD ? n_px : o_sx, vpx = FWD ? o_sx : n_px;
      const T vmy = FWD ? n_py : o_sy, vpy = FWD ? o_sy : n_py;
      const T vmz = FWD ? n_pz : o_sz, vpz = FWD ? o_sz : n_pz;
      // diagonal in the reference's visiting order (algorithm.hpp:106-226)
      const T ci = CI::value(s, cc);
      T d = T(0);
      if (x == 0) d += two_q0 * ci;
      if (x + 1 == nx) d += two_q0 * ci;
      d += wmx;
      d += wmy;
      d += wmz;
      d += wpz;
      d += wpy;
      d += wpx;
      d += adv;
      if (s.zero_to_one && d == T(0)) d = T(1);
      const T xi = xw[k];
      T acc = bw[k];
      acc = fma(x > 0 ? wmx + adv : T(0), vmx, acc);
      acc = fma(wmy, vmy, acc);
      acc = fma(wmz, vmz, acc);
      acc = fma(wpz, vpz, acc);
      acc = fma(wpy, vpy, acc);
      acc = fma(wpx, vpx, acc);
      T xn;
      if (ULTRA) {
        acc = fma(-d, xi, acc);
        const T ad = d < T(0) ? -d : d;
        xn = xi + (omega * acc) / (ad + T(0.5) * T(0));
      } else {
        const T quot = acc / d;
        xn = fma(omega, quot, m1w * xi);
      }
      xo[k] = xn;
      // stores without branches (lanes that store nothing use kOff), so that
      // every step issues the same vector memory instructions; the border
      // lines publish their values to the downstream columns with sc1 stores
      rstore<T>(xn, rxo, valid && !run_store ? boff(zorg + dz * lz, ST) : kOff);
      rstore<T, 16>(xn, rox, valid && give_x ? uint32_t(lx_off + lz) * ST : kOff);
      rstore<T, 16>(xn, roy, valid && give_y ? uint32_t(ly_off + lz) * ST : kOff);
      // a lane's values of a step outside its line are read only by steps
      // outside the box
      prev = xn;
      wpx_prev = w_sx;
      wpy_prev = w_sy;
      wpz_prev = w_sz;
    }
    if constexpr (RUNS) store_run<T, K, !FWD>(xo, rxo, line && run_store ? uint32_t(zb) * ST : kRunOff);
  }
  clear_next_box(mbox_next, col, nz, lane);
}

// ---------------------------------------------------------------------------
// The same sweep with the operator's coefficients cached in the sweep's own
// order.  Every cell the forward sweep visits gets a record {w+x, w+y, w+z,
// d}: the weights of the faces to its physical +x / +y / +z neighbours (0
// where there is none) and its diagonal, computed by sor_records_kernel with
// the operations the on-the-fly sweep performs (face weights are symmetric
// bit for bit; the diagonal sums in the same order), so both sweeps give the
// same bits.  A cell's -x / -y / -z faces are the +x / +y / +z faces of its
// neighbours: inside a column they come from the neighbouring lane (its
// record one step earlier forward, one step later backward) or from the own
// record of the previous / next step; the lanes on a column's -x and -y
// borders (physical) read them from a halo array.
//
// Records are stored per column and per chunk of K = 4 steps, in 16-byte
// slices interleaved over the 64 lanes:
//     [column][chunk][slice q][lane][16 bytes]
// so that one 16-byte load of a wave reads 1 KiB of consecutive memory (a
// lane's K records are its NQ slices).  Halo values are stored per column and
// chunk as 16 entries of K values: entry b is the -x face of lane (0, b),
// entry 8 + a the -y face of lane (a, 0), so eight lanes read 128 consecutive
// bytes.  The backward sweep reads the same arrays: with its column grid
// anchored at the box's far corner (its first logical column starts
// rx = (8 - nx % 8) % 8 lines before the box), backward column C, lane L and
// step t' visit the cell of forward column ncols - 1 - C, lane 63 - L, step
// nsteps - 1 - t'.  Forward chunks start at step -s with
// s = (4 - nsteps % 4) % 4, so that the chunks of both sweeps are the same
// sets of four steps.

constexpr int kRecK = 4;     // steps per chunk of the record sweep
constexpr int kRecVals = 4;  // values per record
constexpr int kHaloEntries = 2 * CB;

struct RecGeom {
  int64_t cx, cy, ncols, nsteps, s, nc, rx, ry;
};
__host__ __device__ inline RecGeom rec_geom(const StencilDesc& d) {
  RecGeom g;
  g.cx = (d.nx + CB - 1) / CB;
  g.cy = (d.ny + CB - 1) / CB;
  g.ncols = g.cx * g.cy;
  g.nsteps = d.nz + 2 * (CB - 1);
  g.s = (kRecK - g.nsteps % kRecK) % kRecK;
  g.nc = (g.nsteps + g.s) / kRecK;
  g.rx = (CB - d.nx % CB) % CB;
  g.ry = (CB - d.ny % CB) % CB;
  return g;
}
// 16-byte slices per lane and chunk
template <typename T>
constexpr int kRecSlices = kRecK * kRecVals * int(sizeof(T)) / 16;
// elements of the record array; the halo array follows it
__host__ __device__ inline int64_t rec_elems(const RecGeom& g) {
  return g.ncols * g.nc * kRecK * 64 * kRecVals;
}
__host__ __device__ inline int64_t halo_elems(const RecGeom& g) {
  return g.ncols * g.nc * kHaloEntries * kRecK;
}

template <typename T, int KIND, bool FD>
__global__ __launch_bounds__(256) void sor_records_kernel(StencilDesc s, RecGeom g, T* __restrict__ w) {
  using CI = CellInfo<T, KIND>;
  constexpr int NQ = kRecSlices<T>;
  constexpr int per16 = 16 / int(sizeof(T));
  const int64_t nx = s.nx, ny = s.ny, nz = s.nz, nyz = ny * nz;
  const T qx = T(s.q[0]), qy = T(s.q[1]), qz = T(s.q[2]);
  const T two_q0 = T(2) * qx;
  const T adv = T(s.adv);
  T* const halo = w + rec_elems(g);
  const int64_t total = g.ncols * g.nc * kRecK * 64;
  for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int L = int(idx % 64);
    int64_t r = idx / 64;
    const int k = int(r % kRecK);
    r /= kRecK;
    const int64_t c = r % g.nc, col = r / g.nc;
    const int64_t I = col / g.cy, J = col % g.cy;
    const int a = L / CB, b = L % CB;
    const int64_t x = I * CB + a, y = J * CB + b, z = kRecK * c + k - g.s - a - b;
    T v[kRecVals] = {};
    T hmx = T(0), hmy = T(0);
    if (x < nx && y < ny && z >= 0 && z < nz) {
      const int64_t p = x * nyz + y * nz + z;
      const typename CI::Raw cc = CI::load(s, p);
      auto face = [&](bool has, int64_t q, T qd, int d) {
        return has ? conn_of<T, KIND, FD>(s, cc, CI::load(s, q), qd, d) : T(0);
      };
      const T wmx = face(x > 0, p - nyz, qx, 0), wpx = face(x + 1 < nx, p + nyz, qx, 0);
      const T wmy = face(y > 0, p - nz, qy, 1), wpy = face(y + 1 < ny, p + nz, qy, 1);
      const T wmz = face(z > 0, p - 1, qz, 2), wpz = face(z + 1 < nz, p + 1, qz, 2);
      // the diagonal in the reference's visiting order (algorithm.hpp:106-226)
      const T ci = CI::value(s, cc);
      T d = T(0);
      if (x == 0) d += two_q0 * ci;
      if (x + 1 == nx) d += two_q0 * ci;
      d += wmx;
      d += wmy;
      d += wmz;
      d += wpz;
      d += wpy;
      d += wpx;
      d += adv;
      if (s.zero_to_one && d == T(0)) d = T(1);
      v[0] = wpx;
      v[1] = wpy;
      v[2] = wpz;
      v[3] = d;
      hmx = wmx;
      hmy = wmy;
    }
#pragma unroll
    for (int j = 0; j < kRecVals; ++j) {
      const int i = k * kRecVals + j;  // value index within the lane's chunk
      const int q = i / per16, within = i % per16;
      w[(((col * g.nc + c) * NQ + q) * 64 + L) * per16 + within] = v[j];
    }
    const int64_t hbase = (col * g.nc + c) * kHaloEntries;
    if (a == 0) halo[(hbase + b) * kRecK + k] = hmx;
    if (b == 0) halo[(hbase + CB + a) * kRecK + k] = hmy;
  }
}

template <typename T, bool FWD, bool ULTRA, bool HAS_X>
__global__ __launch_bounds__(64) void sor_wavefront_r(StencilDesc s, RecGeom g,
                                                      const T* __restrict__ w,
                                                      const T* __restrict__ xin,
                                                      T* __restrict__ xout,
                                                      const T* __restrict__ bv, T omega,
                                                      int* ticket, int* ticket_next, int* err,
                                                      const uint32_t* order, T* mbox, T* mbox_next) {
  constexpr int K = kRecK;
  constexpr int NQ = kRecSlices<T>;
  constexpr int W = int(sizeof(T) / 4);  // dwords per value
  constexpr uint32_t ST = sizeof(T);
  const int lane = threadIdx.x;
  const int a = lane / CB, b = lane % CB;
  const int64_t cy = g.cy;
  const int64_t col = order[claim_column(ticket, ticket_next)];
  const int64_t I = col / cy, J = col % cy;
  const int64_t nx = s.nx, ny = s.ny, nz = s.nz, nyz = ny * nz;
  // logical line coordinates (the backward grid starts rx / ry lines before
  // the box)
  const int64_t lx = I * CB + a - (FWD ? 0 : g.rx), ly = J * CB + b - (FWD ? 0 : g.ry);
  const bool line = lx >= 0 && lx < nx && ly >= 0 && ly < ny;
  const int64_t x = FWD ? lx : nx - 1 - lx;  // physical x
  const bool has_sx = lx + 1 < nx, has_px = lx > 0;
  const bool has_sy = ly + 1 < ny, has_py = ly > 0;
  // the physical planes of logical x in [lx(a = 0) - 1, lx(a = 7) + 1]
  const int64_t lxa = I * CB - (FWD ? 0 : g.rx);
  const int64_t l0 = lxa > 0 ? lxa - 1 : 0;
  const int64_t l1 = lxa + CB < nx - 1 ? lxa + CB : nx - 1;
  const int64_t p0 = FWD ? l0 : nx - 1 - l1, p1 = FWD ? l1 : nx - 1 - l0;
  const int64_t relems = (p1 - p0 + 1) * nyz;
  const Rsrc rxi = col_rsrc(xin, p0 * nyz, HAS_X ? relems : 0);
  const Rsrc rbv = col_rsrc(bv, p0 * nyz, relems);
  const Rsrc rxo = col_rsrc(static_cast<const T*>(xout), p0 * nyz, relems);
  // the column's records and halo values (forward column ncols - 1 - col
  // when backward) and this lane's slots in them
  const int64_t rcol = FWD ? col : g.ncols - 1 - col;
  const int64_t rbytes = g.nc * NQ * 64 * 16;
  const Rsrc rrec = col_rsrc(w, rcol * (rbytes / ST), rbytes / ST);
  const uint32_t rlane = uint32_t(FWD ? lane : 63 - lane) * 16u;
  const int64_t hcol = g.nc * kHaloEntries * kRecK;  // halo elements per column
  const Rsrc rhalo = col_rsrc(w + rec_elems(g), rcol * hcol, hcol);
  // the lanes on the physical -x / -y border and their halo entries
  const bool hx_lane = FWD ? a == 0 : a == CB - 1;
  const bool hy_lane = FWD ? b == 0 : b == CB - 1;
  const int hx_entry = FWD ? b : CB - 1 - b, hy_entry = CB + (FWD ? a : CB - 1 - a);
  // element offset (in the resources) of logical z = 0 on this line (lines
  // outside the box are clamped into it), and the element steps of logical
  // +z / +x / +y
  const int64_t cxl = lx < 0 ? 0 : (lx < nx ? lx : nx - 1);
  const int64_t cyl = ly < 0 ? 0 : (ly < ny ? ly : ny - 1);
  const int64_t pxl = FWD ? cxl : nx - 1 - cxl, pyl = FWD ? cyl : ny - 1 - cyl;
  const int32_t dz = FWD ? 1 : -1;
  const int32_t zorg = int32_t((pxl - p0) * nyz + pyl * nz + (FWD ? 0 : nz - 1));
  const int32_t hsx = int32_t(FWD ? nyz : -nyz), hsy = int32_t(FWD ? nz : -nz);
  const int32_t emax = int32_t(relems - 1);
  auto boff = [emax](int32_t e, uint32_t sz) {
    return uint32_t(e < 0 ? 0 : (e > emax ? emax : e)) * sz;
  };
  const T m1w = T(1) - omega;
  const T adv = T(s.adv);  // upwind advection: the -x coupling
  // Mailboxes: 15 lines of nz words per column, the new values of its
  // downstream border lines: line b for lane (7, b), line 8 + a for lane
  // (a, 7) with a < 7 (lane (7, 7)'s values, line 7, serve both
  // directions).  One store per step publishes both borders.
  const int32_t nzi = int32_t(nz);
  const int64_t col_x = (I > 0 ? I - 1 : 0) * cy + J, col_y = I * cy + (J > 0 ? J - 1 : 0);
  const bool take_x = line && a == 0 && has_px;
  const bool take_y = line && b == 0 && has_py;
  const bool give_x = a == CB - 1, give_y = b == CB - 1;
  const int out_line = give_x ? b : (give_y ? CB + a : -1);
  const int in_y_line = a < CB - 1 ? CB + a : CB - 1;
  const Rsrc rmo = col_rsrc(static_cast<const T*>(mbox), col * 15 * nz, 15 * nz);
  const Rsrc rmx = col_rsrc(static_cast<const T*>(mbox), col_x * 15 * nz, 15 * nz);
  const Rsrc rmy = col_rsrc(static_cast<const T*>(mbox), col_y * 15 * nz, 15 * nz);
  const T* const in_x = mbox + col_x * 15 * nz + int64_t(b) * nz;
  const T* const in_y = mbox + col_y * 15 * nz + int64_t(in_y_line) * nz;
  const int32_t ix_off = b * nzi, iy_off = in_y_line * nzi, o_off = out_line * nzi;
  T prev = T(0);
  // carried from the previous step (forward): the own +x / +y / +z face
  // weights, which the +x / +y lanes and the next step read as their -x /
  // -y / -z faces
  T wpx_prev = T(0), wpy_prev = T(0), wpz_prev = T(0);
  // A chunk's windows (steps t0 .. t0 + K - 1, logical z0 = t0 - a - b):
  //   rec [K][4]   the cells' records, in step order
  //   hx, hy [K]   -x / -y faces of the border lanes (halo)
  //   xw [K]       old values of the own line, z0 .. z0 + K - 1 (the value
  //                at z0 + K is the next chunk's xw[0])
  //   exx, exy [K] old values of the +x / +y line beyond the column
  //   bw [K]       right-hand side
  struct Win {
    T rec[K][kRecVals];
    T hx[K], hy[K];
    T xw[K], exx[K], exy[K], bw[K];
  };
  auto load_elements = [&](Win& v, int32_t z0) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int32_t e = zorg + dz * (z0 + j);
      v.bw[j] = rload<T>(rbv, boff(e, ST));
      v.exx[j] = HAS_X ? rload<T>(rxi, give_x && has_sx ? boff(e + hsx, ST) : kOff) : T(0);
      v.exy[j] = HAS_X ? rload<T>(rxi, give_y && has_sy ? boff(e + hsy, ST) : kOff) : T(0);
    }
#pragma unroll
    for (int j = 0; j < K; ++j)
      v.xw[j] = HAS_X ? rload<T>(rxi, boff(zorg + dz * (z0 + j), ST)) : T(0);
  };
  // chunk index c (0 .. nc - 1) -> its windows; the natural-layout windows
  // come as 16-byte runs, patched element by element in a masked block for
  // the lanes whose runs reach past the column's resource (no path skips
  // the runs)
  auto load_windows = [&](Win& v, int64_t c) {
    const int64_t t0 = FWD ? kRecK * c - g.s : kRecK * c;
    const int64_t fc = FWD ? c : g.nc - 1 - c;
    {
      uint32_t d[NQ * 4];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        using V4 = unsigned int __attribute__((ext_vector_type(4)));
        const V4 u = __builtin_amdgcn_raw_buffer_load_b128(
            rrec, uint32_t(((fc * NQ + q) * 64) * 16) + rlane, 0u, 0);
        d[4 * q] = u[0];
        d[4 * q + 1] = u[1];
        d[4 * q + 2] = u[2];
        d[4 * q + 3] = u[3];
      }
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int j = 0; j < kRecVals; ++j) {
          const int m = ((FWD ? k : K - 1 - k) * kRecVals + j) * W;
          if constexpr (W == 1) v.rec[k][j] = __builtin_bit_cast(T, d[m]);
          else v.rec[k][j] = __builtin_bit_cast(T, uint64_t(d[m]) | (uint64_t(d[m + 1]) << 32));
        }
    }
    {
      const int64_t hb = fc * kHaloEntries;
      load_run<T, K, !FWD>(rhalo, hx_lane ? uint32_t((hb + hx_entry) * kRecK) * ST : kRunOff, v.hx);
      load_run<T, K, !FWD>(rhalo, hy_lane ? uint32_t((hb + hy_entry) * kRecK) * ST : kRunOff, v.hy);
    }
    const int32_t z0 = int32_t(t0) - a - b;
    const int32_t zk = FWD ? zorg + z0 : zorg - z0 - (K - 1);  // K elements
    auto out = [emax](int32_t lo, int n) { return lo < 0 || lo + n - 1 > emax; };
    const bool bad = line && (out(zk, K) ||
                              (HAS_X && give_x && has_sx && out(zk + hsx, K)) ||
                              (HAS_X && give_y && has_sy && out(zk + hsy, K)));
    const bool ok = line && !bad;
    load_run<T, K, !FWD>(rbv, ok ? uint32_t(zk) * ST : kRunOff, v.bw);
    if constexpr (HAS_X) {
      load_run<T, K, !FWD>(rxi, ok ? uint32_t(zk) * ST : kRunOff, v.xw);
      load_run<T, K, !FWD>(rxi, ok && give_x && has_sx ? uint32_t(zk + hsx) * ST : kRunOff, v.exx);
      load_run<T, K, !FWD>(rxi, ok && give_y && has_sy ? uint32_t(zk + hsy) * ST : kRunOff, v.exy);
    } else {
#pragma unroll
      for (int j = 0; j < K; ++j) v.xw[j] = T(0);
#pragma unroll
      for (int j = 0; j < K; ++j) v.exx[j] = v.exy[j] = T(0);
    }
    if (bad) load_elements(v, z0);
  };
  // One chunk: its steps from `cur` (the backward sweep also reads the first
  // record of `nxt`) while the windows two chunks ahead load into `fill`.
  // Three buffers rotate by name in the loop below, so no window is copied
  // between registers.
  auto chunk = [&](const Win& cur, const Win& nxt, Win& fill, int64_t c) {
    const int64_t t0 = FWD ? kRecK * c - g.s : kRecK * c;
    const int32_t z0 = int32_t(t0) - a - b;
    const int32_t zb = FWD ? zorg + z0 : zorg - z0 - (K - 1);  // lowest result element
    // lanes whose K steps are all in their line store the chunk as one run;
    // the others store element by element in a masked block
    const bool full = line && z0 >= 0 && z0 + K <= nz;
    T xo[K];
    // new border values of the upstream columns, prefetched (as in
    // sor_wavefront)
    T nwx[K], nwy[K];
    const bool m_in = z0 >= 0 && z0 + K <= nz;
    load_run<T, K, false, 16>(rmx, take_x && m_in ? uint32_t(ix_off + z0) * ST : kRunOff, nwx);
    load_run<T, K, false, 16>(rmy, take_y && m_in ? uint32_t(iy_off + z0) * ST : kRunOff, nwy);
    if ((take_x || take_y) && !m_in) {
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int32_t lz = z0 + j;
        const bool okz = lz >= 0 && lz < nz;
        const int32_t cz = lz < 0 ? 0 : (lz >= nz ? nzi - 1 : lz);
        const T vx = __hip_atomic_load(in_x + cz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const T vy = __hip_atomic_load(in_y + cz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        nwx[j] = take_x && okz ? vx : T(0);
        nwy[j] = take_y && okz ? vy : T(0);
      }
    }
    load_windows(fill, c + 2);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int32_t lz = z0 + k;
      const bool valid = line && lz >= 0 && lz < nz;
      if (valid && ((take_x && is_sentinel(nwx[k])) || (take_y && is_sentinel(nwy[k])))) {
#pragma unroll
        for (int j = k + 1; j < K; ++j) {
          const int32_t lzj = z0 + j;
          const bool okj = lzj >= 0 && lzj < nz;
          const int32_t czj = lzj < 0 ? 0 : (lzj >= nz ? nzi - 1 : lzj);
          const T vx = __hip_atomic_load(in_x + czj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const T vy = __hip_atomic_load(in_y + czj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (take_x && okj && is_sentinel(nwx[j])) nwx[j] = vx;
          if (take_y && okj && is_sentinel(nwy[j])) nwy[j] = vy;
        }
        if (take_x) nwx[k] = mailbox_poll(in_x + lz, err);
        if (take_y) nwy[k] = mailbox_poll(in_y + lz, err);
        // leave with no load of this block in flight (see sor_wavefront)
        __builtin_amdgcn_s_waitcnt(0);
      }
      // old values of the logical +x / +y neighbours (same logical z) and the
      // new values of the logical -x / -y neighbours (one step ago)
      // the old value one step ahead in z: the next chunk's first at the
      // chunk's last step
      const T xz1 = k + 1 < K ? cur.xw[k + 1 < K ? k + 1 : 0] : nxt.xw[0];
      T xsx = __shfl(xz1, lane + CB < 64 ? lane + CB : lane, 64);
      T xsy = shift<kWaveShl1>(xz1, xz1);
      const T npx = __shfl(prev, lane >= CB ? lane - CB : lane, 64);
      const T npy = shift<kWaveShr1>(prev, prev);
      if (give_x) xsx = cur.exx[k];
      if (give_y) xsy = cur.exy[k];
      const bool has_sz = lz + 1 < nz, has_pz = lz > 0;
      // (backward, the grid's first column may hold lanes left of the box:
      // a lane at logical x or y = 0 takes no value from its neighbour)
      T n_px = a == 0 ? nwx[k] : npx;
      T n_py = b == 0 ? nwy[k] : npy;
      if constexpr (!FWD) {
        n_px = has_px ? n_px : T(0);
        n_py = has_py ? n_py : T(0);
      }
      const T n_pz = has_pz ? prev : T(0);
      const T o_sx = has_sx ? xsx : T(0);
      const T o_sy = has_sy ? xsy : T(0);
      const T o_sz = has_sz ? xz1 : T(0);
      const T vmx = FWD ? n_px : o_sx, vpx = FWD ? o_sx : n_px;
      const T vmy = FWD ? n_py : o_sy, vpy = FWD ? o_sy : n_py;
      const T vmz = FWD ? n_pz : o_sz, vpz = FWD ? o_sz : n_pz;
      // physical faces: +x / +y / +z and the diagonal from the own record;
      // -x / -y from the lane holding that neighbour's record (forward: lane
      // -8 / -1, one step ago; backward: lane +8 / +1, one step later) or the
      // halo; -z from the own record one step ago (forward) / later
      // (backward).  Records of cells outside the box are 0.
      const T* r = cur.rec[k];
      const T wpx = r[0], wpy = r[1], wpz = r[2], d = r[3];
      T wmx, wmy, wmz;
      if constexpr (FWD) {
        wmx = __shfl(wpx_prev, lane >= CB ? lane - CB : lane, 64);
        wmy = shift<kWaveShr1>(wpy_prev, wpy_prev);
        wmz = wpz_prev;
      } else {
        const T* rn = k + 1 < K ? cur.rec[k + 1 < K ? k + 1 : 0] : nxt.rec[0];
        wmx = __shfl(rn[0], lane + CB < 64 ? lane + CB : lane, 64);
        wmy = shift<kWaveShl1>(rn[1], rn[1]);
        wmz = rn[2];
      }
      if (hx_lane) wmx = cur.hx[k];
      if (hy_lane) wmy = cur.hy[k];
      const T xi = cur.xw[k];
      T acc = cur.bw[k];
      acc = fma(x > 0 ? wmx + adv : T(0), vmx, acc);
      acc = fma(wmy, vmy, acc);
      acc = fma(wmz, vmz, acc);
      acc = fma(wpz, vpz, acc);
      acc = fma(wpy, vpy, acc);
      acc = fma(wpx, vpx, acc);
      T xn;
      if (ULTRA) {
        acc = fma(-d, xi, acc);
        const T ad = d < T(0) ? -d : d;
        xn = xi + (omega * acc) / (ad + T(0.5) * T(0));
      } else {
        const T quot = acc / d;
        xn = fma(omega, quot, m1w * xi);
      }
      xo[k] = xn;
      rstore<T, 16>(xn, rmo, valid && out_line >= 0 ? uint32_t(o_off + lz) * ST : kOff);
      prev = xn;
      if constexpr (FWD) {
        wpx_prev = wpx;
        wpy_prev = wpy;
        wpz_prev = wpz;
      }
    }
    store_run<T, K, !FWD>(xo, rxo, full ? uint32_t(zb) * ST : kRunOff);
    if (line && !full) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int32_t lz = z0 + k;
        rstore<T>(xo[k], rxo, lz >= 0 && lz < nz ? boff(zorg + dz * lz, ST) : kOff);
      }
    }
  };
  Win w0, w1, w2;
  load_windows(w0, 0);
  load_windows(w1, 1);
  const int64_t nc = g.nc;
  for (int64_t c = 0;; c += 3) {
    chunk(w0, w1, w2, c);
    if (c + 1 >= nc) break;
    chunk(w1, w2, w0, c + 1);
    if (c + 2 >= nc) break;
    chunk(w2, w0, w1, c + 2);
    if (c + 3 >= nc) break;
  }
  clear_next_box(mbox_next, col, nz, lane);
}

// one sweep's tickets and mailbox sets: its own and the other (cleared by it)
struct SweepBoxes {
  int* ticket;
  int* ticket_next;
  int* err;
  const uint32_t* order;
  int64_t ncols;
  void* mailbox;
  void* mailbox_next;
};

template <typename T, int KIND, bool FWD, bool ULTRA, bool FD>
void launch5(const StencilDesc& s, const void* x_in, void* x_out, const void* b, double omega,
             const SweepBoxes& sched, hipStream_t st) {
  const int64_t cy = (s.ny + CB - 1) / CB;
  const dim3 grid(unsigned(sched.ncols)), block(64);
  if (x_in) {
    hipLaunchKernelGGL((sor_wavefront<T, KIND, FWD, ULTRA, true, FD>), grid, block, 0, st, s,
                       static_cast<const T*>(x_in), static_cast<T*>(x_out),
                       static_cast<const T*>(b), T(omega), sched.ticket, sched.ticket_next, sched.err,
                       sched.order, cy, static_cast<T*>(sched.mailbox), static_cast<T*>(sched.mailbox_next));
  } else {
    hipLaunchKernelGGL((sor_wavefront<T, KIND, FWD, ULTRA, false, FD>), grid, block, 0, st, s,
                       static_cast<const T*>(x_in), static_cast<T*>(x_out),
                       static_cast<const T*>(b), T(omega), sched.ticket, sched.ticket_next, sched.err,
                       sched.order, cy, static_cast<T*>(sched.mailbox), static_cast<T*>(sched.mailbox_next));
  }
  LGPU_HIP_CHECK(hipGetLastError());
}

template <typename T, int KIND, bool FWD, bool ULTRA>
void launch4(const StencilDesc& s, const void* x_in, void* x_out, const void* b, double omega,
             const SweepBoxes& sched, hipStream_t st) {
  const bool fd = KIND == 0 && (sizeof(T) == 8 ? s.fast_div64 : s.fast_div32);
  if (fd) launch5<T, KIND, FWD, ULTRA, true>(s, x_in, x_out, b, omega, sched, st);
  else launch5<T, KIND, FWD, ULTRA, false>(s, x_in, x_out, b, omega, sched, st);
}

template <typename T, int KIND>
void launch2(const StencilDesc& s, const void* x_in, void* x_out, const void* b, double omega,
             bool fwd, bool ultra, const SweepBoxes& sched, hipStream_t st) {
  if (fwd) {
    if (ultra) launch4<T, KIND, true, true>(s, x_in, x_out, b, omega, sched, st);
    else launch4<T, KIND, true, false>(s, x_in, x_out, b, omega, sched, st);
  } else {
    if (ultra) launch4<T, KIND, false, true>(s, x_in, x_out, b, omega, sched, st);
    else launch4<T, KIND, false, false>(s, x_in, x_out, b, omega, sched, st);
  }
}

template <typename T, bool FWD, bool ULTRA>
void launch_r(const StencilDesc& s, const void* w, const void* x_in, void* x_out, const void* b,
              double omega, const SweepBoxes& sched, hipStream_t st) {
  const RecGeom g = rec_geom(s);
  const dim3 grid(unsigned(sched.ncols)), block(64);
  if (x_in) {
    hipLaunchKernelGGL((sor_wavefront_r<T, FWD, ULTRA, true>), grid, block, 0, st, s, g,
                       static_cast<const T*>(w), static_cast<const T*>(x_in), static_cast<T*>(x_out),
                       static_cast<const T*>(b), T(omega), sched.ticket, sched.ticket_next, sched.err,
                       sched.order, static_cast<T*>(sched.mailbox), static_cast<T*>(sched.mailbox_next));
  } else {
    hipLaunchKernelGGL((sor_wavefront_r<T, FWD, ULTRA, false>), grid, block, 0, st, s, g,
                       static_cast<const T*>(w), static_cast<const T*>(x_in), static_cast<T*>(x_out),
                       static_cast<const T*>(b), T(omega), sched.ticket, sched.ticket_next, sched.err,
                       sched.order, static_cast<T*>(sched.mailbox), static_cast<T*>(sched.mailbox_next));
  }
  LGPU_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_r2(const StencilDesc& s, const void* w, const void* x_in, void* x_out, const void* b,
               double omega, bool fwd, bool ultra, const SweepBoxes& sched, hipStream_t st) {
  if (fwd) {
    if (ultra) launch_r<T, true, true>(s, w, x_in, x_out, b, omega, sched, st);
    else launch_r<T, true, false>(s, w, x_in, x_out, b, omega, sched, st);
  } else {
    if (ultra) launch_r<T, false, true>(s, w, x_in, x_out, b, omega, sched, st);
    else launch_r<T, false, false>(s, w, x_in, x_out, b, omega, sched, st);
  }
}

template <typename T, int KIND>
void launch_records(const StencilDesc& s, void* w, hipStream_t st) {
  const RecGeom g = rec_geom(s);
  const int64_t n = g.ncols * g.nc * kRecK * 64;
  const dim3 grid(unsigned(std::min<int64_t>((n + 255) / 256, 65536))), block(256);
  const bool fd = KIND == 0 && (sizeof(T) == 8 ? s.fast_div64 : s.fast_div32);
  if (fd) hipLaunchKernelGGL((sor_records_kernel<T, KIND, true>), grid, block, 0, st, s, g, static_cast<T*>(w));
  else hipLaunchKernelGGL((sor_records_kernel<T, KIND, false>), grid, block, 0, st, s, g, static_cast<T*>(w));
  LGPU_HIP_CHECK(hipGetLastError());
}

}  // namespace

// a record per column, chunk slot and lane; 0 when a column's records would
// not fit 32-bit byte offsets (the sweeps then evaluate the stencil)
size_t sor_weights_bytes(const StencilDesc& s, bool f32) {
  const RecGeom g = rec_geom(s);
  const int64_t es = int64_t(f32 ? sizeof(float) : sizeof(double));
  const int64_t per_col = g.nc * kRecK * 64 * kRecVals * es;
  if (per_col >= (int64_t(1) << 31)) return 0;
  return size_t((rec_elems(g) + halo_elems(g)) * es);
}

void stencil_sor_weights(const StencilDesc& s, bool f32, void* w, hipStream_t st) {
  if (f32) {
    if (s.kind == 0) launch_records<float, 0>(s, w, st);
    else launch_records<float, 1>(s, w, st);
  } else {
    if (s.kind == 0) launch_records<double, 0>(s, w, st);
    else launch_records<double, 1>(s, w, st);
  }
}

size_t sor_sched_ints(const StencilDesc&) { return 2; }

// one mailbox set: 16 lines of nz words per column (the kernels use 15 or 16),
// sized for f64 words; an operator holds two sets
size_t sor_mailbox_bytes(const StencilDesc& s) {
  const int64_t cx = (s.nx + CB - 1) / CB, cy = (s.ny + CB - 1) / CB;
  return size_t(2 * cx * cy * CB * s.nz) * sizeof(double);
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
                 double omega, bool forward, bool ultra, const SorSched& sched, hipStream_t st,
                 const void* weights) {
  LGPU_REQUIRE(x_in != x_out, "SOR sweep needs distinct input and output vectors");
  // a column's ten planes are addressed with 32-bit byte offsets below 2^31
  LGPU_REQUIRE(10 * s.ny * s.nz * int64_t(f32 ? 4 : 8) < (int64_t(1) << 31),
               "SOR sweep: the y-z plane is too large for 32-bit column offsets");
  LGPU_REQUIRE(sched.box != nullptr, "SOR sweep: no mailbox state");
  SorBoxState& bs = *sched.box;
  const int set = bs.next;
  const int words = f32 ? 4 : 8;
  const size_t set_bytes = sor_mailbox_bytes(s);
  char* const box0 = static_cast<char*>(sched.mailbox);
  SweepBoxes sb{sched.state + set, sched.state + (1 - set), sched.err, sched.order, sched.ncols,
                box0 + size_t(set) * set_bytes, box0 + size_t(1 - set) * set_bytes};
  // the set is cleared for this precision by the previous sweep, or here (the
  // error word, sched.err, is sticky until the host clears it)
  if (bs.clean[set] < words) {
    const int64_t dwords = int64_t(set_bytes / sizeof(uint32_t));
    const dim3 rg(unsigned(std::min<int64_t>((dwords + 255) / 256, 4096))), rb(256);
    hipLaunchKernelGGL(mailbox_reset, rg, rb, 0, st, static_cast<uint32_t*>(sb.mailbox), dwords,
                       sb.ticket);
    LGPU_HIP_CHECK(hipGetLastError());
  }
  bs.clean[0] = bs.clean[1] = 0;
  if (weights) {
    if (f32) launch_r2<float>(s, weights, x_in, x_out, b, omega, forward, ultra, sb, st);
    else launch_r2<double>(s, weights, x_in, x_out, b, omega, forward, ultra, sb, st);
  } else if (f32) {
    if (s.kind == 0) launch2<float, 0>(s, x_in, x_out, b, omega, forward, ultra, sb, st);
    else launch2<float, 1>(s, x_in, x_out, b, omega, forward, ultra, sb, st);
  } else {
    if (s.kind == 0) launch2<double, 0>(s, x_in, x_out, b, omega, forward, ultra, sb, st);
    else launch2<double, 1>(s, x_in, x_out, b, omega, forward, ultra, sb, st);
  }
  // the sweep leaves the other set and its ticket cleared in its word size
  bs.clean[1 - set] = words;
  bs.next = 1 - set;
}

}  // namespace lgpu
