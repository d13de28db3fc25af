// Matrix-free stencil passes: SpMV, residual, fused SpMV+dot, fused
// residual+norm, Jacobi, and the Dirichlet right-hand side.
//
// Layout: cell (x, y, z) at index (x*ny + y)*nz + z (z contiguous).
//
// A wavefront owns YR consecutive y rows of 64 z cells (lane = z) and marches
// along x over XS planes; a block stacks NW wavefronts along y.  Every face
// weight is evaluated once and shared by the two cells it connects (the
// harmonic mean is symmetric bit for bit, and so are the material tables):
//   +x faces are carried from one plane to the next in registers,
//   y faces between the wavefront's rows are shared in registers,
//   z faces move one lane over by a DPP whole-wave shift; the face between
//   lane 0 and the previous z cell is evaluated for all YR rows by one batched
//   evaluation in lanes 0..YR-1.
// That is 3 YR + 2 face evaluations per YR x 64 cells and plane instead of 6
// per cell.  Neighbour values along z also come from DPP shifts, so a plane of
// a wavefront costs YR coefficient and value loads, 2 halo rows along y
// (cache hits of the neighbouring wavefronts' rows) and a few single-lane halo
// loads along z.  All loads are issued unconditionally at clamped in-box
// indices (absent neighbours are masked to the zero the row evaluation expects)
// and the rows of the next plane are fetched one plane ahead.
//
// Every cell performs exactly the operations of the reference's valued,
// unordered row iterate (tensor/dynamic/stencil/matrix/algorithm.hpp:91-227) in
// its order, so the results equal the row-by-row evaluation bit for bit.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "device.hpp"

namespace lgpu {

namespace {

constexpr int WZ = 64;        // lanes of a wavefront, along z
constexpr int kMaxWaves = 8;   // wavefronts per block (at most; caps VGPRs at 256)
// y rows per wavefront, and plane slots held in registers (the current plane,
// the next one, and kSlots - 2 planes in flight ahead of them)
template <typename T>
constexpr int kRows = sizeof(T) == 4 ? 4 : 4;
template <typename T>
constexpr int kSlots = sizeof(T) == 4 ? 3 : 3;

__device__ __forceinline__ float read_lane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// Buffer resources (V#) over one plane of an array: the base address sits in
// scalar registers, loads and stores take a 32-bit per-lane byte offset and a
// 32-bit wave-uniform byte offset, so no 64-bit address arithmetic runs in
// vector registers.
constexpr int kRsrcWord3 = 0x00020000;  // gfx9: 32-bit data format, no swizzle

using Rsrc = __amdgpu_buffer_rsrc_t;

template <typename E>
__device__ __forceinline__ Rsrc plane_rsrc(const E* base, int64_t elem_offset, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<E*>(base + elem_offset), short(0),
                                           int(bytes), kRsrcWord3);
}

template <typename E>
__device__ __forceinline__ E bload(Rsrc r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(E) == 1) {
    return E(__builtin_amdgcn_raw_buffer_load_b8(r, voff, soff, 0));
  } else if constexpr (sizeof(E) == 4) {
    return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  } else {
    static_assert(sizeof(E) == 8);
    return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  }
}

template <typename E>
__device__ __forceinline__ void bstore(E v, Rsrc r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(E) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
  } else {
    static_assert(sizeof(E) == 8);
    using V2 = __attribute__((ext_vector_type(2))) unsigned int;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, voff, soff, 0);
  }
}

// The two valuators.  Cell data E is held in registers as R (the coefficient,
// or the material id widened to int); face(tab, a, b, q_d, d) is the
// connection value and value(tab, r) the cell's coefficient, where tab is the
// block's LDS copy of the lookup tables (unused by the values valuator).  FD
// selects the scaling-free division for the harmonic mean.
template <typename T, int K, bool FD>
struct Cells;
template <typename T, bool FD>
struct Cells<T, 0, FD> {
  using E = T;
  using R = T;
  static constexpr int kTab = 1;
  __device__ __forceinline__ static const E* ptr(const StencilDesc& s) { return coeff_ptr<T>(s); }
  __device__ __forceinline__ static R widen(E e) { return e; }
  __device__ __forceinline__ static T value(const T*, R r) { return r; }
  // stationary-diffusion-values.hpp:131-135: H(c_i, c_j) * q_d
  __device__ __forceinline__ static T face(const T*, R a, R b, T qd, int) {
    return hmean_sel<T, FD>(a, b) * qd;
  }
  __device__ __forceinline__ static void fill(const StencilDesc&, T*) {}
};
template <typename T, bool FD>
struct Cells<T, 1, FD> {
  using E = uint8_t;
  using R = int;
  static constexpr int kTab = 3 * 16 + 4;  // T_x, T_y, T_z (4 x 4 each), then C[0..3]
  __device__ __forceinline__ static const E* ptr(const StencilDesc& s) { return s.mat; }
  __device__ __forceinline__ static R widen(E e) { return int(e); }
  __device__ __forceinline__ static T value(const T* tab, R r) { return tab[48 + r]; }
  // stationary-diffusion-lookup.hpp:129-136: T_d[m_i * 4 + m_j]
  __device__ __forceinline__ static T face(const T* tab, R a, R b, T, int d) {
    return tab[d * 16 + a * 4 + b];
  }
  // thread 0 copies the tables into LDS with constant indices (kernel
  // argument reads stay scalar); the caller synchronises
  __device__ __forceinline__ static void fill(const StencilDesc& s, T* tab) {
    if (threadIdx.x != 0) return;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        if constexpr (sizeof(T) == 8) tab[d * 16 + m] = T(s.tab64[d][m]);
        else tab[d * 16 + m] = T(s.tab32[d][m]);
      }
#pragma unroll
    for (int r = 0; r < 4; ++r) tab[48 + r] = T(s.mc[r]);
  }
};

template <int V>
using Slot = std::integral_constant<int, V>;

// Launch shape: a block holds nwz x nwy wavefronts (z fastest) covering
// 64 nwz z cells by kRows * nwy y rows, and marches xs planes; tx x tz x ty
// blocks with x fastest, so that x neighbours are dispatched together.
struct Shape {
  int xs, nwz, nwy, tz, ty, tx;
  __host__ __device__ int64_t blocks() const { return int64_t(tz) * ty * tx; }
  __host__ __device__ int waves() const { return nwz * nwy; }
};

// LDS exchange area of one block, double buffered by plane parity: the first
// and last row of every wavefront (its y neighbours' halo rows) and the lane 0
// and lane 63 values of every row (its z neighbours' halo cells).
template <typename T, typename R, int YR>
struct Exchange {
  T* rx;   // [2 parity][2 first/last][waves][64]
  R* rc;
  T* zx;   // [2 parity][2 lane 0/63][waves][YR]
  R* zc;
  int w;
  __device__ Exchange(void* base, int waves) : w(waves) {
    char* b = static_cast<char*>(base);
    rx = reinterpret_cast<T*>(b);
    b += sizeof(T) * 4 * waves * WZ;
    zx = reinterpret_cast<T*>(b);
    b += sizeof(T) * 4 * waves * YR;
    rc = reinterpret_cast<R*>(b);
    b += sizeof(R) * 4 * waves * WZ;
    zc = reinterpret_cast<R*>(b);
  }
  __host__ __device__ static size_t bytes(int waves) {
    return (sizeof(T) + sizeof(R)) * 4 * size_t(waves) * (WZ + YR);
  }
  __device__ int row(int par, int side, int wave, int lane) const {
    return ((par * 2 + side) * w + wave) * WZ + lane;
  }
  __device__ int cell(int par, int side, int wave, int k) const {
    return ((par * 2 + side) * w + wave) * YR + k;
  }
};

template <typename T, int K, bool FD, int MODE, bool HAS_X>
__global__ __launch_bounds__(WZ * kMaxWaves) void stencil_kernel(StencilDesc s,
                                                                 const T* __restrict__ xv,
                                                                 const T* __restrict__ bv,
                                                                 T* __restrict__ yv, T omega,
                                                                 Shape g,
                                                                 double* __restrict__ partials) {
  constexpr int YR = kRows<T>;
  constexpr int NS = kSlots<T>;
  static_assert(NS >= 3 && NS <= 4, "plane slots");
  constexpr bool NEED_B = MODE == SM_RESID || MODE == SM_RESID_NORM || MODE == SM_JACOBI;
  // Values valuator: a cell outside the box reads as coefficient 0 and value 0
  // (out-of-range buffer accesses return 0 and drop stores), and H(c, 0) = 0,
  // so absent faces and neighbours come out as the +0 the row iterate expects
  // without any masking.  The lookup valuator's material 0 is a real material,
  // so it masks explicitly.
  constexpr bool ZL = K == 0;
  using C = Cells<T, K, FD>;
  using E = typename C::E;
  using R = typename C::R;
  constexpr uint32_t SE = sizeof(E), ST = sizeof(T);
  constexpr uint32_t kOut = 0x80000000u;  // per-lane offset beyond every plane
  extern __shared__ double lds_raw[];
  const Exchange<T, R, YR> ex(lds_raw, g.waves());
  __shared__ T tab[C::kTab];
  C::fill(s, tab);
  __syncthreads();
  const int lane = threadIdx.x & (WZ - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WZ);
  const int64_t hw = blockIdx.x;
  const int64_t tile = xcd_tile(hw, g.blocks());
  const int bx = int(tile % g.tx), bz = int((tile / g.tx) % g.tz);
  const int by = int(tile / (int64_t(g.tx) * g.tz));
  const int wz = wave % g.nwz, wy = wave / g.nwz;
  const int nx = int(s.nx), ny = int(s.ny), nz = int(s.nz);
  const int64_t nyz = int64_t(ny) * nz;
  const int z0 = (bz * g.nwz + wz) * WZ;
  const int z = z0 + lane;
  const bool zin = z < nz;
  const bool has_pz = z + 1 < nz;
  const uint32_t zc = uint32_t(min(z, nz - 1));
  const int y0 = (by * g.nwy + wy) * YR;
  const T q0 = T(s.q[0]), q1 = T(s.q[1]), q2 = T(s.q[2]);
  const T bnd = T(2) * q0;
  const T adv = T(s.adv);
  const T m1w = T(1) - omega;
  const E* const cbase = C::ptr(s);
  const uint32_t pbe = uint32_t(nyz) * SE, pbt = uint32_t(nyz) * ST;  // plane bytes
  const int x0 = bx * g.xs;
  const int x1 = min(x0 + g.xs, nx);
  // Every wavefront of the block runs the march (the exchange needs all of
  // them at each barrier); wavefronts or rows outside the box work on clamped
  // (lookup) or empty (values) data and store nothing.
  uint32_t roe[YR], rot[YR];
  bool rin[YR];
#pragma unroll
  for (int k = 0; k < YR; ++k) {
    rin[k] = y0 + k < ny;
    const uint32_t r = uint32_t(min(y0 + k, ny - 1)) * uint32_t(nz);
    roe[k] = r * SE;
    rot[k] = r * ST;
  }
  const bool has_ym = y0 > 0 && y0 - 1 < ny, has_yp = y0 + YR < ny;
  const uint32_t om = uint32_t(min(max(y0 - 1, 0), ny - 1)) * uint32_t(nz);
  const uint32_t op = uint32_t(min(y0 + YR, ny - 1)) * uint32_t(nz);
  // plane resources: planes outside the box (ZL), rows outside it (ZL) and
  // planes outside [lo, hi] get an empty range, so their loads return 0
  // without touching memory and their stores are dropped
  auto crs = [&](int p, bool ok, int lo, int hi) {
    const bool in = p >= lo && p <= hi;
    if constexpr (ZL) {
      return plane_rsrc(cbase, int64_t(p) * nyz, ok && in && p >= 0 && p < nx ? pbe : 0u);
    } else {
      return plane_rsrc(cbase, int64_t(min(max(p, 0), nx - 1)) * nyz, in ? pbe : 0u);
    }
  };
  auto trs = [&](const T* a, int p, bool ok, int lo, int hi) {
    const bool in = p >= lo && p <= hi;
    if constexpr (ZL) {
      return plane_rsrc(a, int64_t(p) * nyz, ok && in && p >= 0 && p < nx ? pbt : 0u);
    } else {
      return plane_rsrc(a, int64_t(min(max(p, 0), nx - 1)) * nyz, in ? pbt : 0u);
    }
  };
  // per-lane byte offsets (lanes beyond nz: out of range when ZL)
  const uint32_t vze = ZL && !zin ? kOut : zc * SE, vzt = ZL && !zin ? kOut : zc * ST;
  // z halo cells that lie outside the block: lane 0 of the block's first
  // wavefront column (cell z0-1) and lane 63 of its last (cell z0+64)
  const bool zhalo_mem = (lane == 0 && wz == 0 && z0 > 0) ||
                         (lane == WZ - 1 && wz == g.nwz - 1 && z0 + WZ < nz);
  const uint32_t hz = uint32_t(lane == 0 ? max(z0 - 1, 0) : min(z0 + WZ, nz - 1));
  const uint32_t hze = hz * SE, hzt = hz * ST;
  // batched face between z0-1 and z0: lane k < YR evaluates it for row k
  const bool fh_lane = lane < YR && z0 > 0;
  const uint32_t fh_cell =
      uint32_t(min(y0 + min(lane, YR - 1), ny - 1)) * uint32_t(nz) + uint32_t(max(z0 - 1, 0));

  double part = 0.0;
  // NS plane slots of the wavefront's own rows, used round robin: the current
  // plane, the next plane (in march direction) and planes in flight; the last
  // slot holds the plane behind the march start at first
  R pc[NS][YR];
  T px[NS][YR];
  // carried state: the face weight and value behind the march
  T wbh[YR], xbh[YR];
  // values from outside the block's own rows (right-hand side, y halo rows,
  // z halo cells, the batched z face's cell), loaded one plane ahead into the
  // slot of their plane
  struct Halo {
    R cym, cyp, fa;
    T xym, xyp;
    R hc[YR];
    T hx[YR];
    T pb[YR];
  };
  Halo hl[NS];

  auto march = [&](auto DIR) {
    constexpr int dir = decltype(DIR)::value;
    const int first = dir > 0 ? x0 : x1 - 1;
    auto load_rows = [&](int slot, int p) {
#pragma unroll
      for (int k = 0; k < YR; ++k) {
        pc[slot][k] = C::widen(bload<E>(crs(p, rin[k], x0 - 1, x1), vze, roe[k]));
        if constexpr (HAS_X) px[slot][k] = bload<T>(trs(xv, p, rin[k], x0 - 1, x1), vzt, rot[k]);
        else px[slot][k] = T(0);
      }
    };
    // planes outside [x0, x1) get empty ranges: no memory traffic
    auto load_halo = [&](Halo& h, int p) {
      if constexpr (NEED_B) {
#pragma unroll
        for (int k = 0; k < YR; ++k) h.pb[k] = bload<T>(trs(bv, p, rin[k], x0, x1 - 1), vzt, rot[k]);
      }
      h.cym = h.cyp = h.fa = R(0);
      h.xym = h.xyp = T(0);
#pragma unroll
      for (int k = 0; k < YR; ++k) {
        h.hc[k] = R(0);
        h.hx[k] = T(0);
      }
      if (wy == 0) {
        h.cym = C::widen(bload<E>(crs(p, has_ym, x0, x1 - 1), vze, om * SE));
        if constexpr (HAS_X) h.xym = bload<T>(trs(xv, p, has_ym, x0, x1 - 1), vzt, om * ST);
      }
      if (wy == g.nwy - 1) {
        h.cyp = C::widen(bload<E>(crs(p, has_yp, x0, x1 - 1), vze, op * SE));
        if constexpr (HAS_X) h.xyp = bload<T>(trs(xv, p, has_yp, x0, x1 - 1), vzt, op * ST);
      }
      if (zhalo_mem) {
#pragma unroll
        for (int k = 0; k < YR; ++k) {
          h.hc[k] = C::widen(bload<E>(crs(p, rin[k], x0, x1 - 1), hze, roe[k]));
          if constexpr (HAS_X) h.hx[k] = bload<T>(trs(xv, p, rin[k], x0, x1 - 1), hzt, rot[k]);
        }
      }
      if (wz == 0 && fh_lane) h.fa = C::widen(bload<E>(crs(p, true, x0, x1 - 1), fh_cell * SE, 0));
    };
    load_rows(0, first);
    load_halo(hl[0], first);
#pragma unroll
    for (int j = 1; j + 1 < NS; ++j) load_rows(j, first + j * dir);
    load_rows(NS - 1, first - dir);
#pragma unroll
    for (int k = 0; k < YR; ++k) {
      const T w = C::face(tab, pc[NS - 1][k], pc[0][k], q0, 0);
      const bool behind = dir > 0 ? first > 0 : first + 1 < nx;
      wbh[k] = ZL || behind ? w : T(0);
      xbh[k] = ZL || behind ? px[NS - 1][k] : T(0);
    }

    // plane p, the J-th of a round of NS: current rows in slot J, next rows in
    // slot J+1; prefetches plane p+(NS-1)dir into slot J-1
    auto plane = [&](auto J, int p) {
      constexpr int s0 = decltype(J)::value % NS, s1 = (decltype(J)::value + 1) % NS;
      constexpr int s2 = (decltype(J)::value + NS - 1) % NS;
      // This plane's values from outside the block were loaded one plane
      // ago.  The next plane's are issued before the prefetch of plane
      // p+(NS-1)dir: memory operations complete in issue order, so waiting for
      // them one plane later does not wait for that prefetch.
      load_halo(hl[s1], p + dir);
      load_rows(s2, p + (NS - 1) * dir);
      const Halo& h = hl[s0];
      T pb[YR];
      R hc[YR];
      T hx[YR];
#pragma unroll
      for (int k = 0; k < YR; ++k) {
        pb[k] = h.pb[k];
        hc[k] = h.hc[k];
        hx[k] = h.hx[k];
      }
      R cym = h.cym, cyp = h.cyp, fa = h.fa;
      T xym = h.xym, xyp = h.xyp;
      // publish this plane's boundary rows and cells
      const int par = p & 1;
      ex.rc[ex.row(par, 0, wave, lane)] = pc[s0][0];
      ex.rc[ex.row(par, 1, wave, lane)] = pc[s0][YR - 1];
      if constexpr (HAS_X) {
        ex.rx[ex.row(par, 0, wave, lane)] = px[s0][0];
        ex.rx[ex.row(par, 1, wave, lane)] = px[s0][YR - 1];
      }
      if (lane == 0 || lane == WZ - 1) {
        const int side = lane == 0 ? 0 : 1;
#pragma unroll
        for (int k = 0; k < YR; ++k) {
          ex.zc[ex.cell(par, side, wave, k)] = pc[s0][k];
          if constexpr (HAS_X) ex.zx[ex.cell(par, side, wave, k)] = px[s0][k];
        }
      }
      __syncthreads();
      // halo data inside the block from the neighbouring wavefronts
      if (wy > 0) {
        cym = ex.rc[ex.row(par, 1, wave - g.nwz, lane)];
        if constexpr (HAS_X) xym = ex.rx[ex.row(par, 1, wave - g.nwz, lane)];
      }
      if (wy < g.nwy - 1) {
        cyp = ex.rc[ex.row(par, 0, wave + g.nwz, lane)];
        if constexpr (HAS_X) xyp = ex.rx[ex.row(par, 0, wave + g.nwz, lane)];
      }
      if ((lane == 0 && wz > 0) || (lane == WZ - 1 && wz < g.nwz - 1)) {
        const int src = lane == 0 ? wave - 1 : wave + 1, side = lane == 0 ? 1 : 0;
#pragma unroll
        for (int k = 0; k < YR; ++k) {
          hc[k] = ex.zc[ex.cell(par, side, src, k)];
          if constexpr (HAS_X) hx[k] = ex.zx[ex.cell(par, side, src, k)];
        }
      }
      R fb = R(0);
      if (lane < YR) {
        if (wz > 0) fa = ex.zc[ex.cell(par, 1, wave - 1, lane)];
        fb = ex.zc[ex.cell(par, 0, wave, lane)];
      }
      const bool xlo = p == 0, xhi = p + 1 == nx;
      if constexpr (!ZL) {
        xym = has_ym ? xym : T(0);
        xyp = has_yp ? xyp : T(0);
      }

      // faces: z halo (batched), y faces
      T fh = C::face(tab, fa, fb, q2, 2);
      if constexpr (!ZL) fh = fh_lane ? fh : T(0);
      T gy[YR + 1];
      gy[0] = C::face(tab, cym, pc[s0][0], q1, 1);
#pragma unroll
      for (int k = 1; k < YR; ++k) gy[k] = C::face(tab, pc[s0][k - 1], pc[s0][k], q1, 1);
      gy[YR] = C::face(tab, pc[s0][YR - 1], cyp, q1, 1);
      if constexpr (!ZL) {
        gy[0] = has_ym ? gy[0] : T(0);
#pragma unroll
        for (int k = 1; k < YR; ++k) gy[k] = rin[k] ? gy[k] : T(0);
        gy[YR] = has_yp ? gy[YR] : T(0);
      }

#pragma unroll
      for (int k = 0; k < YR; ++k) {
        const R cc = pc[s0][k];
        const T xc = px[s0][k];
        // z neighbours from the adjacent lanes; the halo lanes keep the halo
        const R cpz = shift<kWaveShl1>(hc[k], cc);
        T wpz = C::face(tab, cc, cpz, q2, 2);
        const T xmz = shift<kWaveShr1>(hx[k], xc);
        T xpz = shift<kWaveShl1>(hx[k], xc);
        // x face and neighbour ahead of the march
        T wah = C::face(tab, cc, pc[s1][k], q0, 0);
        T xah = px[s1][k];
        const int km = k > 0 ? k - 1 : 0, kp = k + 1 < YR ? k + 1 : k;
        const T xmy = k == 0 ? xym : px[s0][km];
        T xpy = k == YR - 1 ? xyp : px[s0][kp];
        if constexpr (!ZL) {
          const bool edge = dir > 0 ? xhi : xlo;
          wpz = has_pz ? wpz : T(0);
          xpz = has_pz ? xpz : T(0);
          wah = edge ? T(0) : wah;
          xah = edge ? T(0) : xah;
          if (k < YR - 1) xpy = rin[kp] ? xpy : T(0);
        }
        const T wmx = dir > 0 ? wbh[k] : wah, wpx = dir > 0 ? wah : wbh[k];
        // -x coupling: face weight plus the upwind inflow (no -x neighbour at x = 0)
        const T cmx = xlo ? T(0) : wmx + adv;
        const T xmx = dir > 0 ? xbh[k] : xah, xpx = dir > 0 ? xah : xbh[k];
        const T wmz = shift<kWaveShr1>(read_lane(fh, k), wpz);
        const T wmy = gy[k], wpy = gy[k + 1];
        // diagonal in the row iterate's order
        const T ci = C::value(tab, cc);
        T d = T(0);
        if (xlo) d += bnd * ci;
        if (xhi) d += bnd * ci;
        d += wmx;
        d += wmy;
        d += wmz;
        d += wpz;
        d += wpy;
        d += wpx;
        d += adv;
        if (s.zero_to_one && d == T(0)) d = T(1);
        // ZL: rows and lanes outside the box store out of range (dropped) and
        // contribute exact zeros to the partial sums
        const bool live = ZL || (zin && rin[k]);
        const Rsrc ry = trs(yv, p, rin[k], p, p);
        if constexpr (MODE == SM_JACOBI) {
          // jacobi/sweep.hpp:66-87: accum = b - sum a_ij x_j (fnmadd with a_ij = -w)
          T acc = pb[k];
          acc = fma(cmx, xmx, acc);
          acc = fma(wmy, xmy, acc);
          acc = fma(wmz, xmz, acc);
          acc = fma(wpz, xpz, acc);
          acc = fma(wpy, xpy, acc);
          acc = fma(wpx, xpx, acc);
          const T quot = acc / d;
          if (live) bstore<T>(fma(omega, quot, m1w * xc), ry, vzt, rot[k]);
        } else {
          // matrix-vector-product.hpp:71-82: sum = fmadd(a_ij, x_j, sum) in iterate order
          T sum = T(0);
          sum = fma(-cmx, xmx, sum);
          sum = fma(-wmy, xmy, sum);
          sum = fma(-wmz, xmz, sum);
          sum = fma(-wpz, xpz, sum);
          sum = fma(-wpy, xpy, sum);
          sum = fma(-wpx, xpx, sum);
          sum = fma(d, xc, sum);
          if constexpr (MODE == SM_SPMV) {
            if (live) bstore<T>(sum, ry, vzt, rot[k]);
          } else if constexpr (MODE == SM_SPMV_DOT) {
            if (live) {
              bstore<T>(sum, ry, vzt, rot[k]);
              part += double(xc) * double(sum);
            }
          } else if constexpr (MODE == SM_RESID) {
            if (live) bstore<T>(pb[k] - sum, ry, vzt, rot[k]);
          } else {
            const T rr = pb[k] - sum;
            if (live) {
              bstore<T>(rr, ry, vzt, rot[k]);
              part += double(rr) * double(rr);
            }
          }
        }
        wbh[k] = wah;
        xbh[k] = xc;
      }
    };

    for (int p = first, left = x1 - x0;;) {
      if (left-- <= 0) break;
      plane(Slot<0>{}, p);
      p += dir;
      if (left-- <= 0) break;
      plane(Slot<1>{}, p);
      p += dir;
      if (left-- <= 0) break;
      plane(Slot<2>{}, p);
      p += dir;
      if constexpr (NS > 3) {
        if (left-- <= 0) break;
        plane(Slot<3>{}, p);
        p += dir;
      }
    }
  };
  // x neighbours march in opposite directions, so the plane both of them read
  // at their common border is read by both at the same time (start or end)
  if (bx & 1) march(std::integral_constant<int, -1>{});
  else march(std::integral_constant<int, 1>{});

  if constexpr (MODE == SM_SPMV_DOT || MODE == SM_RESID_NORM) {
    const double t = block_sum_dyn(part);
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
  if (x == 0) sum += T(s.adv) * T(s.p_start);
  b[i] = sum;
}

int env_int(const char* name, int lo, int hi, int dflt) {
  if (const char* e = std::getenv(name)) {
    const int v = std::atoi(e);
    if (v >= lo && v <= hi) return v;
  }
  return dflt;
}

template <typename T>
Shape shape_of(const StencilDesc& s) {
  constexpr int yr = kRows<T>;
  Shape g;
  const int waves = env_int("LINEAL_AMD_STENCIL_WAVES", 1, kMaxWaves, 8);
  const int zt = int((s.nz + WZ - 1) / WZ);  // z tiles of one wavefront
  g.nwz = env_int("LINEAL_AMD_STENCIL_NWZ", 1, waves, std::min(zt, 4));
  while (g.nwz > 1 && (waves % g.nwz != 0 || g.nwz > zt)) --g.nwz;
  g.nwy = waves / g.nwz;
  g.tz = int((zt + g.nwz - 1) / g.nwz);
  g.ty = int((s.ny + int64_t(g.nwy) * yr - 1) / (int64_t(g.nwy) * yr));
  // longest march (fewest boundary planes) that still gives a block per CU
  // (f64) or two (f32); measured best at 256^3 (8 waves, 4 along z)
  const int64_t cols = int64_t(g.tz) * g.ty;
  const int64_t min_blocks = sizeof(T) == 8 ? 256 : 512;
  int xs = 32;
  while (xs > 4 && cols * ((s.nx + xs - 1) / xs) < min_blocks) xs /= 2;
  g.xs = env_int("LINEAL_AMD_STENCIL_XS", 1, 256, xs);
  g.tx = int((s.nx + g.xs - 1) / g.xs);
  return g;
}

template <typename T, int K, bool FD, int MODE>
void launch_mode(const StencilDesc& s, const void* x, const void* b, void* y, double omega,
                 double* partials, hipStream_t st) {
  const Shape g = shape_of<T>(s);
  const dim3 block(unsigned(WZ * g.waves()));
  using R = typename Cells<T, K, FD>::R;
  const size_t lds = Exchange<T, R, kRows<T>>::bytes(g.waves());
  const auto xp = static_cast<const T*>(x);
  const auto bp = static_cast<const T*>(b);
  const auto yp = static_cast<T*>(y);
  if (x) {
    hipLaunchKernelGGL((stencil_kernel<T, K, FD, MODE, true>), dim3(g.blocks()), block, lds, st,
                       s, xp, bp, yp, T(omega), g, partials);
  } else {
    hipLaunchKernelGGL((stencil_kernel<T, K, FD, MODE, false>), dim3(g.blocks()), block, lds, st,
                       s, xp, bp, yp, T(omega), g, partials);
  }
  LGPU_HIP_CHECK(hipGetLastError());
}

template <typename T, int K, bool FD>
void dispatch_mode(const StencilDesc& s, int mode, const void* x, const void* b, void* y,
                   double omega, double* partials, hipStream_t st) {
  switch (mode) {
    case SM_SPMV: launch_mode<T, K, FD, SM_SPMV>(s, x, b, y, omega, partials, st); break;
    case SM_RESID: launch_mode<T, K, FD, SM_RESID>(s, x, b, y, omega, partials, st); break;
    case SM_SPMV_DOT: launch_mode<T, K, FD, SM_SPMV_DOT>(s, x, b, y, omega, partials, st); break;
    case SM_RESID_NORM:
      launch_mode<T, K, FD, SM_RESID_NORM>(s, x, b, y, omega, partials, st);
      break;
    case SM_JACOBI: launch_mode<T, K, FD, SM_JACOBI>(s, x, b, y, omega, partials, st); break;
    default: throw Error{LGPU_INVALID_ARG, "unknown stencil mode"};
  }
}

}  // namespace

int64_t stencil_partials(const StencilDesc& s, bool f32) {
  // every block of the pass writes exactly one partial
  return f32 ? shape_of<float>(s).blocks() : shape_of<double>(s).blocks();
}

void stencil_apply(const StencilDesc& s, int mode, bool f32, const void* x, const void* b,
                   void* y, double omega, double* partials, hipStream_t st) {
  if (mode != SM_JACOBI) LGPU_REQUIRE(x != nullptr, "stencil pass needs an input vector");
  // the kernels address a plane of an array through a buffer resource with
  // 32-bit byte offsets
  LGPU_REQUIRE(s.nx < (int64_t(1) << 31) && s.ny * s.nz < (int64_t(1) << 28),
               "stencil plane too large (ny * nz must stay below 2^28)");
  if (f32) {
    if (s.kind != 0) dispatch_mode<float, 1, false>(s, mode, x, b, y, omega, partials, st);
    else if (s.fast_div32) dispatch_mode<float, 0, true>(s, mode, x, b, y, omega, partials, st);
    else dispatch_mode<float, 0, false>(s, mode, x, b, y, omega, partials, st);
  } else {
    if (s.kind != 0) dispatch_mode<double, 1, false>(s, mode, x, b, y, omega, partials, st);
    else if (s.fast_div64) dispatch_mode<double, 0, true>(s, mode, x, b, y, omega, partials, st);
    else dispatch_mode<double, 0, false>(s, mode, x, b, y, omega, partials, st);
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
