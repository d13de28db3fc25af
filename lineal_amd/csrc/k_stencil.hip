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
constexpr int NW = 4;         // wavefronts per block, along y
constexpr int NT = WZ * NW;
// y rows per wavefront: f32 carries more rows (register budget)
template <typename T>
constexpr int kRows = sizeof(T) == 4 ? 4 : 4;

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

__device__ __forceinline__ int read_lane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
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
// or the material id widened to int); face(a, b, d) is the connection value.
// FD selects the scaling-free division for the harmonic mean.
template <typename T, int K, bool FD>
struct Cells;
template <typename T, bool FD>
struct Cells<T, 0, FD> {
  using E = T;
  using R = T;
  __device__ __forceinline__ static const E* ptr(const StencilDesc& s) { return coeff_ptr<T>(s); }
  __device__ __forceinline__ static R widen(E e) { return e; }
  __device__ __forceinline__ static T value(const StencilDesc&, R r) { return r; }
  // stationary-diffusion-values.hpp:131-135: H(c_i, c_j) * q_d
  __device__ __forceinline__ static T face(const StencilDesc&, R a, R b, T qd, int) {
    return hmean_sel<T, FD>(a, b) * qd;
  }
};
template <typename T, bool FD>
struct Cells<T, 1, FD> {
  using E = uint8_t;
  using R = int;
  __device__ __forceinline__ static const E* ptr(const StencilDesc& s) { return s.mat; }
  __device__ __forceinline__ static R widen(E e) { return int(e); }
  __device__ __forceinline__ static T value(const StencilDesc& s, R r) { return T(s.mc[r]); }
  // stationary-diffusion-lookup.hpp:129-136: T_d[m_i * 4 + m_j]
  __device__ __forceinline__ static T face(const StencilDesc& s, R a, R b, T, int d) {
    return CellInfo<T, 1>::conn(s, uint8_t(a), uint8_t(b), T(0), d);
  }
};

int64_t march_len(const StencilDesc& s, int yr) {
  if (const char* e = std::getenv("LINEAL_AMD_STENCIL_XS")) {
    const int64_t v = std::atoll(e);
    if (v >= 1 && v <= 256) return v;
  }
  // longest march (fewest extra boundary planes) that still gives >= 8192
  // wavefronts, i.e. eight per SIMD of the 256 CUs
  const int64_t tiles = ((s.nz + WZ - 1) / WZ) * ((s.ny + yr - 1) / yr);
  int64_t xs = 32;
  while (xs > 4 && tiles * ((s.nx + xs - 1) / xs) < 8192) xs /= 2;
  return xs;
}

template <int V>
using Slot = std::integral_constant<int, V>;

template <typename T, int K, bool FD, int MODE, bool HAS_X>
__global__ __launch_bounds__(NT) void stencil_kernel(StencilDesc s, const T* __restrict__ xv,
                                                     const T* __restrict__ bv, T* __restrict__ yv,
                                                     T omega, int xs, int tz, int ty, int tx,
                                                     double* __restrict__ partials) {
  constexpr int YR = kRows<T>;
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
  const int lane = threadIdx.x & (WZ - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WZ);
  const int64_t hw = blockIdx.x;
  const int64_t tile = xcd_tile(hw, int64_t(tz) * ty * tx);
  const int bz = int(tile % tz), by = int((tile / tz) % ty), bx = int(tile / (int64_t(tz) * ty));
  const int nx = int(s.nx), ny = int(s.ny), nz = int(s.nz);
  const int64_t nyz = int64_t(ny) * nz;
  const int z0 = bz * WZ;
  const int z = z0 + lane;
  const bool zin = z < nz;
  const bool has_pz = z + 1 < nz;
  const uint32_t zc = uint32_t(min(z, nz - 1));
  const int y0 = (by * NW + wave) * YR;
  const T q0 = T(s.q[0]), q1 = T(s.q[1]), q2 = T(s.q[2]);
  const T bnd = T(2) * q0;
  const T m1w = T(1) - omega;
  double part = 0.0;

  if (y0 < ny) {
    const E* const cbase = C::ptr(s);
    const uint32_t pbe = uint32_t(nyz) * SE, pbt = uint32_t(nyz) * ST;  // plane bytes
    const int x0 = bx * xs;
    const int x1 = min(x0 + xs, nx);
    // wave-uniform byte offsets of the rows within a plane
    uint32_t roe[YR], rot[YR];
    bool rin[YR];
#pragma unroll
    for (int k = 0; k < YR; ++k) {
      rin[k] = y0 + k < ny;
      const uint32_t r = uint32_t(min(y0 + k, ny - 1)) * uint32_t(nz);
      roe[k] = r * SE;
      rot[k] = r * ST;
    }
    const bool has_ym = y0 > 0, has_yp = y0 + YR < ny;
    const uint32_t om = uint32_t(max(y0 - 1, 0)) * uint32_t(nz);
    const uint32_t op = uint32_t(min(y0 + YR, ny - 1)) * uint32_t(nz);
    // plane resources: planes outside the box (ZL) and rows outside it get an
    // empty range, so their loads return 0 and their stores are dropped
    auto crs = [&](int p, bool ok) {
      if constexpr (ZL) {
        return plane_rsrc(cbase, int64_t(p) * nyz, ok && p >= 0 && p < nx ? pbe : 0u);
      } else {
        return plane_rsrc(cbase, int64_t(min(max(p, 0), nx - 1)) * nyz, pbe);
      }
    };
    auto trs = [&](const T* a, int p, bool ok) {
      if constexpr (ZL) {
        return plane_rsrc(a, int64_t(p) * nyz, ok && p >= 0 && p < nx ? pbt : 0u);
      } else {
        return plane_rsrc(a, int64_t(min(max(p, 0), nx - 1)) * nyz, pbt);
      }
    };
    // per-lane byte offsets (lanes beyond nz: out of range when ZL)
    const uint32_t vze = ZL && !zin ? kOut : zc * SE, vzt = ZL && !zin ? kOut : zc * ST;
    // z halo: lane 0 holds cell z0-1, lane 63 cell z0+64 (zero where they do not exist)
    const bool halo_lane = (lane == 0 && z0 > 0) || (lane == WZ - 1 && z0 + WZ < nz);
    const uint32_t hz = uint32_t(lane == 0 ? max(z0 - 1, 0) : min(z0 + WZ, nz - 1));
    const uint32_t hze = hz * SE, hzt = hz * ST;
    // batched face between z0-1 and z0: lane k < YR evaluates it for row k
    const bool fh_lane = lane < YR && z0 > 0;
    const uint32_t fh_cell =
        uint32_t(min(y0 + min(lane, YR - 1), ny - 1)) * uint32_t(nz) + uint32_t(max(z0 - 1, 0));

    // three plane slots of the wavefront's own rows, used round robin: the
    // current plane, the next plane, and the prefetch target (plane x0-1 at
    // the start); the right-hand side runs one plane ahead in the same slots
    R pc[3][YR];
    T px[3][YR];
    T pb[3][YR];
    auto load_rows = [&](int slot, int p) {
#pragma unroll
      for (int k = 0; k < YR; ++k) {
        pc[slot][k] = C::widen(bload<E>(crs(p, rin[k]), vze, roe[k]));
        if constexpr (HAS_X) px[slot][k] = bload<T>(trs(xv, p, rin[k]), vzt, rot[k]);
        else px[slot][k] = T(0);
      }
    };
    auto load_b = [&](int slot, int p) {
      if constexpr (NEED_B) {
#pragma unroll
        for (int k = 0; k < YR; ++k) pb[slot][k] = bload<T>(trs(bv, p, rin[k]), vzt, rot[k]);
      }
    };

    // carried state: -x face weights and -x neighbour values of the rows
    T wmx[YR], xmx[YR];
    load_rows(0, x0);
    load_rows(1, x0 + 1);
    load_b(0, x0);
    load_rows(2, x0 - 1);
#pragma unroll
    for (int k = 0; k < YR; ++k) {
      const T w = C::face(s, pc[2][k], pc[0][k], q0, 0);
      wmx[k] = ZL || x0 > 0 ? w : T(0);
      xmx[k] = ZL || x0 > 0 ? px[2][k] : T(0);
    }

    // one plane p with current rows in slot S0, next rows in S1; prefetches
    // plane p+2 into S2 and the right-hand side of plane p+1 into S1
    auto plane = [&](auto S0, auto S1, auto S2, int p) {
      constexpr int s0 = decltype(S0)::value, s1 = decltype(S1)::value, s2 = decltype(S2)::value;
      load_rows(s2, p + 2);
      load_b(s1, p + 1);
      // halo data of this plane (rows of the neighbouring wavefronts: cache hits)
      const R cym = C::widen(bload<E>(crs(p, has_ym), vze, om * SE));
      const R cyp = C::widen(bload<E>(crs(p, has_yp), vze, op * SE));
      T xym = T(0), xyp = T(0);
      R hc[YR];
      T hx[YR];
#pragma unroll
      for (int k = 0; k < YR; ++k) {
        hc[k] = R(0);
        hx[k] = T(0);
      }
      R fa = R(0), fb = R(0);
      if constexpr (HAS_X) {
        xym = bload<T>(trs(xv, p, has_ym), vzt, om * ST);
        xyp = bload<T>(trs(xv, p, has_yp), vzt, op * ST);
        if (halo_lane) {
#pragma unroll
          for (int k = 0; k < YR; ++k) hx[k] = bload<T>(trs(xv, p, rin[k]), hzt, rot[k]);
        }
      }
      if (halo_lane) {
#pragma unroll
        for (int k = 0; k < YR; ++k) hc[k] = C::widen(bload<E>(crs(p, rin[k]), hze, roe[k]));
      }
      if (fh_lane) {
        const Rsrc rf = crs(p, true);
        fa = C::widen(bload<E>(rf, fh_cell * SE, 0));
        fb = C::widen(bload<E>(rf, fh_cell * SE + SE, 0));
      }
      const bool xlo = p == 0, xhi = p + 1 == nx;
      if constexpr (!ZL) {
        xym = has_ym ? xym : T(0);
        xyp = has_yp ? xyp : T(0);
      }

      // faces: z halo (batched), y faces
      T fh = C::face(s, fa, fb, q2, 2);
      if constexpr (!ZL) fh = fh_lane ? fh : T(0);
      T gy[YR + 1];
      gy[0] = C::face(s, cym, pc[s0][0], q1, 1);
#pragma unroll
      for (int k = 1; k < YR; ++k) gy[k] = C::face(s, pc[s0][k - 1], pc[s0][k], q1, 1);
      gy[YR] = C::face(s, pc[s0][YR - 1], cyp, q1, 1);
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
        T wpz = C::face(s, cc, cpz, q2, 2);
        const T xmz = shift<kWaveShr1>(hx[k], xc);
        T xpz = shift<kWaveShl1>(hx[k], xc);
        // +x face and neighbour
        T wpx = C::face(s, cc, pc[s1][k], q0, 0);
        T xpx = px[s1][k];
        const int km = k > 0 ? k - 1 : 0, kp = k + 1 < YR ? k + 1 : k;
        const T xmy = k == 0 ? xym : px[s0][km];
        T xpy = k == YR - 1 ? xyp : px[s0][kp];
        if constexpr (!ZL) {
          wpz = has_pz ? wpz : T(0);
          xpz = has_pz ? xpz : T(0);
          wpx = xhi ? T(0) : wpx;
          xpx = xhi ? T(0) : xpx;
          if (k < YR - 1) xpy = rin[kp] ? xpy : T(0);
        }
        const T wmz = shift<kWaveShr1>(read_lane(fh, k), wpz);
        const T wmy = gy[k], wpy = gy[k + 1];
        // diagonal in the row iterate's order
        const T ci = C::value(s, cc);
        T d = T(0);
        if (xlo) d += bnd * ci;
        if (xhi) d += bnd * ci;
        d += wmx[k];
        d += wmy;
        d += wmz;
        d += wpz;
        d += wpy;
        d += wpx;
        if (s.zero_to_one && d == T(0)) d = T(1);
        // ZL: rows and lanes outside the box store out of range (dropped) and
        // contribute exact zeros to the partial sums
        const bool live = ZL || (zin && rin[k]);
        const Rsrc ry = trs(yv, p, rin[k]);
        if constexpr (MODE == SM_JACOBI) {
          // jacobi/sweep.hpp:66-87: accum = b - sum a_ij x_j (fnmadd with a_ij = -w)
          T acc = pb[s0][k];
          acc = fma(wmx[k], xmx[k], acc);
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
          sum = fma(-wmx[k], xmx[k], sum);
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
            if (live) bstore<T>(pb[s0][k] - sum, ry, vzt, rot[k]);
          } else {
            const T rr = pb[s0][k] - sum;
            if (live) {
              bstore<T>(rr, ry, vzt, rot[k]);
              part += double(rr) * double(rr);
            }
          }
        }
        wmx[k] = wpx;
        xmx[k] = xc;
      }
    };

    for (int p = x0;;) {
      if (p >= x1) break;
      plane(Slot<0>{}, Slot<1>{}, Slot<2>{}, p++);
      if (p >= x1) break;
      plane(Slot<1>{}, Slot<2>{}, Slot<0>{}, p++);
      if (p >= x1) break;
      plane(Slot<2>{}, Slot<0>{}, Slot<1>{}, p++);
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

struct Grid {
  int64_t xs, tz, ty, tx;
  int64_t blocks() const { return tz * ty * tx; }
};

template <typename T>
Grid grid_of(const StencilDesc& s) {
  constexpr int yr = kRows<T>;
  Grid g;
  g.xs = march_len(s, yr);
  g.tz = (s.nz + WZ - 1) / WZ;
  g.ty = (s.ny + NW * yr - 1) / (NW * yr);
  g.tx = (s.nx + g.xs - 1) / g.xs;
  return g;
}

template <typename T, int K, bool FD, int MODE>
void launch_mode(const StencilDesc& s, const void* x, const void* b, void* y, double omega,
                 double* partials, hipStream_t st) {
  const Grid g = grid_of<T>(s);
  const auto xp = static_cast<const T*>(x);
  const auto bp = static_cast<const T*>(b);
  const auto yp = static_cast<T*>(y);
  if (x) {
    hipLaunchKernelGGL((stencil_kernel<T, K, FD, MODE, true>), dim3(g.blocks()), dim3(NT), 0, st,
                       s, xp, bp, yp, T(omega), int(g.xs), int(g.tz), int(g.ty), int(g.tx),
                       partials);
  } else {
    hipLaunchKernelGGL((stencil_kernel<T, K, FD, MODE, false>), dim3(g.blocks()), dim3(NT), 0, st,
                       s, xp, bp, yp, T(omega), int(g.xs), int(g.tz), int(g.ty), int(g.tx),
                       partials);
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

int64_t stencil_partials(const StencilDesc& s) {
  return std::max(grid_of<float>(s).blocks(), grid_of<double>(s).blocks());
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
