// Internal declarations shared by the HIP kernels and the C ABI of lineal_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "lineal_amd.h"

namespace lgpu {

namespace setup {
struct Accel;
}
// the AMG setup's criterion and Galerkin phases on the device (k_setup.hip)
std::unique_ptr<setup::Accel> make_device_setup(hipStream_t st);

// ---------------------------------------------------------------------------
// error reporting

void set_error(const std::string& msg);

struct Error {
  int status;
  std::string msg;
};

#define LGPU_HIP_CHECK(expr)                                                             \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::lgpu::Error{e_ == hipErrorOutOfMemory ? LGPU_OUT_OF_MEMORY : LGPU_HIP_ERROR, \
                          std::string(#expr) + ": " + hipGetErrorString(e_)};            \
  } while (0)

#define LGPU_REQUIRE(cond, msg)                                  \
  do {                                                           \
    if (!(cond)) throw ::lgpu::Error{LGPU_INVALID_ARG, (msg)};   \
  } while (0)

// ---------------------------------------------------------------------------
// device buffers

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      bytes = o.bytes;
      o.ptr = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t b) {
    release();
    if (b == 0) b = 16;
    LGPU_HIP_CHECK(hipMalloc(&ptr, b));
    bytes = b;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(ptr); }
};

// ---------------------------------------------------------------------------
// stencil description passed to kernels by value

struct StencilDesc {
  int64_t nx, ny, nz;
  double q[3];           // axis quotients, discretization/box-equidistant.hpp:62-79
  int kind;              // 0: per-cell coefficients, 1: u8 material lookup
  int zero_to_one;
  double p_start, p_end;
  double adv;            // upwind advection v * h_y * h_z along +x (0: none)
  const double* c64;     // kind 0
  const float* c32;      // kind 0, f32 copy of c64
  const uint8_t* mat;    // kind 1
  double mc[4];          // kind 1 material coefficients
  double tab64[3][16];   // kind 1 connection tables H(C[m],C[m'])*q_d in f64
  float tab32[3][16];    //                               in f32
  // kind 0: every coefficient is +0 or within [2^-100, 2^100] (f64 values) /
  // [2^-24, 2^24] (their f32 roundings), so harmonic means may use the
  // scaling-free division
  int fast_div64, fast_div32;
};

// kernel modes of the matrix-free stencil pass
enum StencilMode : int {
  SM_SPMV = 0,       // y = A x
  SM_RESID = 1,      // y = b - A x
  SM_SPMV_DOT = 2,   // y = A x,     partial += x_i y_i
  SM_RESID_NORM = 3, // y = b - A x, partial += y_i^2
  SM_JACOBI = 4,     // y = Jacobi(x -> y), b rhs, scalar omega
};

// CSR on the device
struct CsrDesc {
  int64_t n;
  const int64_t* rowptr;
  const int32_t* cols;
  const void* vals;    // f64 or f32
  int f32;
  int64_t seg_entries;  // most entries in any 64-row segment (rows 64s .. 64s+63)
  int64_t blk_entries;  // most entries in any block of kCsrBlockRows rows
  // 64-row segment sweeps: the longest run of a row's new-value entries (in
  // its own segment, on the swept side) and of its trailing old-value
  // entries, over both sweep directions
  int seg_new_max, seg_trail_max;
  // scheduled segment sweeps (csr_sor_scheduled): segment length (the whole
  // matrix when it is n), {row, level} of every position for each direction,
  // and the longest row
  int64_t sched_seg = 0;
  const int2* sched_fwd = nullptr;
  const int2* sched_bwd = nullptr;
  int sched_max_len = 0;
  int64_t nnz = 0;  // stored entries
};
// rows per block of the CSR passes
constexpr int kCsrBlockRows = 256;

// CSR with B x B block values (B = 1..4), row-major blocks
struct BcsrDesc {
  int64_t nb;            // block rows
  int block;             // B
  const int64_t* rowptr;
  const int32_t* cols;
  const void* vals;      // f64 or f32, nnz_blocks * B * B
  int f32;
};
void bcsr_spmv(const BcsrDesc& a, const void* x, void* y, hipStream_t st);

// ---------------------------------------------------------------------------
// kernel launchers (stream ordered)

// number of partial sums a stencil reduction pass of that precision writes
int64_t stencil_partials(const StencilDesc& s, bool f32);
void stencil_apply(const StencilDesc& s, int mode, bool f32, const void* x, const void* b,
                   void* y, double omega, double* partials, hipStream_t st);
void stencil_rhs(const StencilDesc& s, bool f32, void* b, hipStream_t st);

// lexicographic SOR sweep on the stencil (tile-column wavefront)
// host-side state of a stencil operator's two mailbox sets (with their
// tickets): the set the next sweep uses, and the word size (bytes) each set is
// cleared for: 0 not cleared, 4 f32 words, 8 f64 words (which covers f32)
struct SorBoxState {
  int next = 0;
  int clean[2] = {0, 0};
};
struct SorSched {
  int* state = nullptr;        // [0], [1] the tickets of the two mailbox sets
  int* err = nullptr;          // sticky wait-error word (the context's)
  const uint32_t* order = nullptr;  // ticket -> column id (forward topological order)
  int64_t ncols = 0;
  void* mailbox = nullptr;     // two sets of sor_mailbox_bytes: border-line values passed between columns
  SorBoxState* box = nullptr;  // which set is next and what is cleared
};
size_t sor_sched_ints(const StencilDesc& s);
size_t sor_mailbox_bytes(const StencilDesc& s);
void sor_sched_order(const StencilDesc& s, std::vector<uint32_t>& order);
// weights: the per-cell face weights and diagonal of stencil_sor_weights (the
// sweep reads them), or nullptr (the sweep evaluates them)
void stencil_sor(const StencilDesc& s, bool f32, const void* x_in, void* x_out, const void* b,
                 double omega, bool forward, bool ultra, const SorSched& sched, hipStream_t st,
                 const void* weights = nullptr);
// face weights and diagonal of every cell in the cached-coefficient sweep's
// order (eight values per cell slot, in the sweep's precision); 0 bytes when
// the box is too deep in z for them
size_t sor_weights_bytes(const StencilDesc& s, bool f32);
void stencil_sor_weights(const StencilDesc& s, bool f32, void* w, hipStream_t st);

// CSR kernels
int64_t csr_partials(const CsrDesc& a);
void csr_apply(const CsrDesc& a, int mode, const void* x, const void* b, void* y, double omega,
               double* partials, hipStream_t st);
// A small level swept in one segment as a program of rounds: the rows of one
// dependency level of sched_positions (no row of a round reads another's
// value, and every value a row reads holds what the sequential sweep would
// read), at most 64 per round, one wavefront lane per row.  Each round's rows
// carry their entries in column order padded to the round's chain length (a
// multiple of 8, at least the register slots) with value 0 at column n, where
// the sweep keeps a 0, so
// every lane runs the same FMA chain and padding changes no bit.  Columns are
// stored as byte offsets into the sweep's vector (column x element size);
// the diagonal stays at its column in `cols_ultra` (ultra SOR applies it
// there) and reads column n in `cols_plain`.  Host-built per direction at
// upload, in one buffer: the staging region (rows, diagonals, values, round
// headers), then the two column arrays, each section 16-byte aligned.
struct RoundSweep {
  int nrounds = 0;
  int slots = 0;  // entries per row held in registers (the shortest chain)
  int64_t entries = 0;
  const int4* rounds = nullptr;         // {first entry, first position, rows, chain length}
  const int32_t* rows = nullptr;        // row of each position (the staging region starts here)
  const void* diag = nullptr;           // each position's diagonal (1 where none), T
  const void* vals = nullptr;           // T[entries + kRoundPad]
  const uint32_t* cols_plain = nullptr;  // byte offsets, entries + kRoundPad
  const uint32_t* cols_ultra = nullptr;
};
// levels of at most this many rows may be swept by a round program
constexpr int64_t kRoundMaxRows = 512;
// padding entries (value 0 at column n) after the last row's, so that a lane
// may read a full register block of entries from any row
constexpr int64_t kRoundPad = 32;
// entries a round sweep holds in registers per row: 16, 24 or 32 in f32
// (chosen per level, RoundSweep::slots), 16 in f64; every round's chain is
// at least this long
constexpr int kRoundSlots64 = 16;
// bytes of a round program's staging region: rows (i32), diagonals (T),
// values (T; the entries and the padding), then the round headers; each
// section 16-byte aligned (es = sizeof(T)).  The column array (u32) follows
// it in LDS.
__host__ __device__ constexpr int64_t round_stage_bytes(int64_t n, int64_t entries, int nrounds, int64_t es) {
  return ((n * 4 + 15) & ~int64_t(15)) + ((n * es + 15) & ~int64_t(15)) + (entries + kRoundPad) * es +
         int64_t(nrounds) * 16;
}
// LDS bytes of a round sweep (x, b, the staging region and one column array)
size_t round_lds_bytes(int64_t n, int64_t entries, int nrounds, bool f32);
// one-segment SOR of a level by the round program of the sweep's direction,
// one workgroup
void csr_sor_rounds(const CsrDesc& a, const RoundSweep& rs, const void* x_in, void* x_out,
                    const void* b, double omega, bool ultra, hipStream_t st);

// One level of the multigrid cycle's tail, the small levels that
// amg_tail runs in one workgroup: the level operator (CSR, values of the
// cycle's precision) and vectors, its 64-row segment sweeps' staging size and
// register-list variant, and the transfers to the next level.  The last
// level of a tail is the coarsest (n, rhs and sol only).
struct TailLevel {
  int64_t n;
  const int64_t* rowptr;
  const int32_t* cols;
  const void* vals;
  int64_t cap;         // entries of the level's largest 64-row segment
  int shortlists;      // rows fit the short register lists
  const uint32_t* agg; // this level -> next (prolongation)
  const int64_t* c2f_ptr;  // next level's coarse -> fine lists (restriction into it)
  const int32_t* c2f_idx;
  void* rhs;
  void* sol;
  void* tmp;
  void* aux;
  // one-segment levels: the round programs of both directions (nrounds > 0),
  // else 64-row segments
  RoundSweep rounds_fwd, rounds_bwd;
};
// register-list sizes of the 64-row segment sweeps' short variant (new-value
// and trailing entries per row)
constexpr int kSegNewShort = 8, kSegTrailShort = 16;
// the tail's levels have at most kTailMaxRows rows; its workgroup at most
// kTailMaxWaves wavefronts
constexpr int64_t kTailMaxRows = 512;
constexpr int kTailMaxWaves = 8;
// bytes of LDS one wavefront of amg_tail needs for a level's segments
size_t tail_wave_lds_bytes(int64_t cap, bool f32);
// The V-cycle below the tail's first level (whose rhs is set): down sweeps,
// residuals and restrictions, the coarsest solve (Cholesky, or LU with U),
// prolongations (scaled by inv_relax) and backward sweeps, all in one
// workgroup of `waves` wavefronts with `wave_lds` bytes of LDS each.
void amg_tail(const TailLevel* levels, int count, int64_t coarsest_n, const double* L,
              const double* U, bool f32, bool ultra, double inv_relax, int waves, size_t wave_lds,
              hipStream_t st);
size_t tail_lds_budget();

// the longest segment a scheduled sweep keeps in one workgroup's LDS (its new
// values, f64)
constexpr int64_t kSchedMaxRows = 8192;
// thread-segment SOR over the matrix's sweep schedule (segments of
// a.sched_seg rows, each swept in order: rows grouped by dependency depth)
void csr_sor_scheduled(const CsrDesc& a, const void* x_in, void* x_out, const void* b,
                       double omega, bool forward, bool ultra, hipStream_t st);
// {row, level} positions of the scheduled sweeps of a CSR matrix: per segment
// of `seg` rows, the rows sorted by their depth in the segment's dependency
// graph for the forward (new values below the row) and the backward sweep;
// returns the longest row
int sched_positions(int64_t n, const int64_t* rowptr, const int32_t* cols, int64_t seg,
                    std::vector<int2>& fwd, std::vector<int2>& bwd);
void csr_sor(const CsrDesc& a, const void* x_in, void* x_out, const void* b, double omega,
             bool forward, bool ultra, int* sched, int* err, int* flags, int epoch, hipStream_t st);
// thread-segment SOR: segments of `seg` consecutive rows, 1 <= seg <= kMaxSegmentRows
constexpr int kMaxSegmentRows = 64;
void csr_sor_segmented(const CsrDesc& a, const void* x_in, void* x_out, const void* b,
                       double omega, bool forward, bool ultra, int seg, hipStream_t st);

// outlet flux J = sum_{x = nx-1} (2 q_0 c_i)(p_i - p_end), its adjoint
// right-hand side dJ/dp and the gradient dJ/dlog c (per-cell coefficients, f64)
int64_t outlet_flux_partials(const StencilDesc& s);
void outlet_flux(const StencilDesc& s, const double* p, double* partials, hipStream_t st);
void outlet_flux_adjoint_rhs(const StencilDesc& s, double* g, hipStream_t st);
void outlet_flux_gradient(const StencilDesc& s, const double* p, const double* lam, double* grad,
                          hipStream_t st);

// effective diffusivity through the mid plane x = nx / 2 (needs nx >= 3)
struct EffDiffTables {
  double filled[4];   // lookup: 1 for a filled material, 0 for a "none" material
  double factor[16];  // lookup: H(C[m], C[m']) at m * 4 + m'
};
int64_t eff_diff_partials(const StencilDesc& s);
// partials += filled(i) + filled(i + ny nz) over the mid plane
void eff_diff_filled(const StencilDesc& s, const EffDiffTables& t, double* partials,
                     hipStream_t st);
// partials += D(i, i + ny nz) (sol[i + ny nz] - sol[i]); with cmap (cell -> row
// or -1) sol is indexed by rows and pairs with a dropped cell are skipped
void eff_diff_flux(const StencilDesc& s, const EffDiffTables& t, const void* sol, bool f32,
                   const int64_t* cmap, double* partials, hipStream_t st);
// dst[r] = src[idx[r]], r < n
void vec_gather(const void* src, bool f32, const int64_t* idx, int64_t n, void* dst,
                hipStream_t st);

// matrix and vector files (host side, io.cpp)
namespace io {
struct SmfHeader {
  uint64_t rows = 0, cols = 0, nnz = 0;
};
SmfHeader smf_header(const char* path);
void smf_read(const char* path, SmfHeader& h, std::vector<int64_t>& rowptr,
              std::vector<int64_t>& cols, std::vector<double>& vals);
// stored_count: the header counts every stored entry (CSR), else the entries written
void smf_write(const char* path, uint64_t rows, uint64_t cols_num, const int64_t* rowptr,
               const int64_t* cols, const double* vals, bool stored_count);
uint64_t dvf_size(const char* path);
void dvf_read(const char* path, std::vector<double>& v);
void dvf_write(const char* path, const double* v, uint64_t n);
}  // namespace io

// Per-device static scratch of the allocation-free entry points: 2^22 f64
// (32 MiB) per GPU.  scratch_init() runs once per device with that device
// current (at context creation); scratch_ptr() is that device's address.
constexpr int64_t kScratchDoubles = int64_t(1) << 22;
void scratch_init(int device);
double* scratch_ptr(int device);

// vector kernels
void vec_fill(void* x, int64_t n, bool f32, double v, hipStream_t st);
void vec_fill_int(int* x, int64_t n, int v, hipStream_t st);
void vec_convert(const void* src, bool src_f32, void* dst, bool dst_f32, int64_t n, hipStream_t st);
int64_t dot_partials(int64_t n);
void vec_dot(const void* a, const void* b, bool f32, int64_t n, double* partials, hipStream_t st);
// out[slot] = sum of partials[0..np)  (deterministic order)
void reduce_partials(const double* partials, int64_t np, double* out, hipStream_t st);

// AMG transfer: coarse[c] = sum_{k in c2f[c]} fine[idx[k]]
// blk_entries: most list entries of any 256 consecutive coarse rows (staging
// size; 0 runs the unstaged kernel)
void restrict_sum(const void* fine, bool f32, const int64_t* ptr, const int32_t* idx, int64_t nc,
                  int64_t blk_entries, void* coarse, hipStream_t st);
// out[v] = x[v] + (agg[v] aggregate ? xc[agg[v]] : 0) * inv_relax
void prolong_add(const void* x, const void* xc, bool f32, const uint32_t* agg, int64_t n,
                 double inv_relax, void* out, hipStream_t st);
// dense lower-triangular Cholesky solve, n <= 4096, f64 factor; rhs/x in f64 or f32
void dense_cholesky_solve(const double* L, int64_t n, const void* b, void* x, bool f32,
                          hipStream_t st);
// dense LU solve with unit lower L and upper U (f64 factors), n <= 4096
void dense_lu_solve(const double* L, const double* U, int64_t n, const void* b, void* x, bool f32,
                    hipStream_t st);

// Krylov helper kernels operating on device scalars
// scal layout indices
enum : int {
  SC_RR = 0,      // ||r||^2
  SC_PQ = 1,      // <p, A p>
  SC_RHO0 = 2,    // rho ping
  SC_RHO1 = 3,    // rho pong
  SC_R0V = 4,     // <r0, v>
  SC_TR = 5,      // <t, r>
  SC_TT = 6,      // <t, t>
  SC_ALPHA = 7,
  SC_OMEGA = 8,
  SC_FLUX = 9,    // outlet flux
  SC_EFFDIFF = 10, // mid-plane filled count / flux sum
  SC_ERR = 11,     // the int at this slot: sticky wait-error word of the
                   // context's wavefront sweeps (copied with the scalars)
  SC_COUNT = 16,
};
// x += (num/den) p, r -= (num/den) q, partial += r^2; r32 (may be NULL)
// receives r rounded to f32
void cg_update(double* x, double* r, float* r32, const double* p, const double* q,
               const double* scal, int num, int den, int64_t n, double* partials, hipStream_t st);
// x += (num/den) p only
void axpy_scal(double* x, const double* p, const double* scal, int num, int den, double sign,
               int64_t n, hipStream_t st);
// p = z + (num/den) p, z in f64 or f32 (widened exactly)
void xpby_scal(double* p, const double* z, const double* scal, int num, int den, int64_t n,
               hipStream_t st);
void xpby_scal(double* p, const float* z, const double* scal, int num, int den, int64_t n,
               hipStream_t st);
// partials of <a, b> for an f32 b, equal to the f64 dot with b widened
void vec_dot_mixed(const double* a, const float* b, int64_t n, double* partials, hipStream_t st);
// BiCGStab: p = r + ((rho1/rho0)*(alpha/omega)) (p - omega v)
void bicg_p_update(double* p, const double* r, const double* v, const double* scal, int rho_new,
                   int rho_old, int64_t n, hipStream_t st);
// r = r - a v (a = scal[num]/scal[den] stored into scal[out]), partial += r^2
void bicg_r_update(double* r, const double* v, double* scal, int num, int den, int out,
                   int64_t n, double* partials, hipStream_t st);
// dual dot: partial0 += a.b, partial1 += a.a   (partials holds 2*np)
void vec_dot2(const double* a, const double* b, int64_t n, double* partials, hipStream_t st);
void reduce_partials2(const double* partials, int64_t np, double* out0, double* out1,
                      hipStream_t st);

}  // namespace lgpu
