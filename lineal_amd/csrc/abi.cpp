// C ABI of lineal_amd: object lifetimes, dispatch, the V-cycle and the Krylov
// drivers.  Every entry point catches every exception and returns a status.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <new>
#include <thread>

#include "internal.hpp"
#include "setup.hpp"

namespace lgpu {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace lgpu

using namespace lgpu;

#define API_BEGIN try {
#define API_END                                           \
  }                                                       \
  catch (const lgpu::Error& e) {                          \
    lgpu::set_error(e.msg);                               \
    return e.status;                                      \
  }                                                       \
  catch (const std::bad_alloc&) {                         \
    lgpu::set_error("host allocation failed");            \
    return LGPU_OUT_OF_MEMORY;                            \
  }                                                       \
  catch (const std::exception& e) {                       \
    lgpu::set_error(e.what());                            \
    return LGPU_HIP_ERROR;                                \
  }                                                       \
  catch (...) {                                           \
    lgpu::set_error("unknown error");                     \
    return LGPU_HIP_ERROR;                                \
  }                                                       \
  return LGPU_OK;

// ---------------------------------------------------------------------------
// objects

struct lgpu_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf partials;        // reduction scratch
  int64_t partial_cap = 0;
  DevBuf scal;            // SC_COUNT device scalars
  double* host_scal = nullptr;  // pinned mirror
  hipEvent_t scal_ev = nullptr;  // marks the latest copy of the scalars to the host mirror
  // opt-in timing of kernel classes (lgpu_ctx_kernel_timing): start/stop
  // event pairs recorded around every launch of a class whose bit is set
  int ktime_mask = 0;
  std::vector<hipEvent_t> kev[LGPU_KT_COUNT];
  size_t kused[LGPU_KT_COUNT] = {};
  // Krylov work vectors (f64), kept between solves and grown on demand, so
  // that a repeated solve allocates nothing
  std::vector<DevBuf> work;
  int64_t work_n = 0;
  // the sticky wait-error word of the wavefront sweeps: an int in a scalar
  // slot, so that every mirror copy of the scalars carries it
  int* err() const { return reinterpret_cast<int*>(scal.as<double>() + SC_ERR); }
  bool host_err() const { return reinterpret_cast<const int*>(host_scal + SC_ERR)[0] != 0; }
  double* work_vec(int i, int64_t n) {
    if (n > work_n) {
      work.clear();
      work_n = n;
    }
    while (int(work.size()) <= i) work.emplace_back(size_t(work_n) * sizeof(double));
    return work[size_t(i)].as<double>();
  }
  double* partials_for(int64_t need) {
    need *= 2;  // dual reductions
    if (need > partial_cap) {
      partials.alloc(size_t(need) * sizeof(double));
      partial_cap = need;
    }
    return partials.as<double>();
  }
};

namespace {

// Brackets one launch of a kernel class with HIP events on the context's
// stream while kernel timing is enabled.
struct KTimed {
  lgpu_ctx ctx;
  hipEvent_t stop = nullptr;
  KTimed(lgpu_ctx c, int cls) : ctx(c) {
    if (!((ctx->ktime_mask >> cls) & 1)) return;
    auto& v = ctx->kev[cls];
    size_t& u = ctx->kused[cls];
    if (u + 2 > v.size()) {
      hipEvent_t a = nullptr, b = nullptr;
      LGPU_HIP_CHECK(hipEventCreate(&a));
      LGPU_HIP_CHECK(hipEventCreate(&b));
      v.push_back(a);
      v.push_back(b);
    }
    LGPU_HIP_CHECK(hipEventRecord(v[u], ctx->stream));
    stop = v[u + 1];
    u += 2;
  }
  ~KTimed() {
    if (stop) (void)hipEventRecord(stop, ctx->stream);
  }
  KTimed(const KTimed&) = delete;
  KTimed& operator=(const KTimed&) = delete;
};

}  // namespace

struct lgpu_vec_s {
  lgpu_ctx ctx = nullptr;
  int64_t n = 0;
  bool f32 = false;
  void* ptr = nullptr;
  DevBuf own;
};

struct CsrDev {
  int64_t n = 0;
  bool f32 = false;
  bool symmetric = true;
  DevBuf rowptr, cols, vals, sched, flags;
  int epoch = 0;
  int64_t seg_entries = 0, blk_entries = 0;
  int seg_new_max = 0, seg_trail_max = 0;
  // the scheduled segment sweeps' positions (build_schedule)
  int64_t sched_seg = 0;
  int sched_max_len = 0;
  DevBuf sched_fwd, sched_bwd;
  int64_t nnz = 0;
  CsrDesc desc() const {
    return CsrDesc{n,           rowptr.as<int64_t>(), cols.as<int32_t>(), vals.ptr, f32 ? 1 : 0,
                   seg_entries, blk_entries,          seg_new_max,       seg_trail_max,
                   sched_seg,   sched_fwd.as<int2>(), sched_bwd.as<int2>(), sched_max_len,
                   nnz};
  }
  // positions of the scheduled sweeps with segments of `seg` rows
  void build_schedule(const setup::RowMatrix& m, int64_t seg) {
    std::vector<int2> fwd, bwd;
    sched_max_len = sched_positions(m.n, m.rowptr.data(), m.cols.data(), seg, fwd, bwd);
    sched_fwd.alloc(fwd.size() * sizeof(int2));
    sched_bwd.alloc(bwd.size() * sizeof(int2));
    LGPU_HIP_CHECK(hipMemcpy(sched_fwd.ptr, fwd.data(), fwd.size() * sizeof(int2), hipMemcpyHostToDevice));
    LGPU_HIP_CHECK(hipMemcpy(sched_bwd.ptr, bwd.data(), bwd.size() * sizeof(int2), hipMemcpyHostToDevice));
    sched_seg = seg;
  }
  // one-segment round programs of both directions ([0] forward, [1]
  // backward; nrounds = 0 when the level is not swept that way)
  RoundSweep rounds[2];
  DevBuf rounds_buf[2];
  // Builds the round programs of a level of at most kRoundMaxRows rows swept
  // in one segment (the dependency levels of sched_positions with the whole
  // matrix as the segment).  Returns false, building nothing, when a program
  // would exceed the LDS budget.
  bool build_rounds(const setup::RowMatrix& m, size_t lds_budget) {
    rounds[0] = rounds[1] = RoundSweep{};
    if (m.n <= 0 || m.n > kRoundMaxRows) return false;
    std::vector<int2> pos[2];
    sched_positions(m.n, m.rowptr.data(), m.cols.data(), m.n, pos[0], pos[1]);
    const int64_t es = f32 ? 4 : 8;
    // register slots per row: in f32 the smallest of 16, 24, 32 that holds
    // nine rows in ten of the level (longer rows read the rest from LDS)
    int slots = kRoundSlots64;
    if (f32) {
      std::vector<int64_t> len(static_cast<size_t>(n));
      for (int64_t i = 0; i < n; ++i) len[size_t(i)] = m.rowptr[size_t(i + 1)] - m.rowptr[size_t(i)];
      std::sort(len.begin(), len.end());
      const int64_t p90 = len[size_t(std::min<int64_t>(n - 1, (9 * n) / 10))];
      slots = p90 <= 16 ? 16 : (p90 <= 24 ? 24 : 32);
    }
    auto r16 = [](int64_t v) { return (v + 15) & ~int64_t(15); };
    for (int dir = 0; dir < 2; ++dir) {
      const std::vector<int2>& ps = pos[dir];
      // rounds: one dependency level at a time, at most 64 rows each, the
      // chain length the round's longest row rounded up to a multiple of 8,
      // at least the sweep's register slots
      std::vector<int4> rs;
      int64_t entries = 0;
      for (int64_t p = 0; p < n;) {
        int64_t q = p;
        int longest = 0;
        while (q < n && ps[size_t(q)].y == ps[size_t(p)].y && q - p < 64) {
          const int64_t i = ps[size_t(q)].x;
          longest = std::max(longest, int(m.rowptr[size_t(i + 1)] - m.rowptr[size_t(i)]));
          ++q;
        }
        const int chain = std::max(slots, (longest + 7) / 8 * 8);
        rs.push_back(int4{int(entries), int(p), int(q - p), chain});
        entries += (q - p) * chain;
        p = q;
      }
      if (round_lds_bytes(n, entries, int(rs.size()), f32) > lds_budget) {
        rounds[0] = rounds[1] = RoundSweep{};
        return false;
      }
      // the staging region the sweep copies whole (rows, diagonals, values,
      // round headers), then the two column arrays (byte offsets)
      const int64_t ep = entries + kRoundPad;
      const int64_t o_rows = 0;
      const int64_t o_diag = o_rows + r16(n * 4);
      const int64_t o_vals = o_diag + r16(n * es);
      const int64_t o_hdr = o_vals + ep * es;
      const int64_t o_cp = o_hdr + int64_t(rs.size()) * 16;
      const int64_t o_cu = o_cp + 4 * ep;
      const int64_t total = o_cu + 4 * ep;
      if (o_cp != round_stage_bytes(n, entries, int(rs.size()), es))
        throw Error{LGPU_HIP_ERROR, "round program layout mismatch"};
      std::vector<unsigned char> h(size_t(total), 0);
      auto put = [&](int64_t off, double v) {
        if (f32) {
          const float f = float(v);
          std::memcpy(h.data() + off, &f, 4);
        } else {
          std::memcpy(h.data() + off, &v, 8);
        }
      };
      auto put32 = [&](int64_t off, int64_t col) {
        const uint32_t b = uint32_t(col * es);
        std::memcpy(h.data() + off, &b, 4);
      };
      std::memcpy(h.data() + o_hdr, rs.data(), rs.size() * sizeof(int4));
      for (int64_t e = entries; e < ep; ++e) {  // the padding: value 0 at column n
        put32(o_cp + e * 4, n);
        put32(o_cu + e * 4, n);
      }
      for (const int4& r : rs) {
        for (int j = 0; j < r.z; ++j) {
          const int64_t p = r.y + j;
          const int32_t i = ps[size_t(p)].x;
          std::memcpy(h.data() + o_rows + p * 4, &i, 4);
          double d = 1.0;  // no stored diagonal: the sweep divides by 1
          const int64_t e0 = int64_t(r.x) + int64_t(j) * r.w;
          const int64_t k0 = m.rowptr[size_t(i)], k1 = m.rowptr[size_t(i + 1)];
          for (int64_t u = 0; u < r.w; ++u) {
            const int64_t e = e0 + u;
            if (k0 + u < k1) {
              const int32_t c = m.cols[size_t(k0 + u)];
              const double v = m.vals[size_t(k0 + u)];
              if (c == i) d = v;
              put(o_vals + e * es, v);
              put32(o_cp + e * 4, c == i ? n : c);
              put32(o_cu + e * 4, c);
            } else {  // padding: value 0 at column n
              put(o_vals + e * es, 0.0);
              put32(o_cp + e * 4, n);
              put32(o_cu + e * 4, n);
            }
          }
          put(o_diag + p * es, d);
        }
      }
      rounds_buf[dir].alloc(size_t(total));
      LGPU_HIP_CHECK(hipMemcpy(rounds_buf[dir].ptr, h.data(), size_t(total), hipMemcpyHostToDevice));
      const auto b0 = static_cast<unsigned char*>(rounds_buf[dir].ptr);
      RoundSweep& g = rounds[dir];
      g.nrounds = int(rs.size());
      g.slots = slots;
      g.entries = entries;
      g.rounds = reinterpret_cast<const int4*>(b0 + o_hdr);
      g.rows = reinterpret_cast<const int32_t*>(b0 + o_rows);
      g.diag = b0 + o_diag;
      g.vals = b0 + o_vals;
      g.cols_plain = reinterpret_cast<const uint32_t*>(b0 + o_cp);
      g.cols_ultra = reinterpret_cast<const uint32_t*>(b0 + o_cu);
    }
    return true;
  }
  void upload(const setup::RowMatrix& m, bool as_f32) {
    n = m.n;
    auto most = [&](int64_t rows) {
      int64_t e = 0;
      for (int64_t s0 = 0; s0 < n; s0 += rows)
        e = std::max(e, m.rowptr[size_t(std::min(s0 + rows, n))] - m.rowptr[size_t(s0)]);
      return e;
    };
    seg_entries = most(kMaxSegmentRows);
    blk_entries = most(kCsrBlockRows);
    // per row of a 64-row segment [s0, s1): forward new [s0, i), trailing
    // (i, n); backward new (i, s1), trailing [s1, n)
    seg_new_max = seg_trail_max = 0;
    for (int64_t i = 0; i < n; ++i) {
      const int64_t s0 = i / kMaxSegmentRows * kMaxSegmentRows;
      const int64_t s1 = std::min(s0 + kMaxSegmentRows, n);
      int fn = 0, ft = 0, bn = 0, bt = 0;
      for (int64_t k = m.rowptr[size_t(i)]; k < m.rowptr[size_t(i + 1)]; ++k) {
        const int64_t c = m.cols[size_t(k)];
        fn += c >= s0 && c < i;
        ft += c > i;
        bn += c > i && c < s1;
        bt += c >= s1;
      }
      seg_new_max = std::max({seg_new_max, fn, bn});
      seg_trail_max = std::max({seg_trail_max, ft, bt});
    }
    f32 = as_f32;
    symmetric = m.symmetric;
    rowptr.alloc(size_t(n + 1) * sizeof(int64_t));
    LGPU_HIP_CHECK(hipMemcpy(rowptr.ptr, m.rowptr.data(), size_t(n + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    const size_t nnz = m.cols.size();
    this->nnz = int64_t(nnz);
    cols.alloc(nnz * sizeof(int32_t));
    LGPU_HIP_CHECK(hipMemcpy(cols.ptr, m.cols.data(), nnz * sizeof(int32_t), hipMemcpyHostToDevice));
    if (as_f32) {
      std::vector<float> v(m.vals.begin(), m.vals.end());
      vals.alloc(nnz * sizeof(float));
      LGPU_HIP_CHECK(hipMemcpy(vals.ptr, v.data(), nnz * sizeof(float), hipMemcpyHostToDevice));
    } else {
      vals.alloc(nnz * sizeof(double));
      LGPU_HIP_CHECK(hipMemcpy(vals.ptr, m.vals.data(), nnz * sizeof(double), hipMemcpyHostToDevice));
    }
    sched.alloc(4 * sizeof(int));
    flags.alloc(size_t(n) * sizeof(int));
    LGPU_HIP_CHECK(hipMemset(sched.ptr, 0, 4 * sizeof(int)));
    LGPU_HIP_CHECK(hipMemset(flags.ptr, 0, size_t(n) * sizeof(int)));
    epoch = 0;
  }
};

struct lgpu_op_s {
  lgpu_ctx ctx = nullptr;
  bool is_csr = false;
  // stencil
  lgpu_box box{};
  StencilDesc desc{};
  DevBuf c64, c32, mat;
  std::vector<double> host_coeff;
  std::vector<uint8_t> host_mat;
  DevBuf sor_state, sor_order, sor_mailbox;
  int64_t sor_cols = 0;
  SorBoxState sor_box;
  // csr
  CsrDev csr;
  setup::RowMatrix host_csr;
  // compressed system from a stencil: row of every cell or -1 (i64 per cell)
  DevBuf cmap;
  int64_t cmap_cells = 0;
  // intermediate vector of SSOR (the forward sweep's output), f64-sized
  DevBuf ssor_aux;
  // per-cell face weights and diagonal read by the stencil SOR sweeps, per
  // precision, built on first use (sor_cache = false: the sweeps evaluate
  // them on the fly and no cache is kept)
  DevBuf sor_w64, sor_w32;
  bool sor_cache = true;

  int64_t rows() const { return is_csr ? csr.n : desc.nx * desc.ny * desc.nz; }
  void ensure_f32() {
    if (is_csr || desc.kind != 0 || c32.ptr) return;
    const int64_t n = rows();
    c32.alloc(size_t(n) * sizeof(float));
    vec_convert(c64.ptr, false, c32.ptr, true, n, ctx->stream);
    desc.c32 = c32.as<float>();
  }
  // the SOR weights of a stencil operator in that precision: built here when
  // `build` (and the cache is on), else only when already there
  const void* sor_weights(bool f32, bool build) {
    if (is_csr || !sor_cache) return nullptr;
    DevBuf& w = f32 ? sor_w32 : sor_w64;
    if (!w.ptr && build) {
      const size_t bytes = sor_weights_bytes(desc, f32);
      if (bytes == 0) return nullptr;
      if (f32) ensure_f32();
      w.alloc(bytes);
      stencil_sor_weights(desc, f32, w.ptr, ctx->stream);
    }
    return w.ptr;
  }
};

struct AmgLevel {
  CsrDev a;            // level operator (levels >= 1)
  int64_t seg = 0;     // rows per smoother segment (levels >= 1; 0: one segment)
  DevBuf agg;          // this level -> next level map
  DevBuf c2f_ptr, c2f_idx;
  int64_t c2f_blk = 0;  // most list entries of any 256 consecutive coarse rows
  DevBuf rhs, sol, tmp, aux;
  int64_t n = 0;
};

struct lgpu_amg_s {
  lgpu_ctx ctx = nullptr;
  lgpu_op finest = nullptr;
  lgpu_amg_params params{};
  setup::Hierarchy host;
  std::vector<AmgLevel> lv;
  DevBuf chol;   // coarsest lower factor (Cholesky, or the unit lower LU factor)
  DevBuf upper;  // coarsest upper LU factor (nonsymmetric hierarchies)
  double setup_seconds = 0.0;
  bool f32 = false;
  DevBuf r32, z32;  // finest f32 copies for the mixed-precision cycle
  // the V-cycle's tail (levels tail_first .. coarsest in one workgroup,
  // amg_tail): its TailLevel array on the device, 0 levels when none
  DevBuf tail;
  int tail_first = 0, tail_count = 0, tail_waves = 0;
  size_t tail_wave_lds = 0;
  // an SSOR preconditioner instead of a hierarchy (lgpu_ssor_precon_create):
  // ssor_iters applications of a forward and a backward sweep from zero
  int ssor_iters = 0;
  double ssor_omega = 1.0;
  bool ssor_ultra = false;
};

namespace {

size_t esize(bool f32) { return f32 ? sizeof(float) : sizeof(double); }

void check_vec(lgpu_vec v, int64_t n, const char* what) {
  LGPU_REQUIRE(v != nullptr, std::string(what) + " is NULL");
  LGPU_REQUIRE(v->n == n, std::string(what) + " has the wrong size");
}

double host_hmean(double a, double b) {
  const double p = a * b;
  return p != 0.0 ? (p + p) / (a + b) : 0.0;
}
float host_hmean(float a, float b) {
  const float p = a * b;
  return p != 0.0f ? (p + p) / (a + b) : 0.0f;
}

void fill_desc(lgpu_op op, const lgpu_box* box) {
  StencilDesc& d = op->desc;
  d.nx = box->n[0];
  d.ny = box->n[1];
  d.nz = box->n[2];
  // discretization/box-equidistant.hpp:62-79
  double h[3];
  for (int k = 0; k < 3; ++k) h[k] = box->len[k] / double(box->n[k]);
  for (int k = 0; k < 3; ++k) {
    double area = 1.0;
    for (int e = 0; e < 3; ++e)
      if (e != k) area *= h[e];
    d.q[k] = area / h[k];
  }
  d.zero_to_one = box->zero_to_one;
  d.p_start = box->p_start;
  d.p_end = box->p_end;
  d.adv = box->velocity_x == 0.0 ? 0.0 : box->velocity_x * h[1] * h[2];
}

void init_sor_sched(lgpu_op op) {
  std::vector<uint32_t> order;
  sor_sched_order(op->desc, order);
  op->sor_cols = int64_t(order.size());
  op->sor_order.alloc(order.size() * sizeof(uint32_t));
  LGPU_HIP_CHECK(hipMemcpy(op->sor_order.ptr, order.data(), order.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  const size_t ints = sor_sched_ints(op->desc);
  op->sor_state.alloc(ints * sizeof(int));
  LGPU_HIP_CHECK(hipMemset(op->sor_state.ptr, 0, ints * sizeof(int)));
  op->sor_mailbox.alloc(2 * sor_mailbox_bytes(op->desc));
  op->sor_box = SorBoxState{};
}

SorSched sor_sched_of(lgpu_op op) {
  SorSched s;
  s.state = op->sor_state.as<int>();
  s.err = op->ctx->err();
  s.order = op->sor_order.as<uint32_t>();
  s.ncols = op->sor_cols;
  s.mailbox = op->sor_mailbox.ptr;
  s.box = &op->sor_box;
  return s;
}

// Host evaluation of the stencil rows in f64 in iterate order (-x,-y,-z,+z,+y,+x,
// diagonal last), tensor/dynamic/stencil/matrix/algorithm.hpp:91-227, plus the
// upwind advection terms (the -x coupling and the diagonal); the same
// operations as the device's row evaluation.
void stencil_host_rows(const lgpu_op op, setup::RowMatrix& m) {
  const StencilDesc& d = op->desc;
  const int64_t nx = d.nx, ny = d.ny, nz = d.nz, n = nx * ny * nz, nyz = ny * nz;
  m = setup::RowMatrix{};
  m.n = n;
  m.f32 = false;
  m.symmetric = d.adv == 0.0;
  m.rowptr.resize(size_t(n + 1));
  m.rowptr[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t x = i / nyz, y = (i / nz) % ny, z = i % nz;
    const int cnt = 1 + (x > 0) + (x + 1 < nx) + (y > 0) + (y + 1 < ny) + (z > 0) + (z + 1 < nz);
    m.rowptr[i + 1] = m.rowptr[i] + cnt;
  }
  m.cols.resize(size_t(m.rowptr[n]));
  m.vals.resize(size_t(m.rowptr[n]));
  auto coeff = [&](int64_t i) {
    return d.kind == 0 ? op->host_coeff[size_t(i)] : d.mc[op->host_mat[size_t(i)]];
  };
  auto conn = [&](int64_t i, int64_t j, int ax) {
    if (d.kind == 0) return host_hmean(op->host_coeff[size_t(i)], op->host_coeff[size_t(j)]) * d.q[ax];
    return d.tab64[ax][op->host_mat[size_t(i)] * 4 + op->host_mat[size_t(j)]];
  };
  auto work = [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const int64_t x = i / nyz, y = (i / nz) % ny, z = i % nz;
      const int64_t pos[3] = {x, y, z}, off[3] = {nyz, nz, 1};
      const double ci = coeff(i);
      double diag = 0.0;
      if (x == 0) diag += (2.0 * d.q[0]) * ci;
      if (x + 1 == nx) diag += (2.0 * d.q[0]) * ci;
      int64_t k = m.rowptr[i];
      for (int ax = 0; ax < 3; ++ax)
        if (pos[ax] > 0) {
          const double w = conn(i, i - off[ax], ax);
          m.cols[size_t(k)] = int32_t(i - off[ax]);
          m.vals[size_t(k++)] = ax == 0 ? -(w + d.adv) : -w;
          diag += w;
        }
      for (int ax = 2; ax >= 0; --ax) {
        const int64_t ext = ax == 0 ? nx : (ax == 1 ? ny : nz);
        if (pos[ax] + 1 < ext) {
          const double w = conn(i, i + off[ax], ax);
          m.cols[size_t(k)] = int32_t(i + off[ax]);
          m.vals[size_t(k++)] = -w;
          diag += w;
        }
      }
      diag += d.adv;
      if (d.zero_to_one && diag == 0.0) diag = 1.0;
      m.cols[size_t(k)] = int32_t(i);
      m.vals[size_t(k)] = diag;
    }
  };
  const int nt = int(std::max(1u, std::thread::hardware_concurrency()));
  if (n < 65536 || nt == 1) {
    work(0, n);
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(work, n * t / nt, n * (t + 1) / nt);
    for (auto& th : pool) th.join();
  }
}

// dispatch of one matrix pass
void op_apply(lgpu_op op, int mode, bool f32, const void* x, const void* b, void* y, double omega,
              double* partials) {
  hipStream_t st = op->ctx->stream;
  if (op->is_csr) {
    LGPU_REQUIRE(op->csr.f32 == f32, "CSR value type and vector type differ");
    KTimed kt(op->ctx, LGPU_KT_CSR_PASS);
    csr_apply(op->csr.desc(), mode, x, b, y, omega, partials, st);
  } else {
    if (f32) op->ensure_f32();
    KTimed kt(op->ctx, LGPU_KT_STENCIL_PASS);
    stencil_apply(op->desc, mode, f32, x, b, y, omega, partials, st);
  }
}

// partial sums written by an f64 reduction pass of the operator
int64_t op_partials(lgpu_op op) {
  return op->is_csr ? csr_partials(op->csr.desc()) : stencil_partials(op->desc, false);
}

void csr_sweep(lgpu_ctx ctx, CsrDev& a, const void* x_in, void* x_out, const void* b, double omega,
               bool fwd, bool ultra) {
  hipStream_t st = ctx->stream;
  if (++a.epoch == 0x7fffffff) {
    vec_fill_int(a.flags.as<int>(), a.n, 0, st);
    a.epoch = 1;
  }
  csr_sor(a.desc(), x_in, x_out, b, omega, fwd, ultra, a.sched.as<int>(), ctx->err(),
          a.flags.as<int>(), a.epoch, st);
}

void op_sor(lgpu_op op, bool f32, const void* x_in, void* x_out, const void* b, double omega,
            bool fwd, bool ultra) {
  hipStream_t st = op->ctx->stream;
  if (op->is_csr) {
    LGPU_REQUIRE(op->csr.f32 == f32, "CSR value type and vector type differ");
    KTimed kt(op->ctx, LGPU_KT_CSR_SOR);
    csr_sweep(op->ctx, op->csr, x_in, x_out, b, omega, fwd, ultra);
  } else {
    if (f32) op->ensure_f32();
    const SorSched s = sor_sched_of(op);
    const void* w = op->sor_weights(f32, true);
    KTimed kt(op->ctx, fwd ? LGPU_KT_STENCIL_SOR_FWD : LGPU_KT_STENCIL_SOR_BWD);
    stencil_sor(op->desc, f32, x_in, x_out, b, omega, fwd, ultra, s, st, w);
  }
}

// sticky timeout words of the wavefront kernels
int read_int(hipStream_t st, const int* dev) {
  int v = 0;
  LGPU_HIP_CHECK(hipMemcpyAsync(&v, dev, sizeof(int), hipMemcpyDeviceToHost, st));
  LGPU_HIP_CHECK(hipStreamSynchronize(st));
  return v;
}

// Reads the context's sticky wait-error word after the queued work and
// clears it: true if a bounded wait of a wavefront sweep expired since the
// last clear (the sweeps after it gave up their waits at once and their
// output is undefined).
bool take_wait_error(lgpu_ctx ctx) {
  if (read_int(ctx->stream, ctx->err()) == 0) return false;
  vec_fill_int(ctx->err(), 1, 0, ctx->stream);
  LGPU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return true;
}

int wait_timeout() {
  set_error("a wavefront kernel's bounded wait expired");
  return LGPU_SYNC_TIMEOUT;
}

double read_scalar(lgpu_ctx ctx, int slot) {
  LGPU_HIP_CHECK(hipMemcpyAsync(ctx->host_scal, ctx->scal.as<double>() + slot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LGPU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return ctx->host_scal[0];
}

// r = b - A x (f64), ||r||^2 into the device scalar *dst
void residual_norm(lgpu_op op, const double* b, const double* x, double* r, double* dst) {
  lgpu_ctx ctx = op->ctx;
  const int64_t np = op_partials(op);
  double* part = ctx->partials_for(np);
  op_apply(op, SM_RESID_NORM, false, x, b, r, 0.0, part);
  reduce_partials(part, np, dst, ctx->stream);
}

void dot_into(lgpu_ctx ctx, const void* a, const void* b, bool f32, int64_t n, double* dst) {
  const int64_t np = dot_partials(n);
  double* part = ctx->partials_for(np);
  vec_dot(a, b, f32, n, part, ctx->stream);
  reduce_partials(part, np, dst, ctx->stream);
}

// smoother of a coarse level: thread segments of the level's segment length
// (level_segment_rows), 0 meaning one lexicographic segment
void coarse_sweep(lgpu_amg amg, CsrDev& a, int64_t seg, const void* x_in, void* x_out,
                  const void* b, double omega, bool fwd, bool ultra) {
  KTimed kt(amg->ctx, LGPU_KT_CSR_SOR);
  if (seg == 0 && a.rounds[fwd ? 0 : 1].nrounds > 0)
    csr_sor_rounds(a.desc(), a.rounds[fwd ? 0 : 1], x_in, x_out, b, omega, ultra, amg->ctx->stream);
  else if (a.sched_seg > 0)
    csr_sor_scheduled(a.desc(), x_in, x_out, b, omega, fwd, ultra, amg->ctx->stream);
  else if (seg > 0)
    csr_sor_segmented(a.desc(), x_in, x_out, b, omega, fwd, ultra, int(seg), amg->ctx->stream);
  else
    csr_sweep(amg->ctx, a, x_in, x_out, b, omega, fwd, ultra);
}

// One multigrid cycle in the hierarchy's precision from zero: z = M r
// (MultiGridSolver::Instance::apply_impl, multigrid/solver.hpp:56-176).  The
// cycle is the reference's stack of DefaultCycleInfo {iter, iter_num}
// (multigrid/cycle.hpp:24-79): a level runs its down step (forward SOR from
// zero, residual, restriction) when iter == iter_num and its up step
// (prolong-add scaled by 1/relax, backward SOR) when iter == 0; otherwise it
// counts down and pushes its coarser level again (REGULAR: {iter_num},
// DECREASING: {iter}).  DefaultCycle{REGULAR, 1} is the V-cycle.
// `fine` smooths and forms the residual on the finest level: the hierarchy's
// own finest operator, or a Krylov operator of the same size (the upwind
// advection operator over the hierarchy of its diffusion part, so that the
// forward sweep follows the flow).
void vcycle(lgpu_amg amg, const void* r, void* z, lgpu_op fine, const void* x0 = nullptr) {
  const bool f32 = amg->f32;
  hipStream_t st = amg->ctx->stream;
  const int L = int(amg->lv.size());
  const bool ultra = amg->params.ultra != 0;
  // the smoothers are SorSolver{1} (test/aux/default.hpp:156-186)
  const double omega = 1.0;
  const double inv_relax = 1.0 / amg->params.relax;
  auto xs = [&](int l) { return l == 0 ? z : amg->lv[size_t(l)].sol.ptr; };
  auto bs = [&](int l) { return l == 0 ? r : amg->lv[size_t(l)].rhs.ptr; };
  auto down = [&](int l) {
    AmgLevel& cur = amg->lv[size_t(l)];
    if (l == 0) {
      op_sor(fine, f32, x0, xs(0), bs(0), omega, true, ultra);
      op_apply(fine, SM_RESID, f32, xs(0), bs(0), cur.tmp.ptr, 0.0, nullptr);
    } else {
      coarse_sweep(amg, cur.a, cur.seg, nullptr, xs(l), bs(l), omega, true, ultra);
      KTimed kt(amg->ctx, LGPU_KT_CSR_PASS);
      csr_apply(cur.a.desc(), SM_RESID, xs(l), bs(l), cur.tmp.ptr, 0.0, nullptr, st);
    }
    const int64_t nc = amg->lv[size_t(l + 1)].n;
    restrict_sum(cur.tmp.ptr, f32, cur.c2f_ptr.as<int64_t>(), cur.c2f_idx.as<int32_t>(), nc,
                 cur.c2f_blk, amg->lv[size_t(l + 1)].rhs.ptr, st);
  };
  auto up = [&](int l) {
    AmgLevel& cur = amg->lv[size_t(l)];
    prolong_add(xs(l), amg->lv[size_t(l + 1)].sol.ptr, f32, cur.agg.as<uint32_t>(), cur.n,
                inv_relax, cur.aux.ptr, st);
    if (l == 0) op_sor(fine, f32, cur.aux.ptr, xs(0), bs(0), omega, false, ultra);
    else coarse_sweep(amg, cur.a, cur.seg, cur.aux.ptr, xs(l), bs(l), omega, false, ultra);
  };
  auto coarsest = [&] {
    AmgLevel& c = amg->lv[size_t(L - 1)];
    if (amg->host.lu)
      dense_lu_solve(amg->chol.as<double>(), amg->upper.as<double>(), c.n, c.rhs.ptr, c.sol.ptr, f32, st);
    else
      dense_cholesky_solve(amg->chol.as<double>(), c.n, c.rhs.ptr, c.sol.ptr, f32, st);
  };
  const bool regular = amg->params.cycle_kind == LGPU_CYCLE_REGULAR;
  if (regular && amg->params.cycle_iterations == 1 && amg->tail_count >= 2) {
    // the V-cycle with its small levels in one workgroup
    const int t0 = amg->tail_first;
    for (int l = 0; l < t0; ++l) down(l);
    {
      KTimed kt(amg->ctx, LGPU_KT_CSR_SOR);
      amg_tail(amg->tail.as<TailLevel>(), amg->tail_count, amg->lv[size_t(L - 1)].n, amg->chol.as<double>(),
               amg->host.lu ? amg->upper.as<double>() : nullptr, f32, ultra, inv_relax, amg->tail_waves,
               amg->tail_wave_lds, st);
    }
    for (int l = t0 - 1; l >= 0; --l) up(l);
    return;
  }
  std::vector<int> iter(size_t(L), 0), num(size_t(L), 0);
  int top = 0;
  iter[0] = num[0] = amg->params.cycle_iterations;
  while (top >= 0) {
    const int l = top;
    if (l == L - 1) {
      coarsest();
      --top;
      continue;
    }
    if (iter[size_t(l)] == num[size_t(l)]) down(l);
    if (iter[size_t(l)] == 0) {
      up(l);
      --top;
      continue;
    }
    --iter[size_t(l)];
    ++top;
    const int next = top == L - 1 ? 0 : (regular ? num[size_t(l)] : iter[size_t(l)]);
    iter[size_t(top)] = num[size_t(top)] = next;
  }
}

// z (f64) = M r (f64)
// SsorBroom::sweep (gauss-seidel-variant/solver.hpp:45-88): a forward sweep
// x_in -> aux, then a backward sweep aux -> x_out; x_out may be x_in
void ssor_sweep(lgpu_op op, bool f32, const void* x_in, void* x_out, const void* b, double omega,
                bool ultra) {
  if (!op->ssor_aux.ptr) op->ssor_aux.alloc(size_t(op->rows()) * sizeof(double));
  op_sor(op, f32, x_in, op->ssor_aux.ptr, b, omega, true, ultra);
  op_sor(op, f32, op->ssor_aux.ptr, x_out, b, omega, false, ultra);
}

// The same two sweeps with the intermediate vector in `aux` (device memory
// of at least rows() elements), and no kernel-class timing (whose events are
// created on demand): nothing here allocates.
void ssor_sweep_into(lgpu_op op, bool f32, const void* x_in, void* x_out, const void* b,
                     double omega, bool ultra, void* aux) {
  hipStream_t st = op->ctx->stream;
  for (int fwd = 1; fwd >= 0; --fwd) {
    const void* in = fwd ? x_in : aux;
    void* out = fwd ? aux : x_out;
    if (op->is_csr) {
      csr_sweep(op->ctx, op->csr, in, out, b, omega, fwd != 0, ultra);
    } else {
      const SorSched s = sor_sched_of(op);
      // the cached weights when they exist (building them would allocate)
      stencil_sor(op->desc, f32, in, out, b, omega, fwd != 0, ultra, s, st,
                  op->sor_weights(f32, false));
    }
  }
}

void precondition(lgpu_amg amg, const double* r, double* z, lgpu_op fine) {
  if (amg->ssor_iters > 0) {
    // DoubleSweepImpl with a fixed iteration count (broom.hpp:159-171), from zero
    ssor_sweep(fine, false, nullptr, z, r, amg->ssor_omega, amg->ssor_ultra);
    for (int k = 1; k < amg->ssor_iters; ++k)
      ssor_sweep(fine, false, z, z, r, amg->ssor_omega, amg->ssor_ultra);
    return;
  }
  if (!amg->f32) {
    vcycle(amg, r, z, fine);
    return;
  }
  const int64_t n = amg->lv[0].n;
  hipStream_t st = amg->ctx->stream;
  vec_convert(r, false, amg->r32.ptr, true, n, st);
  vcycle(amg, amg->r32.ptr, amg->z32.ptr, fine);
  vec_convert(amg->z32.ptr, true, z, false, n, st);
}

}  // namespace

// ---------------------------------------------------------------------------
// errors, context

extern "C" const char* lgpu_status_string(int status) {
  switch (status) {
    case LGPU_OK: return "ok";
    case LGPU_INVALID_ARG: return "invalid argument";
    case LGPU_OUT_OF_MEMORY: return "out of memory";
    case LGPU_HIP_ERROR: return "HIP error";
    case LGPU_BREAKDOWN: return "Krylov breakdown";
    case LGPU_NONFINITE: return "non-finite residual";
    case LGPU_NOT_CONVERGED: return "not converged";
    case LGPU_NOT_SUPPORTED: return "not supported";
    case LGPU_SYNC_TIMEOUT: return "wavefront wait timed out";
    default: return "unknown status";
  }
}

extern "C" const char* lgpu_last_error(void) { return lgpu::g_last_error.c_str(); }

extern "C" int lgpu_ctx_create(int device, lgpu_ctx* out) {
  API_BEGIN
  LGPU_REQUIRE(out != nullptr, "out is NULL");
  auto ctx = std::make_unique<lgpu_ctx_s>();
  ctx->device = device;
  LGPU_HIP_CHECK(hipSetDevice(device));
  LGPU_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  LGPU_HIP_CHECK(hipEventCreate(&ctx->ev0));
  LGPU_HIP_CHECK(hipEventCreate(&ctx->ev1));
  ctx->scal.alloc(SC_COUNT * sizeof(double));
  LGPU_HIP_CHECK(hipMemset(ctx->scal.ptr, 0, SC_COUNT * sizeof(double)));
  LGPU_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&ctx->host_scal), SC_COUNT * sizeof(double), hipHostMallocDefault));
  LGPU_HIP_CHECK(hipEventCreateWithFlags(&ctx->scal_ev, hipEventDisableTiming));
  ctx->partials_for(4096);
  scratch_init(device);
  *out = ctx.release();
  API_END
}

extern "C" int lgpu_ctx_destroy(lgpu_ctx ctx) {
  API_BEGIN
  if (!ctx) return LGPU_OK;
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->host_scal) (void)hipHostFree(ctx->host_scal);
  for (auto& v : ctx->kev)
    for (hipEvent_t e : v) (void)hipEventDestroy(e);
  (void)hipEventDestroy(ctx->ev0);
  (void)hipEventDestroy(ctx->ev1);
  if (ctx->scal_ev) (void)hipEventDestroy(ctx->scal_ev);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  API_END
}

extern "C" int lgpu_ctx_synchronize(lgpu_ctx ctx) {
  API_BEGIN
  LGPU_REQUIRE(ctx, "ctx is NULL");
  LGPU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  API_END
}

extern "C" int lgpu_timer_start(lgpu_ctx ctx) {
  API_BEGIN
  LGPU_REQUIRE(ctx, "ctx is NULL");
  LGPU_HIP_CHECK(hipEventRecord(ctx->ev0, ctx->stream));
  API_END
}

extern "C" int lgpu_timer_stop(lgpu_ctx ctx, double* ms) {
  API_BEGIN
  LGPU_REQUIRE(ctx && ms, "NULL argument");
  LGPU_HIP_CHECK(hipEventRecord(ctx->ev1, ctx->stream));
  LGPU_HIP_CHECK(hipEventSynchronize(ctx->ev1));
  float f = 0;
  LGPU_HIP_CHECK(hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
  *ms = f;
  API_END
}

extern "C" int lgpu_ctx_set_wait_error(lgpu_ctx ctx) {
  API_BEGIN
  LGPU_REQUIRE(ctx, "ctx is NULL");
  vec_fill_int(ctx->err(), 1, 1, ctx->stream);
  API_END
}

extern "C" int lgpu_ctx_kernel_timing(lgpu_ctx ctx, int class_mask) {
  API_BEGIN
  LGPU_REQUIRE(ctx, "ctx is NULL");
  LGPU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  for (size_t& u : ctx->kused) u = 0;
  LGPU_REQUIRE((class_mask & ~LGPU_KT_ALL) == 0, "unknown kernel class in the mask");
  ctx->ktime_mask = class_mask;
  API_END
}

extern "C" int lgpu_ctx_kernel_time(lgpu_ctx ctx, int kernel_class, double* total_ms,
                                    int64_t* launches) {
  API_BEGIN
  LGPU_REQUIRE(ctx && total_ms && launches, "NULL argument");
  LGPU_REQUIRE(kernel_class >= 0 && kernel_class < LGPU_KT_COUNT, "unknown kernel class");
  LGPU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  const auto& v = ctx->kev[kernel_class];
  const size_t u = ctx->kused[kernel_class];
  double t = 0;
  for (size_t k = 0; k < u; k += 2) {
    float f = 0;
    LGPU_HIP_CHECK(hipEventElapsedTime(&f, v[k], v[k + 1]));
    t += f;
  }
  *total_ms = t;
  *launches = int64_t(u / 2);
  API_END
}

// ---------------------------------------------------------------------------
// vectors

extern "C" int lgpu_vec_create(lgpu_ctx ctx, int64_t n, int dtype, lgpu_vec* out) {
  API_BEGIN
  LGPU_REQUIRE(ctx && out && n >= 0, "invalid vector arguments");
  LGPU_REQUIRE(dtype == LGPU_F64 || dtype == LGPU_F32, "unknown dtype");
  auto v = std::make_unique<lgpu_vec_s>();
  v->ctx = ctx;
  v->n = n;
  v->f32 = dtype == LGPU_F32;
  v->own.alloc(size_t(n) * esize(v->f32));
  v->ptr = v->own.ptr;
  *out = v.release();
  API_END
}

extern "C" int lgpu_vec_wrap(lgpu_ctx ctx, void* device_ptr, int64_t n, int dtype, lgpu_vec* out) {
  API_BEGIN
  LGPU_REQUIRE(ctx && out && device_ptr && n >= 0, "invalid vector arguments");
  LGPU_REQUIRE(dtype == LGPU_F64 || dtype == LGPU_F32, "unknown dtype");
  auto v = std::make_unique<lgpu_vec_s>();
  v->ctx = ctx;
  v->n = n;
  v->f32 = dtype == LGPU_F32;
  v->ptr = device_ptr;
  *out = v.release();
  API_END
}

// Destroying an object does not touch its context (which may already be
// gone): hipFree waits for the device work that still uses the memory.
extern "C" int lgpu_vec_destroy(lgpu_vec v) {
  API_BEGIN
  delete v;
  API_END
}

extern "C" int lgpu_vec_size(lgpu_vec v, int64_t* n) {
  API_BEGIN
  LGPU_REQUIRE(v && n, "NULL argument");
  *n = v->n;
  API_END
}

extern "C" void* lgpu_vec_data(lgpu_vec v) { return v ? v->ptr : nullptr; }

extern "C" int lgpu_vec_upload(lgpu_vec v, const void* host, int64_t n) {
  API_BEGIN
  LGPU_REQUIRE(v && (host || n == 0), "NULL argument");
  LGPU_REQUIRE(n == v->n, "size mismatch");
  LGPU_HIP_CHECK(hipMemcpyAsync(v->ptr, host, size_t(n) * esize(v->f32), hipMemcpyHostToDevice, v->ctx->stream));
  LGPU_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
  API_END
}

extern "C" int lgpu_vec_download(lgpu_vec v, void* host, int64_t n) {
  API_BEGIN
  LGPU_REQUIRE(v && (host || n == 0), "NULL argument");
  LGPU_REQUIRE(n == v->n, "size mismatch");
  LGPU_HIP_CHECK(hipMemcpyAsync(host, v->ptr, size_t(n) * esize(v->f32), hipMemcpyDeviceToHost, v->ctx->stream));
  LGPU_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
  API_END
}

extern "C" int lgpu_vec_fill(lgpu_vec v, double value) {
  API_BEGIN
  LGPU_REQUIRE(v, "NULL vector");
  vec_fill(v->ptr, v->n, v->f32, value, v->ctx->stream);
  API_END
}

extern "C" int lgpu_vec_copy(lgpu_vec src, lgpu_vec dst) {
  API_BEGIN
  LGPU_REQUIRE(src && dst && src->n == dst->n, "invalid copy");
  vec_convert(src->ptr, src->f32, dst->ptr, dst->f32, src->n, src->ctx->stream);
  API_END
}

extern "C" int lgpu_vec_dot(lgpu_vec a, lgpu_vec b, double* out) {
  API_BEGIN
  LGPU_REQUIRE(a && b && out && a->n == b->n && a->f32 == b->f32, "invalid dot");
  dot_into(a->ctx, a->ptr, b->ptr, a->f32, a->n, a->ctx->scal.as<double>() + SC_RR);
  *out = read_scalar(a->ctx, SC_RR);
  API_END
}

// ---------------------------------------------------------------------------
// operators

namespace {

void check_box(const lgpu_box* box) {
  LGPU_REQUIRE(box != nullptr, "box is NULL");
  for (int k = 0; k < 3; ++k) {
    LGPU_REQUIRE(box->n[k] > 0, "box extents must be positive");
    LGPU_REQUIRE(box->len[k] > 0, "box lengths must be positive");
  }
  LGPU_REQUIRE(box->n[0] * box->n[1] * box->n[2] < (int64_t(1) << 31), "box has more than 2^31 cells");
  LGPU_REQUIRE(std::isfinite(box->velocity_x) && box->velocity_x >= 0.0,
               "velocity_x must be finite and non-negative");
}

}  // namespace

extern "C" int lgpu_stencil_values_create(lgpu_ctx ctx, const lgpu_box* box, const double* coeff,
                                          lgpu_op* out) {
  API_BEGIN
  LGPU_REQUIRE(ctx && out && coeff, "NULL argument");
  check_box(box);
  auto op = std::make_unique<lgpu_op_s>();
  op->ctx = ctx;
  op->box = *box;
  fill_desc(op.get(), box);
  op->desc.kind = 0;
  const int64_t n = op->rows();
  op->host_coeff.assign(coeff, coeff + n);
  {
    bool f64ok = true, f32ok = true;
    for (int64_t i = 0; i < n; ++i) {
      const double c = coeff[i];
      const float c32 = float(c);
      f64ok = f64ok && ((c == 0.0 && !std::signbit(c)) || (c >= 0x1p-100 && c <= 0x1p100));
      f32ok = f32ok && ((c32 == 0.0f && !std::signbit(c32)) || (c32 >= 0x1p-24f && c32 <= 0x1p24f));
    }
    op->desc.fast_div64 = f64ok;
    op->desc.fast_div32 = f32ok;
  }
  op->c64.alloc(size_t(n) * sizeof(double));
  LGPU_HIP_CHECK(hipMemcpy(op->c64.ptr, coeff, size_t(n) * sizeof(double), hipMemcpyHostToDevice));
  op->desc.c64 = op->c64.as<double>();
  init_sor_sched(op.get());
  LGPU_HIP_CHECK(hipDeviceSynchronize());
  *out = op.release();
  API_END
}

extern "C" int lgpu_stencil_lookup_create(lgpu_ctx ctx, const lgpu_box* box,
                                          const uint8_t* materials, const double mat_coeffs[4],
                                          lgpu_op* out) {
  API_BEGIN
  LGPU_REQUIRE(ctx && out && materials && mat_coeffs, "NULL argument");
  check_box(box);
  auto op = std::make_unique<lgpu_op_s>();
  op->ctx = ctx;
  op->box = *box;
  fill_desc(op.get(), box);
  StencilDesc& d = op->desc;
  d.kind = 1;
  const int64_t n = op->rows();
  for (int64_t i = 0; i < n; ++i) LGPU_REQUIRE(materials[i] < 4, "material id out of range");
  for (int m = 0; m < 4; ++m) d.mc[m] = mat_coeffs[m];
  // connection tables, lookup-wright.hpp:73-87
  for (int ax = 0; ax < 3; ++ax)
    for (int m = 0; m < 16; ++m) {
      d.tab64[ax][m] = host_hmean(mat_coeffs[m / 4], mat_coeffs[m % 4]) * d.q[ax];
      d.tab32[ax][m] = host_hmean(float(mat_coeffs[m / 4]), float(mat_coeffs[m % 4])) * float(d.q[ax]);
    }
  op->host_mat.assign(materials, materials + n);
  op->mat.alloc(size_t(n));
  LGPU_HIP_CHECK(hipMemcpy(op->mat.ptr, materials, size_t(n), hipMemcpyHostToDevice));
  d.mat = op->mat.as<uint8_t>();
  init_sor_sched(op.get());
  LGPU_HIP_CHECK(hipDeviceSynchronize());
  *out = op.release();
  API_END
}

extern "C" int lgpu_csr_create(lgpu_ctx ctx, int64_t n, const int64_t* rowptr, const int32_t* cols,
                               const double* vals, int dtype, int symmetric, lgpu_op* out) {
  API_BEGIN
  LGPU_REQUIRE(ctx && out && rowptr && cols && vals && n > 0, "invalid CSR arguments");
  LGPU_REQUIRE(dtype == LGPU_F64 || dtype == LGPU_F32, "unknown dtype");
  auto op = std::make_unique<lgpu_op_s>();
  op->ctx = ctx;
  op->is_csr = true;
  setup::RowMatrix& m = op->host_csr;
  m.n = n;
  m.symmetric = symmetric != 0;
  m.f32 = dtype == LGPU_F32;
  m.rowptr.assign(rowptr, rowptr + n + 1);
  const int64_t nnz = rowptr[n];
  LGPU_REQUIRE(rowptr[0] == 0 && nnz >= n, "malformed row pointer");
  m.cols.assign(cols, cols + nnz);
  m.vals.resize(size_t(nnz));
  for (int64_t i = 0; i < n; ++i) {
    bool diag = false;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      LGPU_REQUIRE(cols[k] >= 0 && cols[k] < n, "column index out of range");
      if (k > rowptr[i]) LGPU_REQUIRE(cols[k] > cols[k - 1], "columns must be sorted and unique");
      diag = diag || cols[k] == i;
      m.vals[size_t(k)] = m.f32 ? double(float(vals[k])) : vals[k];
    }
    LGPU_REQUIRE(diag, "every row must store its diagonal");
  }
  op->csr.upload(m, m.f32);
  LGPU_HIP_CHECK(hipDeviceSynchronize());
  *out = op.release();
  API_END
}

namespace {

// csr_from_adjacent_stencil (tensor/dynamic/csr/create-from.hpp:37-239): rows in
// order, an off-diagonal kept iff nonzero, values cast to the matrix type,
// columns sorted (csr/shared.hpp:751-763).  With compression a row whose
// diagonal is zero is dropped and the kept rows renumbered in order
// (create-from.hpp:92-99, 197-229); `map` receives the row of every cell or -1
// and `keep` the cell of every row.
std::unique_ptr<lgpu_op_s> csr_copy(lgpu_op stencil, int dtype, bool compress,
                                    std::vector<int64_t>& map, std::vector<int64_t>& keep) {
  LGPU_REQUIRE(stencil && !stencil->is_csr, "needs a stencil operator");
  LGPU_REQUIRE(dtype == LGPU_F64 || dtype == LGPU_F32, "unknown dtype");
  setup::RowMatrix rows;
  stencil_host_rows(stencil, rows);
  const int64_t n = rows.n;
  map.assign(size_t(n), 0);
  keep.clear();
  for (int64_t i = 0; i < n; ++i) {
    const bool kept = !compress || rows.vals[size_t(rows.rowptr[i + 1] - 1)] != 0.0;
    map[size_t(i)] = kept ? int64_t(keep.size()) : -1;
    if (kept) keep.push_back(i);
  }
  auto op = std::make_unique<lgpu_op_s>();
  op->ctx = stencil->ctx;
  op->is_csr = true;
  setup::RowMatrix& m = op->host_csr;
  m.n = int64_t(keep.size());
  m.f32 = dtype == LGPU_F32;
  m.symmetric = rows.symmetric;
  m.rowptr.assign(1, 0);
  std::vector<std::pair<int32_t, double>> row;
  for (const int64_t i : keep) {
    row.clear();
    for (int64_t k = rows.rowptr[i]; k < rows.rowptr[i + 1]; ++k) {
      if (rows.cols[k] != i && rows.vals[k] == 0.0) continue;
      // a kept coupling never reaches a dropped cell: the coupling to a cell
      // with a zero diagonal is itself zero
      const int64_t c = map[size_t(rows.cols[k])];
      LGPU_REQUIRE(c >= 0, "a nonzero coupling reaches a dropped row");
      row.emplace_back(int32_t(c), m.f32 ? double(float(rows.vals[k])) : rows.vals[k]);
    }
    std::sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (const auto& [c, v] : row) {
      m.cols.push_back(c);
      m.vals.push_back(v);
    }
    m.rowptr.push_back(int64_t(m.cols.size()));
  }
  op->csr.upload(m, m.f32);
  return op;
}

}  // namespace

extern "C" int lgpu_csr_from_stencil(lgpu_op stencil, int dtype, lgpu_op* out) {
  API_BEGIN
  LGPU_REQUIRE(out, "NULL output");
  std::vector<int64_t> map, keep;
  auto op = csr_copy(stencil, dtype, false, map, keep);
  LGPU_HIP_CHECK(hipDeviceSynchronize());
  *out = op.release();
  API_END
}

extern "C" int lgpu_csr_system_from_stencil(lgpu_op stencil, int dtype, int compress,
                                            lgpu_op* lhs, lgpu_vec* rhs, int64_t* cmap) {
  API_BEGIN
  LGPU_REQUIRE(lhs && rhs, "NULL output");
  std::vector<int64_t> map, keep;
  auto op = csr_copy(stencil, dtype, compress != 0, map, keep);
  const bool f32 = dtype == LGPU_F32;
  if (f32) stencil->ensure_f32();
  hipStream_t st = stencil->ctx->stream;
  auto b = std::make_unique<lgpu_vec_s>();
  b->ctx = stencil->ctx;
  b->n = op->csr.n;
  b->f32 = f32;
  b->own.alloc(size_t(b->n) * esize(f32));
  b->ptr = b->own.ptr;
  if (compress) {
    const int64_t n = stencil->rows();
    op->cmap.alloc(size_t(n) * sizeof(int64_t));
    op->cmap_cells = n;
    LGPU_HIP_CHECK(hipMemcpy(op->cmap.ptr, map.data(), size_t(n) * sizeof(int64_t), hipMemcpyHostToDevice));
    DevBuf full, rows;
    full.alloc(size_t(n) * esize(f32));
    rows.alloc(keep.size() * sizeof(int64_t));
    LGPU_HIP_CHECK(hipMemcpy(rows.ptr, keep.data(), keep.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    stencil_rhs(stencil->desc, f32, full.ptr, st);
    vec_gather(full.ptr, f32, rows.as<int64_t>(), b->n, b->ptr, st);
    LGPU_HIP_CHECK(hipStreamSynchronize(st));
  } else {
    stencil_rhs(stencil->desc, f32, b->ptr, st);
    LGPU_HIP_CHECK(hipStreamSynchronize(st));
  }
  if (cmap) std::copy(map.begin(), map.end(), cmap);
  *lhs = op.release();
  *rhs = b.release();
  API_END
}

extern "C" int lgpu_op_destroy(lgpu_op op) {
  API_BEGIN
  delete op;
  API_END
}

extern "C" int lgpu_op_rows(lgpu_op op, int64_t* n) {
  API_BEGIN
  LGPU_REQUIRE(op && n, "NULL argument");
  *n = op->rows();
  API_END
}

extern "C" int lgpu_op_rhs(lgpu_op op, lgpu_vec b) {
  API_BEGIN
  LGPU_REQUIRE(op && !op->is_csr, "the Dirichlet right-hand side needs a stencil operator");
  check_vec(b, op->rows(), "b");
  if (b->f32) op->ensure_f32();
  stencil_rhs(op->desc, b->f32, b->ptr, op->ctx->stream);
  API_END
}

extern "C" int lgpu_op_spmv(lgpu_op op, lgpu_vec x, lgpu_vec y) {
  API_BEGIN
  LGPU_REQUIRE(op, "NULL operator");
  check_vec(x, op->rows(), "x");
  check_vec(y, op->rows(), "y");
  LGPU_REQUIRE(x->f32 == y->f32, "x and y must share a dtype");
  LGPU_REQUIRE(x->ptr != y->ptr, "x and y must not alias");
  op_apply(op, SM_SPMV, x->f32, x->ptr, nullptr, y->ptr, 0.0, nullptr);
  API_END
}

extern "C" int lgpu_op_residual(lgpu_op op, lgpu_vec b, lgpu_vec x, lgpu_vec r) {
  API_BEGIN
  LGPU_REQUIRE(op, "NULL operator");
  check_vec(b, op->rows(), "b");
  check_vec(x, op->rows(), "x");
  check_vec(r, op->rows(), "r");
  LGPU_REQUIRE(x->f32 == b->f32 && x->f32 == r->f32, "vectors must share a dtype");
  LGPU_REQUIRE(x->ptr != r->ptr, "x and r must not alias");
  op_apply(op, SM_RESID, x->f32, x->ptr, b->ptr, r->ptr, 0.0, nullptr);
  API_END
}

extern "C" int lgpu_op_jacobi(lgpu_op op, lgpu_vec x_in, lgpu_vec x_out, lgpu_vec b, double omega) {
  API_BEGIN
  LGPU_REQUIRE(op, "NULL operator");
  check_vec(x_out, op->rows(), "x_out");
  check_vec(b, op->rows(), "b");
  if (x_in) {
    check_vec(x_in, op->rows(), "x_in");
    LGPU_REQUIRE(x_in->f32 == x_out->f32 && x_in->ptr != x_out->ptr, "x_in must match x_out and not alias it");
  }
  LGPU_REQUIRE(b->f32 == x_out->f32, "vectors must share a dtype");
  op_apply(op, SM_JACOBI, x_out->f32, x_in ? x_in->ptr : nullptr, b->ptr, x_out->ptr, omega, nullptr);
  API_END
}

extern "C" int lgpu_op_sor(lgpu_op op, lgpu_vec x_in, lgpu_vec x_out, lgpu_vec b, double omega,
                           int forward, int ultra) {
  API_BEGIN
  LGPU_REQUIRE(op, "NULL operator");
  check_vec(x_out, op->rows(), "x_out");
  check_vec(b, op->rows(), "b");
  if (x_in) {
    check_vec(x_in, op->rows(), "x_in");
    LGPU_REQUIRE(x_in->f32 == x_out->f32 && x_in->ptr != x_out->ptr, "x_in must match x_out and not alias it");
  }
  LGPU_REQUIRE(b->f32 == x_out->f32, "vectors must share a dtype");
  op_sor(op, x_out->f32, x_in ? x_in->ptr : nullptr, x_out->ptr, b->ptr, omega, forward != 0, ultra != 0);
  if (take_wait_error(op->ctx)) return wait_timeout();
  API_END
}

extern "C" int lgpu_op_set_sor_cache(lgpu_op op, int enable) {
  API_BEGIN
  LGPU_REQUIRE(op, "NULL operator");
  op->sor_cache = enable != 0;
  if (!op->sor_cache) {
    LGPU_HIP_CHECK(hipStreamSynchronize(op->ctx->stream));
    op->sor_w64.release();
    op->sor_w32.release();
  }
  API_END
}

extern "C" int lgpu_op_ssor(lgpu_op op, lgpu_vec x_in, lgpu_vec x_out, lgpu_vec b, double omega,
                            int ultra) {
  API_BEGIN
  LGPU_REQUIRE(op, "NULL operator");
  check_vec(x_out, op->rows(), "x_out");
  check_vec(b, op->rows(), "b");
  if (x_in) {
    check_vec(x_in, op->rows(), "x_in");
    LGPU_REQUIRE(x_in->f32 == x_out->f32, "x_in and x_out must share a dtype");
  }
  LGPU_REQUIRE(b->f32 == x_out->f32, "vectors must share a dtype");
  ssor_sweep(op, x_out->f32, x_in ? x_in->ptr : nullptr, x_out->ptr, b->ptr, omega, ultra != 0);
  if (take_wait_error(op->ctx)) return wait_timeout();
  API_END
}

extern "C" int lgpu_op_ssor_noalloc(lgpu_op op, lgpu_vec x_in, lgpu_vec x_out, lgpu_vec b,
                                     double omega, int ultra) {
  API_BEGIN
  LGPU_REQUIRE(op, "NULL operator");
  check_vec(x_out, op->rows(), "x_out");
  check_vec(b, op->rows(), "b");
  if (x_in) {
    check_vec(x_in, op->rows(), "x_in");
    LGPU_REQUIRE(x_in->f32 == x_out->f32, "x_in and x_out must share a dtype");
  }
  LGPU_REQUIRE(b->f32 == x_out->f32, "vectors must share a dtype");
  LGPU_REQUIRE(!op->is_csr || op->csr.f32 == x_out->f32, "CSR value type and vector type differ");
  LGPU_REQUIRE(op->rows() <= kScratchDoubles,
               "lgpu_op_ssor_noalloc: the operator has more rows than the static scratch holds "
               "(4194304); use lgpu_op_ssor");
  // the f32 coefficient copy of a per-cell stencil is made by the first f32
  // pass; making it here would allocate
  LGPU_REQUIRE(!x_out->f32 || op->is_csr || op->desc.kind != 0 || op->c32.ptr,
               "lgpu_op_ssor_noalloc: f32 sweeps need the operator's f32 coefficient copy, "
               "which the first f32 pass of another entry point makes");
  lgpu_ctx ctx = op->ctx;
  ssor_sweep_into(op, x_out->f32, x_in ? x_in->ptr : nullptr, x_out->ptr, b->ptr, omega,
                  ultra != 0, scratch_ptr(ctx->device));
  // the wait-error word through the context's pinned mirror
  LGPU_HIP_CHECK(hipMemcpyAsync(ctx->host_scal + SC_ERR, ctx->err(), sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  LGPU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (ctx->host_err()) {
    vec_fill_int(ctx->err(), 1, 0, ctx->stream);
    LGPU_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return wait_timeout();
  }
  API_END
}

extern "C" int lgpu_ssor_precon_create(lgpu_op op, double omega, int ultra, int iterations,
                                       lgpu_amg* out) {
  API_BEGIN
  LGPU_REQUIRE(op && out, "NULL argument");
  LGPU_REQUIRE(iterations >= 1, "iterations must be at least 1");
  auto amg = std::make_unique<lgpu_amg_s>();
  amg->ctx = op->ctx;
  amg->finest = op;
  amg->ssor_iters = iterations;
  amg->ssor_omega = omega;
  amg->ssor_ultra = ultra != 0;
  *out = amg.release();
  API_END
}

extern "C" int lgpu_op_sor_segmented(lgpu_op op, lgpu_vec x_in, lgpu_vec x_out, lgpu_vec b,
                                     double omega, int forward, int ultra, int64_t segment_rows) {
  API_BEGIN
  LGPU_REQUIRE(op && op->is_csr, "segmented SOR needs a CSR operator");
  LGPU_REQUIRE(segment_rows >= 1 && segment_rows <= kSchedMaxRows, "segment_rows must lie in 1..8192");
  check_vec(x_out, op->rows(), "x_out");
  check_vec(b, op->rows(), "b");
  if (x_in) {
    check_vec(x_in, op->rows(), "x_in");
    LGPU_REQUIRE(x_in->f32 == x_out->f32 && x_in->ptr != x_out->ptr, "x_in must match x_out and not alias it");
  }
  LGPU_REQUIRE(b->f32 == x_out->f32 && op->csr.f32 == x_out->f32, "vectors and values must share a dtype");
  const int64_t n = op->rows();
  // one segment over more rows than a wavefront: the round program, built on
  // first use, when the level is small enough
  if (segment_rows >= n && n > kMaxSegmentRows && n <= kRoundMaxRows &&
      (op->csr.rounds[0].nrounds > 0 || op->csr.build_rounds(op->host_csr, tail_lds_budget()))) {
    csr_sor_rounds(op->csr.desc(), op->csr.rounds[forward ? 0 : 1], x_in ? x_in->ptr : nullptr, x_out->ptr,
                   b->ptr, omega, ultra != 0, op->ctx->stream);
  } else if (segment_rows <= kMaxSegmentRows) {
    csr_sor_segmented(op->csr.desc(), x_in ? x_in->ptr : nullptr, x_out->ptr, b->ptr, omega,
                      forward != 0, ultra != 0, int(segment_rows), op->ctx->stream);
  } else {
    // longer segments: the scheduled sweep, its positions built on first use
    const int64_t seg = std::min(segment_rows, op->rows());
    if (op->csr.sched_seg != seg) op->csr.build_schedule(op->host_csr, seg);
    csr_sor_scheduled(op->csr.desc(), x_in ? x_in->ptr : nullptr, x_out->ptr, b->ptr, omega,
                      forward != 0, ultra != 0, op->ctx->stream);
  }
  API_END
}

// ---------------------------------------------------------------------------
// CSR with block values

struct lgpu_bcsr_s {
  lgpu_ctx ctx = nullptr;
  int64_t nb = 0;
  int block = 1;
  bool f32 = false;
  DevBuf rowptr, cols, vals;
  BcsrDesc desc() const {
    return BcsrDesc{nb, block, rowptr.as<int64_t>(), cols.as<int32_t>(), vals.ptr, f32 ? 1 : 0};
  }
};

extern "C" int lgpu_bcsr_create(lgpu_ctx ctx, int64_t block_rows, int block_size,
                                const int64_t* rowptr, const int32_t* cols, const double* vals,
                                int dtype, lgpu_bcsr* out) {
  API_BEGIN
  LGPU_REQUIRE(ctx && out && rowptr && block_rows >= 0, "invalid block CSR arguments");
  LGPU_REQUIRE(block_size >= 1 && block_size <= 4, "block_size must lie in 1..4");
  LGPU_REQUIRE(dtype == LGPU_F64 || dtype == LGPU_F32, "unknown dtype");
  LGPU_REQUIRE(rowptr[0] == 0, "malformed row pointer");
  const int64_t nnz = rowptr[block_rows];
  LGPU_REQUIRE(nnz == 0 || (cols && vals), "NULL column or value array");
  for (int64_t i = 0; i < block_rows; ++i) {
    LGPU_REQUIRE(rowptr[i + 1] >= rowptr[i], "malformed row pointer");
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      LGPU_REQUIRE(cols[k] >= 0 && cols[k] < block_rows, "column index out of range");
  }
  auto m = std::make_unique<lgpu_bcsr_s>();
  m->ctx = ctx;
  m->nb = block_rows;
  m->block = block_size;
  m->f32 = dtype == LGPU_F32;
  const size_t ne = size_t(nnz) * size_t(block_size * block_size);
  m->rowptr.alloc(size_t(block_rows + 1) * sizeof(int64_t));
  LGPU_HIP_CHECK(hipMemcpy(m->rowptr.ptr, rowptr, size_t(block_rows + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  m->cols.alloc(size_t(nnz) * sizeof(int32_t));
  if (nnz) LGPU_HIP_CHECK(hipMemcpy(m->cols.ptr, cols, size_t(nnz) * sizeof(int32_t), hipMemcpyHostToDevice));
  if (m->f32) {
    std::vector<float> v(vals, vals + ne);
    m->vals.alloc(ne * sizeof(float));
    if (ne) LGPU_HIP_CHECK(hipMemcpy(m->vals.ptr, v.data(), ne * sizeof(float), hipMemcpyHostToDevice));
  } else {
    m->vals.alloc(ne * sizeof(double));
    if (ne) LGPU_HIP_CHECK(hipMemcpy(m->vals.ptr, vals, ne * sizeof(double), hipMemcpyHostToDevice));
  }
  *out = m.release();
  API_END
}

extern "C" int lgpu_bcsr_destroy(lgpu_bcsr m) {
  API_BEGIN
  delete m;
  API_END
}

extern "C" int lgpu_bcsr_rows(lgpu_bcsr m, int64_t* block_rows, int* block_size) {
  API_BEGIN
  LGPU_REQUIRE(m, "NULL matrix");
  if (block_rows) *block_rows = m->nb;
  if (block_size) *block_size = m->block;
  API_END
}

extern "C" int lgpu_bcsr_spmv(lgpu_bcsr m, lgpu_vec x, lgpu_vec y) {
  API_BEGIN
  LGPU_REQUIRE(m, "NULL matrix");
  const int64_t n = m->nb * m->block;
  check_vec(x, n, "x");
  check_vec(y, n, "y");
  LGPU_REQUIRE(x->f32 == m->f32 && y->f32 == m->f32, "the vectors need the matrix's value type");
  LGPU_REQUIRE(x->ptr != y->ptr, "x and y must not alias");
  KTimed kt(m->ctx, LGPU_KT_CSR_PASS);
  bcsr_spmv(m->desc(), x->ptr, y->ptr, m->ctx->stream);
  API_END
}

// ---------------------------------------------------------------------------
// AMG

extern "C" void lgpu_amg_default_params(lgpu_amg_params* p) {
  if (!p) return;
  p->max_level_num = 16;
  p->min_coarsen_size = 32;
  p->min_coarsen_factor = 1.2;
  p->strong_threshold = 1.0 / 3.0;
  p->aggressive_strong_threshold = 1.0 / 3.0;
  p->min_aggregate_size = 6;
  p->max_aggregate_size = 9;
  p->precision = LGPU_F64;
  p->ultra = 0;
  p->coarse_segment_rows = 64;
  p->relax = 1.0;
  p->cycle_kind = LGPU_CYCLE_REGULAR;
  p->cycle_iterations = 1;
  p->setup_on_device = 1;
  p->sequential_rows = kRoundMaxRows;
}

namespace {

void check_amg_params(lgpu_op finest, const lgpu_amg_params* params) {
  LGPU_REQUIRE(params->precision == LGPU_F64 || params->precision == LGPU_F32, "unknown precision");
  LGPU_REQUIRE(params->max_aggregate_size >= 1 && params->min_aggregate_size >= 1, "invalid aggregate sizes");
  LGPU_REQUIRE(params->coarse_segment_rows >= 0 && params->coarse_segment_rows <= kSchedMaxRows,
               "coarse_segment_rows must lie in 0..8192");
  LGPU_REQUIRE(params->sequential_rows >= 0, "sequential_rows must not be negative");
  LGPU_REQUIRE(!finest->is_csr || !finest->csr.f32, "the finest CSR operator must hold f64 values");
  LGPU_REQUIRE(std::isfinite(params->relax) && params->relax != 0.0, "relax must be finite and nonzero");
  LGPU_REQUIRE(params->cycle_kind == LGPU_CYCLE_REGULAR || params->cycle_kind == LGPU_CYCLE_DECREASING,
               "unknown cycle kind");
  LGPU_REQUIRE(params->cycle_iterations >= 1, "cycle_iterations must be at least 1");
}

// Rows per smoother segment of a coarse level of n rows (0: one segment):
// one segment when n is at most sequential_rows, else coarse_segment_rows.
int64_t level_segment_rows(const lgpu_amg_params& p, int64_t n) {
  if (n <= p.sequential_rows) return 0;
  return p.coarse_segment_rows >= n ? 0 : p.coarse_segment_rows;
}

// The V-cycle's tail: the longest run of levels before the coarsest whose
// smoothers are 64-row wavefront segments staged in LDS and that have at most
// kTailMaxRows rows, plus the coarsest; amg_tail runs them in one workgroup.
void build_tail(lgpu_amg amg) {
  amg->tail_count = 0;
  const int L = int(amg->lv.size());
  if (L < 2 || amg->lv[size_t(L - 1)].n <= 0) return;
  const size_t budget = tail_lds_budget();
  int t0 = L - 1;
  size_t stride = 0, block_need = 0;  // LDS per wavefront (64-row segments), per workgroup (rounds)
  while (t0 - 1 >= 1) {
    const AmgLevel& a = amg->lv[size_t(t0 - 1)];
    if (a.n <= 0 || a.n > kTailMaxRows) break;
    if (a.seg == 0 && a.a.rounds[0].nrounds > 0) {
      const size_t need = std::max(round_lds_bytes(a.n, a.a.rounds[0].entries, a.a.rounds[0].nrounds, amg->f32),
                                   round_lds_bytes(a.n, a.a.rounds[1].entries, a.a.rounds[1].nrounds, amg->f32));
      if (need + 16 * kTailMaxWaves > budget) break;  // room for the per-wavefront rounding below
      block_need = std::max(block_need, need);
    } else {
      if (a.seg != kMaxSegmentRows || a.a.sched_seg != 0) break;
      const size_t need = tail_wave_lds_bytes(a.a.seg_entries, amg->f32);
      if (need > budget) break;
      stride = std::max(stride, need);
    }
    --t0;
  }
  if (t0 >= L - 1) return;  // no level besides the coarsest
  const int waves = stride > 0 ? int(std::min<size_t>(kTailMaxWaves, budget / stride)) : kTailMaxWaves;
  // the round sweeps use the workgroup's whole LDS from its start
  stride = std::max(stride, (block_need / size_t(waves) + 15) & ~size_t(15));
  std::vector<TailLevel> tl;
  for (int l = t0; l < L; ++l) {
    AmgLevel& a = amg->lv[size_t(l)];
    TailLevel t{};
    t.n = a.n;
    t.rowptr = a.a.rowptr.as<int64_t>();
    t.cols = a.a.cols.as<int32_t>();
    t.vals = a.a.vals.ptr;
    t.cap = a.a.seg_entries;
    t.shortlists = a.a.seg_new_max <= kSegNewShort && a.a.seg_trail_max <= kSegTrailShort;
    t.agg = a.agg.as<uint32_t>();
    t.c2f_ptr = a.c2f_ptr.as<int64_t>();
    t.c2f_idx = a.c2f_idx.as<int32_t>();
    t.rhs = a.rhs.ptr;
    t.sol = a.sol.ptr;
    t.tmp = a.tmp.ptr;
    t.aux = a.aux.ptr;
    if (l < L - 1 && a.seg == 0) {
      t.rounds_fwd = a.a.rounds[0];
      t.rounds_bwd = a.a.rounds[1];
    }
    tl.push_back(t);
  }
  amg->tail.alloc(tl.size() * sizeof(TailLevel));
  LGPU_HIP_CHECK(hipMemcpy(amg->tail.ptr, tl.data(), tl.size() * sizeof(TailLevel), hipMemcpyHostToDevice));
  amg->tail_first = t0;
  amg->tail_count = int(tl.size());
  amg->tail_waves = waves;
  amg->tail_wave_lds = stride;
}

// Uploads the host hierarchy amg->host (levels, transfers, coarse matrices
// and the coarsest factor) and allocates the level vectors.
void upload_hierarchy(lgpu_amg amg) {
  LGPU_REQUIRE(amg->host.spd, amg->host.lu ? "the coarsest LU factorisation met a zero pivot"
                                          : "coarsest matrix is not positive definite");
  const setup::Hierarchy& h = amg->host;
  const int L = int(h.sizes.size());
  LGPU_REQUIRE(h.sizes.back() <= 4096, "coarsest level has more than 4096 rows");
  amg->lv.resize(size_t(L));
  const size_t es = esize(amg->f32);
  for (int l = 0; l < L; ++l) {
    AmgLevel& a = amg->lv[size_t(l)];
    a.n = h.sizes[size_t(l)];
    if (l > 0) {
      const setup::RowMatrix& m = h.coarse[size_t(l - 1)];
      a.a.upload(m, amg->f32);
      a.seg = level_segment_rows(amg->params, a.n);
      // a level swept in one segment that fits one workgroup's LDS, and
      // segments longer than a wavefront, run the scheduled sweep
      if (a.n > 0 && ((a.seg == 0 && a.n <= kSchedMaxRows) || a.seg > kMaxSegmentRows))
        a.a.build_schedule(m, a.seg == 0 ? a.n : a.seg);
      // a small level in one segment: the round programs
      if (a.seg == 0 && a.n > 0 && a.n <= kRoundMaxRows) a.a.build_rounds(m, tail_lds_budget());
      // the level vectors persist between cycles, as the reference's
      // hierarchy vectors do (hierarchy/wright.hpp:227-228): a DECREASING
      // cycle prolongs a coarse solution no visit of the same cycle wrote.
      // They start at zero (the reference leaves them uninitialised).
      a.rhs.alloc(size_t(a.n) * es);
      a.sol.alloc(size_t(a.n) * es);
      LGPU_HIP_CHECK(hipMemset(a.rhs.ptr, 0, size_t(a.n) * es));
      LGPU_HIP_CHECK(hipMemset(a.sol.ptr, 0, size_t(a.n) * es));
    }
    if (l + 1 < L) {
      const setup::Level& t = h.transfer[size_t(l)];
      a.agg.alloc(t.agg.size() * sizeof(uint32_t));
      LGPU_HIP_CHECK(hipMemcpy(a.agg.ptr, t.agg.data(), t.agg.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      a.c2f_ptr.alloc(t.c2f_ptr.size() * sizeof(int64_t));
      LGPU_HIP_CHECK(hipMemcpy(a.c2f_ptr.ptr, t.c2f_ptr.data(), t.c2f_ptr.size() * sizeof(int64_t), hipMemcpyHostToDevice));
      a.c2f_idx.alloc(t.c2f_idx.size() * sizeof(int32_t));
      LGPU_HIP_CHECK(hipMemcpy(a.c2f_idx.ptr, t.c2f_idx.data(), t.c2f_idx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
      const int64_t ncl = int64_t(t.c2f_ptr.size()) - 1;
      a.c2f_blk = 0;
      for (int64_t c0 = 0; c0 < ncl; c0 += kCsrBlockRows)
        a.c2f_blk = std::max(a.c2f_blk, t.c2f_ptr[size_t(std::min(c0 + kCsrBlockRows, ncl))] -
                                            t.c2f_ptr[size_t(c0)]);
      a.tmp.alloc(size_t(a.n) * es);
      a.aux.alloc(size_t(a.n) * es);
    }
  }
  const int64_t nc = h.sizes.back();
  amg->chol.alloc(size_t(nc * nc) * sizeof(double));
  LGPU_HIP_CHECK(hipMemcpy(amg->chol.ptr, h.chol_dense.data(), size_t(nc * nc) * sizeof(double), hipMemcpyHostToDevice));
  if (h.lu) {
    amg->upper.alloc(size_t(nc * nc) * sizeof(double));
    LGPU_HIP_CHECK(hipMemcpy(amg->upper.ptr, h.upper_dense.data(), size_t(nc * nc) * sizeof(double), hipMemcpyHostToDevice));
  }
  if (amg->f32) {
    amg->finest->ensure_f32();
    amg->r32.alloc(size_t(h.sizes[0]) * sizeof(float));
    amg->z32.alloc(size_t(h.sizes[0]) * sizeof(float));
  }
  build_tail(amg);
  // the finest smoother's cached weights, built with the hierarchy
  amg->finest->sor_weights(amg->f32, true);
  LGPU_HIP_CHECK(hipDeviceSynchronize());
}

}  // namespace

extern "C" int lgpu_amg_create(lgpu_op finest, const lgpu_amg_params* params, lgpu_amg* out) {
  API_BEGIN
  LGPU_REQUIRE(finest && params && out, "NULL argument");
  check_amg_params(finest, params);
  const auto t0 = std::chrono::steady_clock::now();
  auto amg = std::make_unique<lgpu_amg_s>();
  amg->ctx = finest->ctx;
  amg->finest = finest;
  amg->params = *params;
  amg->f32 = params->precision == LGPU_F32;
  setup::Params sp;
  sp.max_level_num = params->max_level_num;
  sp.min_coarsen_size = params->min_coarsen_size;
  sp.min_coarsen_factor = params->min_coarsen_factor;
  sp.strong_threshold = params->strong_threshold;
  sp.aggressive_strong_threshold = params->aggressive_strong_threshold;
  sp.min_aggregate_size = params->min_aggregate_size;
  sp.max_aggregate_size = params->max_aggregate_size;
  sp.coarse_f32 = amg->f32;
  std::unique_ptr<setup::Accel> accel;
  if (params->setup_on_device) accel = make_device_setup(finest->ctx->stream);
  if (finest->is_csr) {
    amg->host = setup::build(finest->host_csr, sp, accel.get());
  } else {
    setup::RowMatrix rows;
    stencil_host_rows(finest, rows);
    amg->host = setup::build(rows, sp, accel.get());
  }
  upload_hierarchy(amg.get());
  amg->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *out = amg.release();
  API_END
}

extern "C" int lgpu_amg_create_from_levels(lgpu_op finest, const lgpu_amg_params* params,
                                           int transfer_num, const lgpu_amg_level* levels,
                                           lgpu_amg* out) {
  API_BEGIN
  LGPU_REQUIRE(finest && params && out && levels, "NULL argument");
  LGPU_REQUIRE(transfer_num >= 1 && transfer_num < 32, "transfer_num must lie in 1..31");
  check_amg_params(finest, params);
  const auto t0 = std::chrono::steady_clock::now();
  auto amg = std::make_unique<lgpu_amg_s>();
  amg->ctx = finest->ctx;
  amg->finest = finest;
  amg->params = *params;
  amg->f32 = params->precision == LGPU_F32;
  setup::Hierarchy& h = amg->host;
  const bool symmetric = finest->is_csr ? finest->host_csr.symmetric : finest->desc.adv == 0.0;
  int64_t n = finest->rows();
  h.sizes.push_back(n);
  for (int l = 0; l < transfer_num; ++l) {
    const lgpu_amg_level& in = levels[l];
    LGPU_REQUIRE(in.agg && in.rowptr && in.cols && in.vals, "NULL level array");
    const int64_t nc = in.coarse_n;
    LGPU_REQUIRE(nc >= 1 && nc < n + 1, "a coarse level must have between 1 and the fine level's rows");
    setup::Level t;
    t.agg.assign(in.agg, in.agg + n);
    std::vector<int64_t> count(size_t(nc), 0);
    for (int64_t v = 0; v < n; ++v) {
      const uint32_t a = t.agg[size_t(v)];
      LGPU_REQUIRE(a == setup::kIsolated || int64_t(a) < nc, "aggregate index out of range");
      if (a != setup::kIsolated) ++count[a];
    }
    t.c2f_ptr.assign(size_t(nc + 1), 0);
    for (int64_t c = 0; c < nc; ++c) {
      LGPU_REQUIRE(count[size_t(c)] > 0, "every aggregate needs a member");
      t.c2f_ptr[size_t(c + 1)] = t.c2f_ptr[size_t(c)] + count[size_t(c)];
    }
    t.c2f_idx.resize(size_t(t.c2f_ptr[size_t(nc)]));
    if (in.c2f_ptr && in.c2f_idx) {
      // the caller's coarse -> fine lists fix the order of the restriction sums
      for (int64_t c = 0; c <= nc; ++c)
        LGPU_REQUIRE(in.c2f_ptr[c] == t.c2f_ptr[size_t(c)], "c2f_ptr does not match the aggregate map");
      for (int64_t c = 0; c < nc; ++c)
        for (int64_t k = in.c2f_ptr[c]; k < in.c2f_ptr[c + 1]; ++k) {
          const int32_t v = in.c2f_idx[k];
          LGPU_REQUIRE(v >= 0 && v < n && int64_t(t.agg[size_t(v)]) == c,
                       "c2f_idx does not match the aggregate map");
          t.c2f_idx[size_t(k)] = v;
        }
    } else {
      std::vector<int64_t> pos(t.c2f_ptr.begin(), t.c2f_ptr.end() - 1);
      for (int64_t v = 0; v < n; ++v)
        if (t.agg[size_t(v)] != setup::kIsolated) t.c2f_idx[size_t(pos[t.agg[size_t(v)]]++)] = int32_t(v);
    }
    setup::RowMatrix m;
    m.n = nc;
    m.f32 = amg->f32;
    m.symmetric = symmetric;
    m.rowptr.assign(in.rowptr, in.rowptr + nc + 1);
    LGPU_REQUIRE(m.rowptr[0] == 0 && m.rowptr[size_t(nc)] >= nc, "malformed coarse row pointer");
    const int64_t nnz = m.rowptr[size_t(nc)];
    m.cols.assign(in.cols, in.cols + nnz);
    m.vals.resize(size_t(nnz));
    for (int64_t i = 0; i < nc; ++i) {
      bool diag = false;
      LGPU_REQUIRE(m.rowptr[size_t(i + 1)] >= m.rowptr[size_t(i)], "malformed coarse row pointer");
      for (int64_t k = m.rowptr[size_t(i)]; k < m.rowptr[size_t(i + 1)]; ++k) {
        LGPU_REQUIRE(in.cols[k] >= 0 && in.cols[k] < nc, "coarse column out of range");
        if (k > m.rowptr[size_t(i)]) LGPU_REQUIRE(in.cols[k] > in.cols[k - 1], "coarse columns must be sorted and unique");
        diag = diag || in.cols[k] == i;
        m.vals[size_t(k)] = amg->f32 ? double(float(in.vals[k])) : in.vals[k];
      }
      LGPU_REQUIRE(diag, "every coarse row must store its diagonal");
    }
    h.transfer.push_back(std::move(t));
    h.coarse.push_back(std::move(m));
    h.sizes.push_back(nc);
    n = nc;
  }
  const auto tc = std::chrono::steady_clock::now();
  setup::factor_coarsest(h);
  h.seconds_coarsest = std::chrono::duration<double>(std::chrono::steady_clock::now() - tc).count();
  upload_hierarchy(amg.get());
  amg->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *out = amg.release();
  API_END
}

extern "C" int lgpu_amg_transfer(lgpu_amg amg, int level, int64_t* c2f_ptr, int32_t* c2f_idx) {
  API_BEGIN
  LGPU_REQUIRE(amg && level >= 0 && level + 1 < int(amg->lv.size()), "invalid level");
  const setup::Level& t = amg->host.transfer[size_t(level)];
  if (c2f_ptr) std::memcpy(c2f_ptr, t.c2f_ptr.data(), t.c2f_ptr.size() * sizeof(int64_t));
  if (c2f_idx) std::memcpy(c2f_idx, t.c2f_idx.data(), t.c2f_idx.size() * sizeof(int32_t));
  API_END
}

extern "C" int lgpu_amg_destroy(lgpu_amg amg) {
  API_BEGIN
  delete amg;
  API_END
}

extern "C" int lgpu_amg_level_num(lgpu_amg amg, int* levels) {
  API_BEGIN
  LGPU_REQUIRE(amg && levels, "NULL argument");
  *levels = int(amg->lv.size());
  API_END
}

extern "C" int lgpu_amg_level_size(lgpu_amg amg, int level, int64_t* n) {
  API_BEGIN
  LGPU_REQUIRE(amg && n && level >= 0 && level < int(amg->lv.size()), "invalid level");
  *n = amg->lv[size_t(level)].n;
  API_END
}

extern "C" int lgpu_amg_aggregates(lgpu_amg amg, int level, uint32_t* host_map) {
  API_BEGIN
  LGPU_REQUIRE(amg && host_map && level >= 0 && level + 1 < int(amg->lv.size()), "invalid level");
  const auto& agg = amg->host.transfer[size_t(level)].agg;
  std::memcpy(host_map, agg.data(), agg.size() * sizeof(uint32_t));
  API_END
}

extern "C" int lgpu_amg_matrix(lgpu_amg amg, int level, int64_t* nnz, int64_t* rowptr, int32_t* cols,
                               double* vals) {
  API_BEGIN
  LGPU_REQUIRE(amg && nnz && level >= 1 && level < int(amg->lv.size()), "invalid level");
  const setup::RowMatrix& m = amg->host.coarse[size_t(level - 1)];
  *nnz = int64_t(m.cols.size());
  if (rowptr) std::memcpy(rowptr, m.rowptr.data(), m.rowptr.size() * sizeof(int64_t));
  if (cols) std::memcpy(cols, m.cols.data(), m.cols.size() * sizeof(int32_t));
  if (vals) std::memcpy(vals, m.vals.data(), m.vals.size() * sizeof(double));
  API_END
}

extern "C" int lgpu_amg_setup_seconds(lgpu_amg amg, double* seconds) {
  API_BEGIN
  LGPU_REQUIRE(amg && seconds, "NULL argument");
  *seconds = amg->setup_seconds;
  API_END
}

extern "C" int lgpu_amg_setup_split(lgpu_amg amg, double seconds[5]) {
  API_BEGIN
  LGPU_REQUIRE(amg && seconds, "NULL argument");
  const setup::Hierarchy& h = amg->host;
  const double phases = h.seconds_criterion + h.seconds_aggregation + h.seconds_galerkin +
                        h.seconds_coarsest;
  seconds[0] = h.seconds_criterion;
  seconds[1] = h.seconds_aggregation;
  seconds[2] = h.seconds_galerkin;
  seconds[3] = h.seconds_coarsest;
  seconds[4] = std::max(0.0, amg->setup_seconds - phases);
  API_END
}

extern "C" int lgpu_amg_cycle(lgpu_amg amg, lgpu_vec x_in, lgpu_vec r, lgpu_vec z) {
  API_BEGIN
  LGPU_REQUIRE(amg && amg->ssor_iters == 0, "needs a multigrid hierarchy");
  const int64_t n = amg->lv[0].n;
  check_vec(r, n, "r");
  check_vec(z, n, "z");
  LGPU_REQUIRE(r->ptr != z->ptr, "r and z must not alias");
  LGPU_REQUIRE(r->f32 == amg->f32 && z->f32 == amg->f32, "the vectors need the hierarchy's precision");
  if (x_in) {
    check_vec(x_in, n, "x_in");
    LGPU_REQUIRE(x_in->f32 == amg->f32, "x_in needs the hierarchy's precision");
    LGPU_REQUIRE(x_in->ptr != z->ptr, "x_in and z must not alias");
  }
  vcycle(amg, r->ptr, z->ptr, amg->finest, x_in ? x_in->ptr : nullptr);
  if (take_wait_error(amg->ctx)) return wait_timeout();
  API_END
}

extern "C" int lgpu_amg_vcycle(lgpu_amg amg, lgpu_vec r, lgpu_vec z) {
  API_BEGIN
  LGPU_REQUIRE(amg, "NULL AMG");
  if (amg->ssor_iters > 0) {
    check_vec(r, amg->finest->rows(), "r");
    check_vec(z, amg->finest->rows(), "z");
    LGPU_REQUIRE(!r->f32 && !z->f32, "the SSOR preconditioner works on f64 vectors");
    LGPU_REQUIRE(r->ptr != z->ptr, "r and z must not alias");
    precondition(amg, static_cast<const double*>(r->ptr), static_cast<double*>(z->ptr),
                 amg->finest);
    if (take_wait_error(amg->ctx)) return wait_timeout();
    return LGPU_OK;
  }
  check_vec(r, amg->lv[0].n, "r");
  check_vec(z, amg->lv[0].n, "z");
  LGPU_REQUIRE(r->ptr != z->ptr, "r and z must not alias");
  if (r->f32 || z->f32) {
    LGPU_REQUIRE(r->f32 && z->f32 && amg->f32, "f32 vectors need an f32 hierarchy");
    vcycle(amg, r->ptr, z->ptr, amg->finest);
  } else {
    precondition(amg, static_cast<const double*>(r->ptr), static_cast<double*>(z->ptr),
                 amg->finest);
  }
  if (take_wait_error(amg->ctx)) return wait_timeout();
  API_END
}

// ---------------------------------------------------------------------------
// Krylov solvers

namespace {

bool rere(int iteration, int every) { return every > 0 && iteration > 0 && iteration % every == 0; }

// The device state of a Krylov SystemInstance: ConjugateGradientsSolver's
// (conjugate-gradients/solver.hpp:50-121) or BiCgStabSolver's
// (bicgstab/solver.hpp:51-151).  Scalars (||r||^2, rho, <p, A p>, alpha,
// omega ...) stay in a device block; start() and each iteration are queued on
// the context's stream, and post()/wait() bring the block to the host
// together with the context's wait-error word.
struct KrylovSystem {
  lgpu_ctx ctx = nullptr;
  lgpu_op op = nullptr;
  lgpu_amg precon = nullptr;
  bool bicg = false;
  int64_t n = 0;
  int rere_every = 0;
  double* x = nullptr;
  const double* b = nullptr;
  double* w[6] = {};         // CG: r, p, q; BiCGStab: r, r0, p, v, t, px
  double* scal = nullptr;    // SC_COUNT device scalars
  double* host = nullptr;    // their pinned mirror
  hipEvent_t ev = nullptr;
  int it = 0, cur = SC_RHO0, nxt = SC_RHO1;
  bool head_queued = false;

  hipStream_t st() const { return ctx->stream; }
  bool mixed() const { return precon && precon->f32 && !bicg; }

  // SystemInstance constructor: r = b - A x with ||r||^2, then
  // CG: p = M r, rho = <p, r> (p = r, rho = ||r||^2 without a preconditioner);
  // BiCGStab: r0 = p = r, rho = ||r||^2
  void start() {
    double* r = w[0];
    residual_norm(op, b, x, r, scal + SC_RR);
    if (bicg) {
      LGPU_HIP_CHECK(hipMemcpyAsync(w[1], r, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, st()));
      LGPU_HIP_CHECK(hipMemcpyAsync(w[2], r, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, st()));
      LGPU_HIP_CHECK(hipMemcpyAsync(scal + cur, scal + SC_RR, sizeof(double), hipMemcpyDeviceToDevice, st()));
    } else if (precon) {
      precondition(precon, r, w[1], op);
      dot_into(ctx, w[1], r, false, n, scal + cur);
    } else {
      LGPU_HIP_CHECK(hipMemcpyAsync(w[1], r, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, st()));
      LGPU_HIP_CHECK(hipMemcpyAsync(scal + cur, scal + SC_RR, sizeof(double), hipMemcpyDeviceToDevice, st()));
    }
  }

  // the part of an iteration that reads p only (CG: q = A p and <p, q>;
  // BiCGStab: y = M p and v = A y), so that it can be queued before the
  // host has read the previous iteration's scalars
  void head() {
    if (bicg) {
      double *p = w[2], *v = w[3], *px = w[5];
      double* y = precon ? px : p;
      if (precon) precondition(precon, p, px, op);
      op_apply(op, SM_SPMV, false, y, nullptr, v, 0.0, nullptr);
    } else {
      const int64_t np = op_partials(op);
      double* part = ctx->partials_for(np);
      op_apply(op, SM_SPMV_DOT, false, w[1], nullptr, w[2], 0.0, part);
      reduce_partials(part, np, scal + SC_PQ, st());
    }
  }

  // the rest of iteration `it` (SystemInstance::iterate), then it + 1
  void rest() {
    if (bicg) rest_bicg();
    else rest_cg();
    std::swap(cur, nxt);
    ++it;
  }

  void rest_cg() {
    double *r = w[0], *p = w[1], *q = w[2];
    // FP32 cycle: the update also rounds r into the cycle's input, and the
    // dot and the direction update read the cycle's f32 output directly (it
    // widens exactly), so no conversion pass runs in between
    float* r32 = mixed() ? static_cast<float*>(precon->r32.ptr) : nullptr;
    bool r32_ready = false;
    if (rere(it, rere_every)) {
      axpy_scal(x, p, scal, cur, SC_PQ, 1.0, n, st());
      residual_norm(op, b, x, r, scal + SC_RR);
    } else {
      const int64_t nd = dot_partials(n);
      double* pd = ctx->partials_for(nd);
      cg_update(x, r, r32, p, q, scal, cur, SC_PQ, n, pd, st());
      reduce_partials(pd, nd, scal + SC_RR, st());
      r32_ready = mixed();
    }
    if (mixed()) {
      if (!r32_ready) vec_convert(r, false, r32, true, n, st());
      float* z32 = static_cast<float*>(precon->z32.ptr);
      vcycle(precon, r32, z32, op);
      const int64_t nd = dot_partials(n);
      double* pd = ctx->partials_for(nd);
      vec_dot_mixed(r, z32, n, pd, st());
      reduce_partials(pd, nd, scal + nxt, st());
      xpby_scal(p, z32, scal, nxt, cur, n, st());
    } else if (precon) {
      precondition(precon, r, q, op);
      dot_into(ctx, r, q, false, n, scal + nxt);
      xpby_scal(p, q, scal, nxt, cur, n, st());
    } else {
      LGPU_HIP_CHECK(hipMemcpyAsync(scal + nxt, scal + SC_RR, sizeof(double), hipMemcpyDeviceToDevice, st()));
      xpby_scal(p, r, scal, nxt, cur, n, st());
    }
  }

  void rest_bicg() {
    double *r = w[0], *r0 = w[1], *p = w[2], *v = w[3], *t = w[4], *px = w[5];
    double* y = precon ? px : p;
    const int64_t nd = dot_partials(n);
    dot_into(ctx, r0, v, false, n, scal + SC_R0V);
    axpy_scal(x, y, scal, cur, SC_R0V, 1.0, n, st());
    double* pd = ctx->partials_for(nd);
    bicg_r_update(r, v, scal, cur, SC_R0V, SC_ALPHA, n, pd, st());
    reduce_partials(pd, nd, scal + SC_RR, st());
    double* z = r;
    if (precon) {
      precondition(precon, r, px, op);
      z = px;
    }
    op_apply(op, SM_SPMV, false, z, nullptr, t, 0.0, nullptr);
    pd = ctx->partials_for(nd);
    vec_dot2(t, r, n, pd, st());
    reduce_partials2(pd, nd, scal + SC_TR, scal + SC_TT, st());
    axpy_scal(x, z, scal, SC_TR, SC_TT, 1.0, n, st());
    if (rere(it, rere_every)) {
      // omega still has to be stored for the direction update
      pd = ctx->partials_for(nd);
      bicg_r_update(t, t, scal, SC_TR, SC_TT, SC_OMEGA, 0, pd, st());
      residual_norm(op, b, x, r, scal + SC_RR);
    } else {
      pd = ctx->partials_for(nd);
      bicg_r_update(r, t, scal, SC_TR, SC_TT, SC_OMEGA, n, pd, st());
    }
    pd = ctx->partials_for(nd);
    vec_dot2(r, r0, n, pd, st());
    // (<r, r0>, ||r||^2)
    reduce_partials2(pd, nd, scal + nxt, scal + SC_RR, st());
    bicg_p_update(p, r, v, scal, nxt, cur, n, st());
  }

  // the scalar block (and the wait-error word) into the pinned mirror,
  // without waiting: work queued after this call overlaps the copy
  void post() {
    LGPU_HIP_CHECK(hipMemcpyAsync(host, scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, st()));
    if (scal != ctx->scal.as<double>())
      LGPU_HIP_CHECK(hipMemcpyAsync(host + SC_ERR, ctx->err(), sizeof(int), hipMemcpyDeviceToHost, st()));
    LGPU_HIP_CHECK(hipEventRecord(ev, st()));
  }
  void wait() const { LGPU_HIP_CHECK(hipEventSynchronize(ev)); }
  double rr() const { return host[SC_RR]; }
  double rho() const { return host[cur]; }
  bool err() const { return reinterpret_cast<const int*>(host + SC_ERR)[0] != 0; }
};

void check_krylov_args(lgpu_op op, lgpu_amg precon, lgpu_vec b, lgpu_vec x) {
  LGPU_REQUIRE(op, "NULL operator");
  const int64_t n = op->rows();
  check_vec(b, n, "b");
  check_vec(x, n, "x");
  LGPU_REQUIRE(!b->f32 && !x->f32, "the Krylov vectors are f64");
  LGPU_REQUIRE(x->ptr != b->ptr, "x and b must not alias");
  LGPU_REQUIRE(!precon || (precon->ctx == op->ctx && precon->finest->rows() == n),
               "the preconditioner must share the operator's context and size");
}

// End of a fused Krylov solve.  Every convergence check reads the wait-error
// word with the scalars, so an expiry in the work an iteration used has
// already stopped the loop (LGPU_SYNC_TIMEOUT).  Work queued after the last
// check is the next iteration's speculative head, whose results are
// discarded: an expiry there only clears the word.  A solve that never
// reached a check (a zero or non-finite initial residual) takes the word as
// it stands.
int finish(lgpu_ctx ctx, int status, bool checked) {
  const bool late = take_wait_error(ctx);
  if (status == LGPU_SYNC_TIMEOUT) return wait_timeout();
  if (late && !checked) return wait_timeout();
  return status;
}

// run_iterative_solver (linear-solver/runner.hpp:17-84) with the relative
// residual stop rule, on the context's shared work vectors and scalars.  The
// next iteration's head is queued before the host waits for this one's
// scalars, so the host's turnaround overlaps it.
int krylov_solve(bool bicg, lgpu_op op, lgpu_amg precon, lgpu_vec b, lgpu_vec x,
                 const lgpu_krylov_params* params, int* iterations, double* res_hist) {
  LGPU_REQUIRE(params, "NULL params");
  check_krylov_args(op, precon, b, x);
  lgpu_ctx ctx = op->ctx;
  KrylovSystem s;
  s.ctx = ctx;
  s.op = op;
  s.precon = precon;
  s.bicg = bicg;
  s.n = op->rows();
  s.rere_every = params->rere_every;
  s.x = static_cast<double*>(x->ptr);
  s.b = static_cast<const double*>(b->ptr);
  for (int k = 0; k < (bicg ? 6 : 3); ++k) s.w[k] = ctx->work_vec(k, s.n);
  s.scal = ctx->scal.as<double>();
  s.host = ctx->host_scal;
  s.ev = ctx->scal_ev;
  s.start();
  const double r0 = std::sqrt(read_scalar(ctx, SC_RR));
  if (res_hist) res_hist[0] = r0;
  int status = LGPU_NOT_CONVERGED;
  bool checked = false;
  if (!std::isfinite(r0)) status = LGPU_NONFINITE;
  else if (r0 == 0.0) status = LGPU_OK;
  while (status == LGPU_NOT_CONVERGED && s.it < params->max_it) {
    if (!s.head_queued) s.head();
    s.head_queued = false;
    s.rest();
    s.post();
    if (s.it < params->max_it) {
      s.head();
      s.head_queued = true;
    }
    s.wait();
    checked = true;
    const double rn = std::sqrt(s.rr());
    if (res_hist) res_hist[s.it] = rn;
    if (s.err()) status = LGPU_SYNC_TIMEOUT;
    else if (!std::isfinite(rn)) status = LGPU_NONFINITE;
    else if (rn <= params->rel_tol * r0) status = LGPU_OK;
    else if (s.rho() == 0.0) status = LGPU_BREAKDOWN;
  }
  if (iterations) *iterations = s.it;
  return finish(ctx, status, checked);
}

}  // namespace

// A SystemInstance on its own work vectors and scalar block, iterated one
// step per call (lgpu_krylov_iterate).
struct lgpu_krylov_s {
  KrylovSystem s;
  DevBuf vecs, scal;
  double* host = nullptr;
  hipEvent_t ev = nullptr;
  ~lgpu_krylov_s() {
    if (s.ctx) (void)hipStreamSynchronize(s.ctx->stream);
    if (host) (void)hipHostFree(host);
    if (ev) (void)hipEventDestroy(ev);
  }
};

extern "C" int lgpu_cg_solve(lgpu_op op, lgpu_amg precon, lgpu_vec b, lgpu_vec x,
                             const lgpu_krylov_params* params, int* iterations, double* res_hist) {
  API_BEGIN
  return krylov_solve(false, op, precon, b, x, params, iterations, res_hist);
  API_END
}

extern "C" int lgpu_bicgstab_solve(lgpu_op op, lgpu_amg precon, lgpu_vec b, lgpu_vec x,
                                   const lgpu_krylov_params* params, int* iterations,
                                   double* res_hist) {
  API_BEGIN
  return krylov_solve(true, op, precon, b, x, params, iterations, res_hist);
  API_END
}

namespace {

int krylov_create(bool bicg, lgpu_op op, lgpu_amg precon, lgpu_vec b, lgpu_vec x, int rere_every,
                  lgpu_krylov* out) {
  LGPU_REQUIRE(out, "NULL output");
  check_krylov_args(op, precon, b, x);
  LGPU_REQUIRE(rere_every >= 0, "rere_every must not be negative");
  auto k = std::make_unique<lgpu_krylov_s>();
  KrylovSystem& s = k->s;
  s.ctx = op->ctx;
  s.op = op;
  s.precon = precon;
  s.bicg = bicg;
  s.n = op->rows();
  s.rere_every = rere_every;
  s.x = static_cast<double*>(x->ptr);
  s.b = static_cast<const double*>(b->ptr);
  const int nv = bicg ? 6 : 3;
  k->vecs.alloc(size_t(nv) * size_t(s.n) * sizeof(double));
  for (int i = 0; i < nv; ++i) s.w[i] = k->vecs.as<double>() + size_t(i) * size_t(s.n);
  k->scal.alloc(SC_COUNT * sizeof(double));
  LGPU_HIP_CHECK(hipMemsetAsync(k->scal.ptr, 0, SC_COUNT * sizeof(double), s.ctx->stream));
  s.scal = k->scal.as<double>();
  LGPU_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&k->host), SC_COUNT * sizeof(double), hipHostMallocDefault));
  LGPU_HIP_CHECK(hipEventCreateWithFlags(&k->ev, hipEventDisableTiming));
  s.host = k->host;
  s.ev = k->ev;
  s.start();
  s.post();
  s.wait();
  if (s.err()) {
    take_wait_error(s.ctx);
    return wait_timeout();
  }
  *out = k.release();
  return LGPU_OK;
}

}  // namespace

extern "C" int lgpu_cg_create(lgpu_op op, lgpu_amg precon, lgpu_vec b, lgpu_vec x, int rere_every,
                              lgpu_krylov* out) {
  API_BEGIN
  return krylov_create(false, op, precon, b, x, rere_every, out);
  API_END
}

extern "C" int lgpu_bicgstab_create(lgpu_op op, lgpu_amg precon, lgpu_vec b, lgpu_vec x,
                                    int rere_every, lgpu_krylov* out) {
  API_BEGIN
  return krylov_create(true, op, precon, b, x, rere_every, out);
  API_END
}

extern "C" int lgpu_krylov_iterate(lgpu_krylov k) {
  API_BEGIN
  LGPU_REQUIRE(k, "NULL Krylov system");
  KrylovSystem& s = k->s;
  s.head();
  s.rest();
  s.post();
  s.wait();
  if (s.err()) {
    take_wait_error(s.ctx);
    return wait_timeout();
  }
  API_END
}

extern "C" int lgpu_krylov_residual_norm(lgpu_krylov k, double* out) {
  API_BEGIN
  LGPU_REQUIRE(k && out, "NULL argument");
  *out = std::sqrt(k->s.rr());
  API_END
}

extern "C" int lgpu_krylov_rho(lgpu_krylov k, double* out) {
  API_BEGIN
  LGPU_REQUIRE(k && out, "NULL argument");
  *out = k->s.rho();
  API_END
}

extern "C" int lgpu_krylov_iteration(lgpu_krylov k, int* out) {
  API_BEGIN
  LGPU_REQUIRE(k && out, "NULL argument");
  *out = k->s.it;
  API_END
}

extern "C" int lgpu_krylov_destroy(lgpu_krylov k) {
  API_BEGIN
  delete k;
  API_END
}

// ---------------------------------------------------------------------------
// adjoint of the solve: outlet flux and its gradient w.r.t. log coefficients

namespace {

void check_adjoint_op(lgpu_op op) {
  LGPU_REQUIRE(op, "NULL operator");
  LGPU_REQUIRE(!op->is_csr && op->desc.kind == 0,
               "the outlet-flux adjoint needs a per-cell coefficient stencil operator");
  LGPU_REQUIRE(op->desc.adv == 0.0, "the outlet-flux adjoint needs a symmetric operator (no advection)");
}

void check_f64(lgpu_vec v, int64_t n, const char* what) {
  check_vec(v, n, what);
  LGPU_REQUIRE(!v->f32, std::string(what) + " must be f64");
}

double outlet_flux_value(lgpu_op op, const double* p) {
  lgpu_ctx ctx = op->ctx;
  const int64_t np = outlet_flux_partials(op->desc);
  double* part = ctx->partials_for(np);
  outlet_flux(op->desc, p, part, ctx->stream);
  reduce_partials(part, np, ctx->scal.as<double>() + SC_FLUX, ctx->stream);
  return read_scalar(ctx, SC_FLUX);
}

}  // namespace

extern "C" int lgpu_outlet_flux(lgpu_op op, lgpu_vec p, double* flux) {
  API_BEGIN
  check_adjoint_op(op);
  LGPU_REQUIRE(flux, "NULL flux");
  check_f64(p, op->rows(), "p");
  *flux = outlet_flux_value(op, static_cast<const double*>(p->ptr));
  API_END
}

extern "C" int lgpu_outlet_flux_adjoint_rhs(lgpu_op op, lgpu_vec g) {
  API_BEGIN
  check_adjoint_op(op);
  check_f64(g, op->rows(), "g");
  outlet_flux_adjoint_rhs(op->desc, static_cast<double*>(g->ptr), op->ctx->stream);
  API_END
}

extern "C" int lgpu_outlet_flux_gradient(lgpu_op op, lgpu_vec p, lgpu_vec lambda, lgpu_vec grad) {
  API_BEGIN
  check_adjoint_op(op);
  const int64_t n = op->rows();
  check_f64(p, n, "p");
  check_f64(lambda, n, "lambda");
  check_f64(grad, n, "grad");
  LGPU_REQUIRE(grad->ptr != p->ptr && grad->ptr != lambda->ptr, "grad must not alias p or lambda");
  outlet_flux_gradient(op->desc, static_cast<const double*>(p->ptr),
                       static_cast<const double*>(lambda->ptr), static_cast<double*>(grad->ptr),
                       op->ctx->stream);
  API_END
}

// ---------------------------------------------------------------------------
// matrix and vector files

extern "C" int lgpu_smf_info(const char* path, int64_t* rows, int64_t* cols, int64_t* nnz) {
  API_BEGIN
  LGPU_REQUIRE(rows && cols && nnz, "NULL argument");
  const io::SmfHeader h = io::smf_header(path);
  *rows = int64_t(h.rows);
  *cols = int64_t(h.cols);
  *nnz = int64_t(h.nnz);
  API_END
}

extern "C" int lgpu_smf_read(lgpu_ctx ctx, const char* path, int dtype, int symmetric,
                             lgpu_op* out) {
  API_BEGIN
  LGPU_REQUIRE(ctx && out, "NULL argument");
  io::SmfHeader h;
  std::vector<int64_t> rowptr, cols;
  std::vector<double> vals;
  io::smf_read(path, h, rowptr, cols, vals);
  LGPU_REQUIRE(h.rows == h.cols, "only square SMF matrices are supported");
  LGPU_REQUIRE(h.rows > 0 && h.rows < (uint64_t(1) << 31), "SMF matrix size out of range");
  std::vector<int32_t> c32(cols.begin(), cols.end());
  const int st = lgpu_csr_create(ctx, int64_t(h.rows), rowptr.data(), c32.data(), vals.data(),
                                 dtype, symmetric, out);
  if (st != LGPU_OK) return st;
  API_END
}

extern "C" int lgpu_smf_write(lgpu_op op, const char* path) {
  API_BEGIN
  LGPU_REQUIRE(op, "NULL operator");
  if (op->is_csr) {
    const setup::RowMatrix& m = op->host_csr;
    std::vector<int64_t> cols(m.cols.begin(), m.cols.end());
    io::smf_write(path, uint64_t(m.n), uint64_t(m.n), m.rowptr.data(), cols.data(), m.vals.data(),
                  true);
  } else {
    // the row iterate's entries in column order (ordered_tag)
    setup::RowMatrix rows;
    stencil_host_rows(op, rows);
    std::vector<int64_t> cols(rows.cols.size());
    std::vector<double> vals(rows.vals.size());
    std::vector<std::pair<int64_t, double>> row;
    for (int64_t i = 0; i < rows.n; ++i) {
      row.clear();
      for (int64_t k = rows.rowptr[i]; k < rows.rowptr[i + 1]; ++k)
        row.emplace_back(rows.cols[size_t(k)], rows.vals[size_t(k)]);
      std::sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
      for (size_t e = 0; e < row.size(); ++e) {
        cols[size_t(rows.rowptr[i]) + e] = row[e].first;
        vals[size_t(rows.rowptr[i]) + e] = row[e].second;
      }
    }
    io::smf_write(path, uint64_t(rows.n), uint64_t(rows.n), rows.rowptr.data(), cols.data(),
                  vals.data(), false);
  }
  API_END
}

extern "C" int lgpu_dvf_size(const char* path, int64_t* n) {
  API_BEGIN
  LGPU_REQUIRE(n, "NULL argument");
  *n = int64_t(io::dvf_size(path));
  API_END
}

extern "C" int lgpu_dvf_read(lgpu_ctx ctx, const char* path, int dtype, lgpu_vec* out) {
  API_BEGIN
  LGPU_REQUIRE(ctx && out, "NULL argument");
  LGPU_REQUIRE(dtype == LGPU_F64 || dtype == LGPU_F32, "unknown dtype");
  std::vector<double> v;
  io::dvf_read(path, v);
  lgpu_vec h = nullptr;
  int st = lgpu_vec_create(ctx, int64_t(v.size()), dtype, &h);
  if (st != LGPU_OK) return st;
  std::unique_ptr<lgpu_vec_s> owned(h);
  if (dtype == LGPU_F32) {
    std::vector<float> f(v.begin(), v.end());
    st = lgpu_vec_upload(h, f.data(), int64_t(f.size()));
  } else {
    st = lgpu_vec_upload(h, v.data(), int64_t(v.size()));
  }
  if (st != LGPU_OK) return st;
  *out = owned.release();
  API_END
}

extern "C" int lgpu_dvf_write(lgpu_vec v, const char* path) {
  API_BEGIN
  LGPU_REQUIRE(v, "NULL vector");
  std::vector<double> host(size_t(v->n));
  if (v->f32) {
    std::vector<float> f(size_t(v->n));
    LGPU_HIP_CHECK(hipMemcpyAsync(f.data(), v->ptr, f.size() * sizeof(float), hipMemcpyDeviceToHost, v->ctx->stream));
    LGPU_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
    std::copy(f.begin(), f.end(), host.begin());
  } else {
    LGPU_HIP_CHECK(hipMemcpyAsync(host.data(), v->ptr, host.size() * sizeof(double), hipMemcpyDeviceToHost, v->ctx->stream));
    LGPU_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
  }
  io::dvf_write(path, host.data(), uint64_t(host.size()));
  API_END
}

namespace {

void check_eff_diff_op(lgpu_op op) {
  LGPU_REQUIRE(op && !op->is_csr, "the effective diffusivity needs a stencil operator");
  // the plane x = nx / 2 needs a +x neighbour plane
  LGPU_REQUIRE(op->desc.nx >= 3, "the effective diffusivity needs at least three cells along x");
}

// the valuators' filled and diffusion-factor lookups
// (tensor/dynamic/stencil/valuator/lookup-wright.hpp:91-106)
EffDiffTables eff_diff_tables(lgpu_op op, const uint8_t* none_mats) {
  EffDiffTables t{};
  for (int m = 0; m < 4; ++m) t.filled[m] = none_mats && none_mats[m] ? 0.0 : 1.0;
  for (int m = 0; m < 16; ++m) t.factor[m] = host_hmean(op->desc.mc[m / 4], op->desc.mc[m % 4]);
  return t;
}

double eff_diff_reduce(lgpu_op op, const std::function<void(double*)>& launch) {
  lgpu_ctx ctx = op->ctx;
  const int64_t np = eff_diff_partials(op->desc);
  double* part = ctx->partials_for(np);
  launch(part);
  reduce_partials(part, np, ctx->scal.as<double>() + SC_EFFDIFF, ctx->stream);
  return read_scalar(ctx, SC_EFFDIFF);
}

}  // namespace

extern "C" int lgpu_eff_diff_area_part(lgpu_op op, const uint8_t* none_mats, double* area_part) {
  API_BEGIN
  check_eff_diff_op(op);
  LGPU_REQUIRE(area_part, "NULL area_part");
  const EffDiffTables t = eff_diff_tables(op, none_mats);
  const double filled = eff_diff_reduce(
      op, [&](double* part) { eff_diff_filled(op->desc, t, part, op->ctx->stream); });
  // effective-diffusion.hpp:113-116
  *area_part = filled / double(op->desc.ny * op->desc.nz) * 0.5;
  API_END
}

extern "C" int lgpu_eff_diff(lgpu_op op, lgpu_op map, lgpu_vec sol, double area_part,
                             double* eff_diff) {
  API_BEGIN
  check_eff_diff_op(op);
  LGPU_REQUIRE(eff_diff, "NULL eff_diff");
  const int64_t* cmap = nullptr;
  int64_t n = op->rows();
  if (map) {
    LGPU_REQUIRE(map->is_csr && map->cmap.ptr && map->cmap_cells == n,
                 "map must be a compressed system of this stencil's size");
    cmap = map->cmap.as<int64_t>();
    n = map->rows();
  }
  check_vec(sol, n, "sol");
  const EffDiffTables t = eff_diff_tables(op, nullptr);
  const double sum = eff_diff_reduce(op, [&](double* part) {
    eff_diff_flux(op->desc, t, sol->ptr, sol->f32, cmap, part, op->ctx->stream);
  });
  // effective-diffusion.hpp:320-326
  const lgpu_box& b = op->box;
  double out = sum;
  out *= op->desc.q[0];
  out /= area_part * ((1.0 * b.len[1]) * b.len[2]);
  out *= b.len[0];
  *eff_diff = out / (b.p_end - b.p_start);
  API_END
}

extern "C" int lgpu_adjoint_grad(lgpu_op op, lgpu_amg precon, const lgpu_krylov_params* params,
                                 lgpu_vec p, lgpu_vec lambda, lgpu_vec grad, double* flux,
                                 int* iterations) {
  API_BEGIN
  check_adjoint_op(op);
  LGPU_REQUIRE(params, "NULL params");
  const int64_t n = op->rows();
  check_f64(p, n, "p");
  check_f64(lambda, n, "lambda");
  check_f64(grad, n, "grad");
  LGPU_REQUIRE(p->ptr != lambda->ptr && grad->ptr != p->ptr && grad->ptr != lambda->ptr,
               "p, lambda and grad must not alias");
  lgpu_ctx ctx = op->ctx;
  lgpu_vec_s rhs_v{};
  // the right-hand sides live in a work vector the CG solves do not use
  double* rhs = ctx->work_vec(6, n);
  rhs_v.ctx = ctx;
  rhs_v.n = n;
  rhs_v.f32 = false;
  rhs_v.ptr = rhs;
  int it[2] = {0, 0};
  stencil_rhs(op->desc, false, rhs, ctx->stream);
  int status = lgpu_cg_solve(op, precon, &rhs_v, p, params, &it[0], nullptr);
  if (status == LGPU_OK || status == LGPU_NOT_CONVERGED) {
    outlet_flux_adjoint_rhs(op->desc, rhs, ctx->stream);
    const int s2 = lgpu_cg_solve(op, precon, &rhs_v, lambda, params, &it[1], nullptr);
    if (status == LGPU_OK) status = s2;
    outlet_flux_gradient(op->desc, static_cast<const double*>(p->ptr),
                         static_cast<const double*>(lambda->ptr), static_cast<double*>(grad->ptr),
                         ctx->stream);
    if (flux) *flux = outlet_flux_value(op, static_cast<const double*>(p->ptr));
  }
  if (iterations) {
    iterations[0] = it[0];
    iterations[1] = it[1];
  }
  return status;
  API_END
}
