// Host-side AMG setup (see setup.hpp).  The strength criterion and the Galerkin
// product are independent per vertex / per coarse row and run on host threads;
// the greedy aggregation runs with one thread segment, as the reference does
// with one thread, so the hierarchy does not depend on the host's core count.
#include "setup.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>

namespace lgpu::setup {

namespace {

template <typename F>
void parallel_for(int64_t n, int threads, F&& f) {
  if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
  if (n < 4096 || threads == 1) {
    f(int64_t(0), n);
    return;
  }
  threads = int(std::min<int64_t>(threads, (n + 4095) / 4096));
  std::vector<std::thread> pool;
  pool.reserve(size_t(threads));
  for (int t = 0; t < threads; ++t) {
    const int64_t b = n * t / threads, e = n * (t + 1) / threads;
    pool.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& th : pool) th.join();
}

// Matrix graph (multigrid/graph/matrix-graph.hpp:63-190): the off-diagonal
// entries of each row in ascending column order are its edges.
struct Graph {
  int64_t n = 0;
  std::vector<int64_t> eptr;
  std::vector<int32_t> head;
  std::vector<double> a;     // a_ij
  std::vector<double> diag;  // a_ii
  std::vector<int64_t> rev;  // reverse edge or -1
  std::vector<uint8_t> depends, influences, isolated;
  bool symmetric = true, f32 = false;
};

// The edges (off-diagonal entries in row order) without values or reverse
// edges: what the aggregation reads when the criterion ran elsewhere.
void build_edges(const RowMatrix& m, Graph& g, int threads) {
  g.n = m.n;
  g.symmetric = m.symmetric;
  g.f32 = m.f32;
  g.eptr.assign(size_t(m.n + 1), 0);
  for (int64_t i = 0; i < m.n; ++i) g.eptr[i + 1] = g.eptr[i] + (m.rowptr[i + 1] - m.rowptr[i] - 1);
  g.head.resize(size_t(g.eptr[m.n]));
  parallel_for(m.n, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      int64_t k = g.eptr[i];
      for (int64_t p = m.rowptr[i]; p < m.rowptr[i + 1]; ++p)
        if (m.cols[p] != i) g.head[k++] = m.cols[p];
    }
  });
}

void build_graph(const RowMatrix& m, Graph& g, int threads) {
  g.n = m.n;
  g.symmetric = m.symmetric;
  g.f32 = m.f32;
  g.eptr.assign(size_t(m.n + 1), 0);
  g.diag.assign(size_t(m.n), 0.0);
  for (int64_t i = 0; i < m.n; ++i) g.eptr[i + 1] = g.eptr[i] + (m.rowptr[i + 1] - m.rowptr[i] - 1);
  const int64_t ne = g.eptr[m.n];
  g.head.resize(size_t(ne));
  g.a.resize(size_t(ne));
  g.rev.resize(size_t(ne));
  parallel_for(m.n, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      int64_t k = g.eptr[i];
      for (int64_t p = m.rowptr[i]; p < m.rowptr[i + 1]; ++p) {
        if (m.cols[p] == i) {
          g.diag[i] = m.vals[p];
        } else {
          g.head[k] = m.cols[p];
          g.a[k++] = m.vals[p];
        }
      }
    }
  });
  parallel_for(m.n, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i)
      for (int64_t k = g.eptr[i]; k < g.eptr[i + 1]; ++k) {
        const int64_t j = g.head[k];
        const auto first = g.head.begin() + g.eptr[j], last = g.head.begin() + g.eptr[j + 1];
        const auto it = std::lower_bound(first, last, int32_t(i));
        g.rev[k] = (it != last && *it == i) ? int64_t(it - g.head.begin()) : -1;
      }
  });
}

// DuneDependencyCriterion (multigrid/dependency/criterion.hpp:35-78) with the
// NegativeOffdiagonalTransform (graph/weight-transform.hpp:28-39), evaluated
// in the level's value type W.
template <typename W>
void criterion_rows(Graph& g, W theta, int64_t b, int64_t e) {
  std::vector<W> strength;
  for (int64_t i = b; i < e; ++i) {
    const int64_t k0 = g.eptr[i], k1 = g.eptr[i + 1];
    strength.resize(size_t(k1 - k0));
    W mx = W(0);
    for (int64_t k = k0; k < k1; ++k) {
      const W w = -W(g.a[k]);
      const W nn = std::max(w, W(0));
      const W head_weight = W(g.diag[g.head[k]]);
      W opposite;
      if (g.symmetric) opposite = nn;
      else opposite = g.rev[k] >= 0 ? std::max(-W(g.a[g.rev[k]]), W(0)) : W(0);
      const W s = nn * opposite / head_weight;
      mx = std::max(mx, s);
      strength[size_t(k - k0)] = s;
    }
    const W lb = theta * mx;
    bool any = false;
    for (int64_t k = k0; k < k1; ++k) {
      const bool strong = strength[size_t(k - k0)] > lb;
      any = any || strong;
      g.depends[k] = strong;
      if (g.rev[k] >= 0) g.influences[g.rev[k]] = strong;
    }
    g.isolated[i] = !any;
  }
}

void criterion(Graph& g, double theta, int threads) {
  g.depends.assign(g.head.size(), 0);
  g.influences.assign(g.head.size(), 0);
  g.isolated.assign(size_t(g.n), 0);
  parallel_for(g.n, threads, [&](int64_t b, int64_t e) {
    if (g.f32) criterion_rows<float>(g, float(theta), b, e);
    else criterion_rows<double>(g, theta, b, e);
  });
}

// Greedy aggregation of one level with a single thread segment
// (multigrid/aggregation/base.hpp:448-707, aggregate-map/multithread.hpp:145-214).
class Aggregator {
 public:
  Aggregator(const Graph& g, int64_t min_size, int64_t max_size, bool aggressive)
      : g_(g), min_(min_size), max_(max_size), aggressive_(aggressive),
        agg_(size_t(g.n), kUnaggregated), in_front_(size_t(g.n), 0), conn_(size_t(g.n) + 1, 0),
        kind_(g.head.size()) {
    // bit 0: the edge's head is a strong dependency; bit 1: it is influenced
    for (size_t e = 0; e < kind_.size(); ++e) kind_[e] = uint8_t(g.depends[e] | (g.influences[e] << 1));
  }

  Level run() {
    for (int64_t v = 0; v < g_.n; ++v) handle_vertex(v);
    Level lv;
    const int64_t nagg = int64_t(size_.size());
    lv.c2f_ptr.assign(size_t(nagg + 1), 0);
    for (int64_t c = 0; c < nagg; ++c) lv.c2f_ptr[c + 1] = lv.c2f_ptr[c] + size_[c];
    // the log in order, placed per aggregate: members in insertion order
    lv.c2f_idx.resize(size_t(lv.c2f_ptr[nagg]));
    std::vector<int64_t> pos(lv.c2f_ptr.begin(), lv.c2f_ptr.end() - 1);
    for (size_t k = 0; k < log_agg_.size(); ++k) lv.c2f_idx[pos[log_agg_[k]]++] = log_vertex_[k];
    stats_.aggregates = nagg;
    lv.stats = stats_;
    lv.agg = std::move(agg_);
    return lv;
  }

 private:
  bool is_aggregate(uint32_t a) const { return a < kIsolated; }
  // membership of the aggregate's connected set: conn_[a] holds the stamp of
  // the last vertex whose aggregate reached a
  bool connected(uint32_t a) const { return conn_[a] == stamp_; }
  bool two_way(int64_t e) const { return kind_[e] == 3; }
  bool one_way(int64_t e) const { return kind_[e] == 1; }
  bool strong(int64_t e) const { return g_.depends[e] || g_.influences[e]; }

  void front_insert(int32_t v) {
    front_.insert(std::lower_bound(front_.begin(), front_.end(), v), v);
    in_front_[v] = 1;
  }
  void front_erase(int32_t v) {
    const auto it = std::lower_bound(front_.begin(), front_.end(), v);
    if (it != front_.end() && *it == v) front_.erase(it);
    in_front_[v] = 0;
  }

  // AggregateManager::add_help: store, then scan every neighbour
  void add(int32_t u, uint32_t cur) {
    agg_[u] = cur;
    open_.push_back(u);
    for (int64_t e = g_.eptr[u]; e < g_.eptr[u + 1]; ++e) {
      const int32_t t = g_.head[e];
      const uint32_t ta = agg_[t];
      if (is_aggregate(ta)) conn_[ta] = stamp_;
      if (ta == kUnaggregated && !in_front_[t]) front_insert(t);
    }
  }

  void add_all(uint32_t cur) {
    for (int32_t v : cand_) {
      front_erase(v);
      add(v, cur);
    }
  }

  void grow(uint32_t cur) {
    while (int64_t(open_.size()) < min_) {
      int64_t best0 = 0, best2 = 0;
      double best1 = 0.0;
      bool has_two_way = false;
      cand_.clear();
      for (int32_t v : front_) {
        if (g_.isolated[v]) continue;
        int64_t con = 0, nbs = 0, cons2 = 0, cons1 = 0;
        for (int64_t e = g_.eptr[v]; e < g_.eptr[v + 1]; ++e) {
          const uint32_t ha = agg_[g_.head[e]];
          // conn_[n] is never stamped: non-aggregates count as unconnected
          con += conn_[std::min<uint32_t>(ha, uint32_t(g_.n))] == stamp_;
          nbs += in_front_[g_.head[e]];
          cons2 += ha == cur && two_way(e);
          cons1 += ha == cur && one_way(e);
        }
        const double connectivity = double(con) / double(g_.eptr[v + 1] - g_.eptr[v]);
        if (cons2 > 0 && !has_two_way) {
          has_two_way = true;
          best0 = cons2;
          best1 = connectivity;
          best2 = nbs;
          cand_.assign(1, v);
          continue;
        }
        const int64_t c0 = has_two_way ? cons2 : cons1;
        if (c0 <= 0) continue;
        // lexicographic comparison of (count, connectivity, front neighbours)
        int cmp;
        if (c0 != best0) cmp = c0 > best0 ? 1 : -1;
        else if (connectivity != best1) cmp = connectivity > best1 ? 1 : (connectivity < best1 ? -1 : 2);
        else cmp = nbs > best2 ? 1 : (nbs < best2 ? -1 : 0);
        if (cmp == 1) {
          best0 = c0;
          best1 = connectivity;
          best2 = nbs;
          cand_.assign(1, v);
        } else if (cmp == 0) {
          cand_.push_back(v);
        }
      }
      if (cand_.empty()) break;
      const size_t room = size_t(max_ - int64_t(open_.size()));
      if (cand_.size() > room) cand_.resize(room);
      add_all(cur);
    }
  }

  void round(uint32_t cur) {
    while (int64_t(open_.size()) < max_) {
      cand_.clear();
      for (int32_t v : front_) {
        if (g_.isolated[v]) continue;
        int64_t cons = 0, unagg = 0, inagg = 0;
        for (int64_t e = g_.eptr[v]; e < g_.eptr[v + 1]; ++e) {
          const uint32_t ha = agg_[g_.head[e]];
          cons += ha == cur && g_.depends[e];
          unagg += ha == kUnaggregated;
          inagg += ha == cur;
        }
        if (cons == 0 || unagg >= inagg) continue;
        cand_.push_back(v);
        break;
      }
      if (cand_.empty()) break;
      add_all(cur);
    }
  }

  void handle_vertex(int64_t v) {
    if (agg_[v] != kUnaggregated) return;
    if (g_.isolated[v]) {
      agg_[v] = kIsolated;
      ++stats_.isolated;
      return;
    }
    const uint32_t cur = uint32_t(size_.size());  // the open aggregate's index
    open_.clear();
    front_.clear();
    stamp_ = uint32_t(v) + 1;  // a fresh, empty connected set
    add(int32_t(v), cur);
    grow(cur);
    round(cur);
    for (int32_t f : front_) in_front_[f] = 0;
    front_.clear();
    if (open_.size() == 1 && max_ > 1) {
      // find_merge_aggregate: first strong neighbour with room in its aggregate
      // (a closed one: only v itself is in the open aggregate)
      uint32_t target = kUnaggregated;
      for (int64_t e = g_.eptr[v]; e < g_.eptr[v + 1]; ++e) {
        if (!strong(e)) continue;
        const uint32_t ha = agg_[g_.head[e]];
        if (is_aggregate(ha) && size_[ha] < max_) {
          target = ha;
          break;
        }
      }
      if (target != kUnaggregated) {
        agg_[v] = target;
        ++size_[target];
        log_agg_.push_back(target);
        log_vertex_.push_back(int32_t(v));
        ++stats_.merged;
        return;
      }
      if (aggressive_) {
        agg_[v] = kIsolated;
        ++stats_.isolated;
        return;
      }
      ++stats_.single;
    } else {
      ++stats_.multi;
    }
    // finalize: close the open aggregate
    size_.push_back(int64_t(open_.size()));
    for (int32_t u : open_) {
      log_agg_.push_back(cur);
      log_vertex_.push_back(u);
    }
  }

  const Graph& g_;
  int64_t min_, max_;
  bool aggressive_;
  std::vector<uint32_t> agg_;
  std::vector<uint8_t> in_front_;
  std::vector<int32_t> front_;
  std::vector<uint32_t> conn_;
  std::vector<uint8_t> kind_;
  uint32_t stamp_ = 0;
  std::vector<int32_t> cand_;
  std::vector<int32_t> open_;        // members of the open aggregate, in order
  std::vector<int64_t> size_;        // member count of each closed aggregate
  std::vector<uint32_t> log_agg_;    // every membership in insertion order:
  std::vector<int32_t> log_vertex_;  // aggregate and vertex
  AggregationStats stats_;
};

// Galerkin product (transport/matrix-shared.hpp:188-202 and the non-unique
// row builder of csr/shared.hpp:733-764).
RowMatrix galerkin(const RowMatrix& fine, const Level& lv, bool coarse_f32, int threads) {
  const int64_t nc = int64_t(lv.c2f_ptr.size()) - 1;
  RowMatrix out;
  out.n = nc;
  out.f32 = coarse_f32;
  out.symmetric = fine.symmetric;
  int nthreads = threads <= 0 ? int(std::max(1u, std::thread::hardware_concurrency())) : threads;
  if (nc < 4096) nthreads = 1;
  struct Part {
    std::vector<int64_t> len;
    std::vector<int32_t> cols;
    std::vector<double> vals;
  };
  std::vector<Part> parts(static_cast<size_t>(nthreads));
  auto work = [&](int t) {
    const int64_t b = nc * t / nthreads, e = nc * (t + 1) / nthreads;
    Part& part = parts[size_t(t)];
    std::vector<std::pair<int32_t, double>> row;
    for (int64_t c = b; c < e; ++c) {
      row.clear();
      for (int64_t k = lv.c2f_ptr[c]; k < lv.c2f_ptr[c + 1]; ++k) {
        const int32_t v = lv.c2f_idx[k];
        for (int64_t p = fine.rowptr[v]; p < fine.rowptr[v + 1]; ++p) {
          const uint32_t J = lv.agg[fine.cols[p]];
          if (J >= kIsolated) continue;
          const double val = coarse_f32 ? double(float(fine.vals[p])) : fine.vals[p];
          auto it = std::find_if(row.begin(), row.end(),
                                 [J](const auto& q) { return q.first == int32_t(J); });
          if (it != row.end()) it->second += val;
          else row.emplace_back(int32_t(J), val);
        }
      }
      std::sort(row.begin(), row.end(),
                [](const auto& l, const auto& r) { return l.first < r.first; });
      part.len.push_back(int64_t(row.size()));
      for (const auto& [j, val] : row) {
        part.cols.push_back(j);
        part.vals.push_back(coarse_f32 ? double(float(val)) : val);
      }
    }
  };
  if (nthreads == 1) {
    work(0);
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t) pool.emplace_back(work, t);
    for (auto& th : pool) th.join();
  }
  out.rowptr.assign(1, 0);
  out.rowptr.reserve(size_t(nc + 1));
  for (const Part& part : parts) {
    for (int64_t l : part.len) out.rowptr.push_back(out.rowptr.back() + l);
    out.cols.insert(out.cols.end(), part.cols.begin(), part.cols.end());
    out.vals.insert(out.vals.end(), part.vals.begin(), part.vals.end());
  }
  return out;
}

}  // namespace

// Cholesky-Banachiewicz (linear-solver/direct/cholesky/decomposition.hpp:44-99)
bool cholesky(const RowMatrix& a, RowMatrix& lower) {
  const int64_t n = a.n;
  lower = RowMatrix{};
  lower.n = n;
  lower.rowptr.assign(1, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t row_start = int64_t(lower.cols.size());
    double diag = 0.0;
    auto entry = [&](int64_t j, double base) {
      const int64_t b = lower.rowptr[j], e = lower.rowptr[j + 1];
      double dot = 0.0;
      int64_t it = b;
      for (int64_t k = row_start; k < int64_t(lower.cols.size()); ++k) {
        while (it != e && lower.cols[it] < lower.cols[k]) ++it;
        if (it == e) break;
        if (lower.cols[it] == lower.cols[k]) dot += lower.vals[k] * lower.vals[it];
      }
      const double value = (base - dot) / lower.vals[e - 1];
      if (value != 0.0) {
        diag -= value * value;
        lower.cols.push_back(int32_t(j));
        lower.vals.push_back(value);
      }
    };
    int64_t next = 0;
    bool first = true;
    for (int64_t p = a.rowptr[i]; p < a.rowptr[i + 1]; ++p) {
      const int64_t j = a.cols[p];
      if (j < i) {
        if (first) first = false;
        else
          for (int64_t g = next; g < j; ++g) entry(g, 0.0);
        entry(j, a.vals[p]);
        next = j + 1;
      } else if (j == i) {
        diag += a.vals[p];
      }
    }
    for (int64_t g = next; g < i; ++g) entry(g, 0.0);
    if (!(diag > 0.0)) return false;
    lower.cols.push_back(int32_t(i));
    lower.vals.push_back(std::sqrt(diag));
    lower.rowptr.push_back(int64_t(lower.cols.size()));
  }
  return true;
}

// Row i starts from its own entries; each finished row k < i that holds a
// (possibly filled-in) entry eliminates it in ascending k, l_ik = w_k / u_kk,
// updating the entries of its upper structure (w_j -= l_ik u_kj, or
// w_j = -(l_ik u_kj) as fill-in); then u_ii = w_i, u_ij = w_j (j > i), l_ii = 1.
bool lu_dense(const RowMatrix& a, std::vector<double>& lower, std::vector<double>& upper) {
  const int64_t n = a.n;
  const size_t nn = size_t(n * n);
  lower.assign(nn, 0.0);
  upper.assign(nn, 0.0);
  std::vector<double> w(static_cast<size_t>(n));
  std::vector<char> present(static_cast<size_t>(n)), ustruct(nn, 0);
  for (int64_t i = 0; i < n; ++i) {
    std::fill(w.begin(), w.end(), 0.0);
    std::fill(present.begin(), present.end(), 0);
    for (int64_t p = a.rowptr[size_t(i)]; p < a.rowptr[size_t(i + 1)]; ++p) {
      w[size_t(a.cols[size_t(p)])] = a.vals[size_t(p)];
      present[size_t(a.cols[size_t(p)])] = 1;
    }
    for (int64_t k = 0; k < i; ++k) {
      if (!present[size_t(k)]) continue;
      const double l = w[size_t(k)] / upper[size_t(k * n + k)];
      lower[size_t(i * n + k)] = l;
      for (int64_t j = k + 1; j < n; ++j) {
        if (!ustruct[size_t(k * n + j)]) continue;
        const double sub = l * upper[size_t(k * n + j)];
        if (present[size_t(j)]) {
          w[size_t(j)] -= sub;
        } else {
          w[size_t(j)] = -sub;
          present[size_t(j)] = 1;
        }
      }
    }
    if (!present[size_t(i)] || w[size_t(i)] == 0.0) return false;
    lower[size_t(i * n + i)] = 1.0;
    upper[size_t(i * n + i)] = w[size_t(i)];
    ustruct[size_t(i * n + i)] = 1;
    for (int64_t j = i + 1; j < n; ++j) {
      if (!present[size_t(j)]) continue;
      upper[size_t(i * n + j)] = w[size_t(j)];
      ustruct[size_t(i * n + j)] = 1;
    }
  }
  return true;
}

// HierarchyWrightInstance::create_impl (multigrid/hierarchy/wright.hpp:120-143, 145-309)
Hierarchy build(const RowMatrix& finest, const Params& p, Accel* accel) {
  using Clock = std::chrono::steady_clock;
  auto since = [](Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
  };
  Hierarchy h;
  h.sizes.push_back(finest.n);
  bool aggressive = false;
  const RowMatrix* fine = &finest;
  h.coarse.reserve(size_t(p.max_level_num) + 1);
  for (int level = 0;; ++level) {
    auto t0 = Clock::now();
    Graph g;
    const double theta = aggressive ? p.aggressive_strong_threshold : p.strong_threshold;
    if (accel) {
      build_edges(*fine, g, p.threads);
      StrengthFlags f;
      accel->criterion(*fine, theta, f);
      g.depends = std::move(f.depends);
      g.influences = std::move(f.influences);
      g.isolated = std::move(f.isolated);
    } else {
      build_graph(*fine, g, p.threads);
      criterion(g, theta, p.threads);
    }
    h.seconds_criterion += since(t0);
    t0 = Clock::now();
    Aggregator aggr(g, p.min_aggregate_size, p.max_aggregate_size, aggressive);
    h.transfer.push_back(aggr.run());
    h.seconds_aggregation += since(t0);
    t0 = Clock::now();
    h.coarse.push_back(accel ? accel->galerkin(*fine, h.transfer.back(), p.coarse_f32)
                             : galerkin(*fine, h.transfer.back(), p.coarse_f32, p.threads));
    h.seconds_galerkin += since(t0);
    const int64_t nc = h.coarse.back().n;
    h.sizes.push_back(nc);
    bool stop = false;
    if (nc < p.min_coarsen_size) {
      stop = true;
    } else if (level + 1 >= p.max_level_num) {
      stop = true;
    } else if (p.min_coarsen_factor * double(nc) > double(h.sizes[size_t(level)])) {
      if (aggressive) stop = true;
      else aggressive = true;
    }
    if (stop) break;
    fine = &h.coarse.back();
  }
  const auto tc = Clock::now();
  factor_coarsest(h);
  h.seconds_coarsest = since(tc);
  return h;
}

// The coarsest solver's factorisation: Cholesky for a symmetric coarsest
// matrix, row-wise LU for a nonsymmetric one (BiCgStab's hierarchies,
// test/aux/iterative.hpp:122-181)
void factor_coarsest(Hierarchy& h) {
  if (!h.coarse.back().symmetric) {
    h.lu = true;
    h.spd = lu_dense(h.coarse.back(), h.chol_dense, h.upper_dense);
    return;
  }
  h.lu = false;
  RowMatrix lower;
  h.spd = cholesky(h.coarse.back(), lower);
  const int64_t n = h.coarse.back().n;
  h.chol_dense.assign(size_t(n * n), 0.0);
  if (h.spd)
    for (int64_t i = 0; i < n; ++i)
      for (int64_t k = lower.rowptr[i]; k < lower.rowptr[i + 1]; ++k)
        h.chol_dense[size_t(i * n + lower.cols[k])] = lower.vals[k];
}

}  // namespace lgpu::setup
