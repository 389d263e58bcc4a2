// f2m_oracle.cpp -- CPU restatement of the reference F2M hot path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the f2m_amd HIP library and the
// CPU baseline of bench.py.  The product never links or calls this file.
//
// Build: oracle/Makefile (g++ -O3 -ffp-contract=off, std::thread).
// The arithmetic is written so that a GPU implementation that evaluates the
// same expressions in the same order produces identical bits:
//   d_e = (D_e + 0.5*F[lo]) + 0.5*F[hi],  lo = min(u,w), hi = max(u,w)
//   g_v = -(t2 + t3) / 2   (t1 <= t2 <= t3: three smallest d over the list,
//                            padded with 1e6 sentinels)
//   F_v <- F_v + g_v        (all nodes from the same old F: Jacobi)
// 0.5*F and /2 are exact in binary64, so an FMA-contracted D + 0.5*F rounds
// identically; contraction is disabled anyway.
#include "f2m_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Persistent worker pool: run(f) calls f(tid, nthreads) on every worker and
// returns when all are done.  Workers spin between jobs (the dual descent runs
// thousands of tiny parallel steps, so thread start-up per step is too slow).
// ---------------------------------------------------------------------------
class Pool {
 public:
  explicit Pool(int n) : n_(n < 1 ? 1 : n) {
    for (int t = 1; t < n_; ++t) th_.emplace_back([this, t] { loop(t); });
  }
  ~Pool() {
    stop_.store(true);
    gen_.fetch_add(1);
    for (auto& t : th_) t.join();
  }
  int size() const { return n_; }
  void run(const std::function<void(int, int)>& f) {
    if (n_ == 1) { f(0, 1); return; }
    job_ = &f;
    done_.store(0);
    gen_.fetch_add(1, std::memory_order_acq_rel);
    f(0, n_);
    while (done_.load(std::memory_order_acquire) != n_ - 1) spin();
  }
  template <class B>
  void parallel_for(int64_t n, B body) {  // body(lo, hi)
    std::function<void(int, int)> f = [&](int t, int nt) {
      int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
      if (lo < hi) body(lo, hi);
    };
    run(f);
  }

 private:
  static void spin() {
#if defined(__x86_64__)
    __builtin_ia32_pause();
#endif
  }
  void loop(int t) {
    uint64_t seen = 0;
    for (;;) {
      uint64_t g;
      int k = 0;
      while ((g = gen_.load(std::memory_order_acquire)) == seen) {
        if (++k > 4096) { std::this_thread::yield(); k = 0; } else spin();
      }
      seen = g;
      if (stop_.load()) return;
      (*job_)(t, n_);
      done_.fetch_add(1, std::memory_order_acq_rel);
    }
  }
  int n_;
  std::vector<std::thread> th_;
  const std::function<void(int, int)>* job_ = nullptr;
  std::atomic<uint64_t> gen_{0};
  std::atomic<int> done_{0};
  std::atomic<bool> stop_{false};
};

// Edge cost on 2D points (FXY.hpp:376-383: sqrt(dY*dY + dX*dX) in double).
inline double cost_of(double xu, double yu, double xw, double yw, int mode) {
  double dx = xu - xw, dy = yu - yw;
  double s = dx * dx + dy * dy;
  double c = std::sqrt(s);
  if (mode == ORC_COST_NINT) return std::floor(c + 0.5);
  if (mode == ORC_COST_RAW_FP32) return (double)(float)c;
  return c;
}
// The same rounding applied to a distance lower bound (monotone in dist).
inline double round_like_cost(double dist, int mode) {
  if (mode == ORC_COST_NINT) return std::floor(dist + 0.5);
  if (mode == ORC_COST_RAW_FP32) return (double)(float)dist;
  return dist;
}

struct Key {  // kNN order: (cost, id) -- restates the global (d, min id, max id)
  double c;   // order of the sorted candidate list (FXY.hpp:100-101) restricted
  int32_t w;  // to one node's incident edges.
  bool operator<(const Key& o) const { return c < o.c || (c == o.c && w < o.w); }
};

// Sorted insertion into a bounded best-K list.
inline void push_best(Key* best, int& cnt, int K, Key k) {
  if (cnt == K && !(k < best[K - 1])) return;
  int i = cnt < K ? cnt++ : K - 1;
  while (i > 0 && k < best[i - 1]) { best[i] = best[i - 1]; --i; }
  best[i] = k;
}

// Directed kNN.  brute: O(n^2) scan.  Otherwise a uniform grid with ring
// search; a ring is closed once every unvisited point's cost lower bound
// exceeds the current K-th cost, so the result equals the brute force.
void knn_directed(int n, const double* x, const double* y, int K, int mode,
                  bool brute, Pool& pool, int32_t* out) {
  if (K > n - 1) K = n - 1;  // FXY.hpp:76 (top_k = N-1)
  if (K <= 0) return;
  if (brute) {
    pool.parallel_for(n, [&](int64_t lo, int64_t hi) {
      std::vector<Key> best(K);
      for (int64_t v = lo; v < hi; ++v) {
        int cnt = 0;
        for (int w = 0; w < n; ++w) {
          if (w == v) continue;
          push_best(best.data(), cnt, K, Key{cost_of(x[v], y[v], x[w], y[w], mode), w});
        }
        for (int j = 0; j < K; ++j) out[v * (int64_t)K + j] = best[j].w;
      }
    });
    return;
  }
  double mnx = x[0], mxx = x[0], mny = y[0], mxy = y[0];
  for (int i = 1; i < n; ++i) {
    mnx = std::min(mnx, x[i]); mxx = std::max(mxx, x[i]);
    mny = std::min(mny, y[i]); mxy = std::max(mxy, y[i]);
  }
  double w = mxx - mnx, h = mxy - mny;
  double cs = std::sqrt(std::max(w * h, 1e-300) / n);
  cs = std::max(cs, std::max(w, h) / n);
  if (!(cs > 0)) cs = 1.0;
  int64_t gx = (int64_t)(w / cs) + 1, gy = (int64_t)(h / cs) + 1;
  while (gx * gy > 4 * (int64_t)n + 16) { cs *= 1.5; gx = (int64_t)(w / cs) + 1; gy = (int64_t)(h / cs) + 1; }
  std::vector<int32_t> cell(n), cstart(gx * gy + 1, 0), order(n);
  for (int i = 0; i < n; ++i) {
    int64_t cx = std::min<int64_t>(gx - 1, (int64_t)((x[i] - mnx) / cs));
    int64_t cy = std::min<int64_t>(gy - 1, (int64_t)((y[i] - mny) / cs));
    cell[i] = (int32_t)(cy * gx + cx);
    cstart[cell[i] + 1]++;
  }
  for (int64_t c = 0; c < gx * gy; ++c) cstart[c + 1] += cstart[c];
  {
    std::vector<int32_t> fill(cstart.begin(), cstart.end() - 1);
    for (int i = 0; i < n; ++i) order[fill[cell[i]]++] = i;
  }
  pool.parallel_for(n, [&](int64_t lo, int64_t hi) {
    std::vector<Key> best(K);
    for (int64_t v = lo; v < hi; ++v) {
      int cnt = 0;
      int64_t cx = cell[v] % gx, cy = cell[v] / gx;
      for (int64_t R = 0;; ++R) {
        for (int64_t yy = cy - R; yy <= cy + R; ++yy) {
          if (yy < 0 || yy >= gy) continue;
          bool edge_row = (yy == cy - R || yy == cy + R);
          for (int64_t xx = cx - R; xx <= cx + R; xx += (edge_row ? 1 : 2 * R)) {
            if (xx >= 0 && xx < gx) {
              int64_t c = yy * gx + xx;
              for (int32_t p = cstart[c]; p < cstart[c + 1]; ++p) {
                int32_t u = order[p];
                if (u == v) continue;
                push_best(best.data(), cnt, K, Key{cost_of(x[v], y[v], x[u], y[u], mode), u});
              }
            }
            if (!edge_row && R == 0) break;
            if (R == 0) break;
          }
        }
        if (R > gx && R > gy) break;
        if (cnt == K) {
          // every unvisited point is farther than R*cs from v
          double lb = round_like_cost((double)R * cs * (1.0 - 1e-9), mode);
          if (best[K - 1].c < lb) break;
        }
      }
      for (int j = 0; j < K; ++j) out[v * (int64_t)K + j] = best[j].w;
    }
  });
}

struct Graph {
  int n = 0;
  std::vector<int32_t> rowptr, nbr;
  std::vector<double> cost;
};

// Symmetrised selection: union over v of kNN_K(v).  This is the edge set the
// reference's greedy pass keeps (graph_tour.hpp:28-90): an edge is skipped
// only when both endpoints already hold K kept edges, and since candidates
// are visited in increasing order, a node's count reaches K exactly at its
// K-th nearest edge.  Slots sorted by (cost, nbr) = the node's edge order
// (graph_tour.hpp:88, rankEDo node.cpp:18).
void build_graph(Graph& G, int n, const double* x, const double* y, int K, int mode,
                 bool brute, Pool& pool) {
  G.n = n;
  int Ke = std::min(K, n - 1);
  std::vector<int32_t> knn((size_t)n * std::max(Ke, 0));
  if (Ke > 0) knn_directed(n, x, y, Ke, mode, brute, pool, knn.data());
  std::vector<std::vector<Key>> adj(n);
  for (int v = 0; v < n; ++v)
    for (int j = 0; j < Ke; ++j) {
      int w = knn[(size_t)v * Ke + j];
      double c = cost_of(x[v], y[v], x[w], y[w], mode);
      adj[v].push_back(Key{c, w});
      adj[w].push_back(Key{c, v});
    }
  G.rowptr.assign(n + 1, 0);
  for (int v = 0; v < n; ++v) {
    auto& a = adj[v];
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end(),
                        [](const Key& p, const Key& q) { return p.w == q.w; }),
            a.end());
    G.rowptr[v + 1] = G.rowptr[v] + (int32_t)a.size();
  }
  G.nbr.resize(G.rowptr[n]);
  G.cost.resize(G.rowptr[n]);
  for (int v = 0; v < n; ++v) {
    int32_t b = G.rowptr[v];
    for (size_t j = 0; j < adj[v].size(); ++j) {
      G.nbr[b + j] = adj[v][j].w;
      G.cost[b + j] = adj[v][j].c;
    }
  }
}

// ---------------------------------------------------------------------------
// Dual descent.
// ---------------------------------------------------------------------------
inline double dval(double D, double Fv, double Fw, int v, int w) {
  // graph_cuda_grad.hpp:9 : d = D + 0.5*F[v1] + 0.5*F[v2], v1 < v2
  // (Edges_v1/v2 ordered by node id, graph_cuda_init.hpp:25-27)
  double a = v < w ? Fv : Fw;
  double b = v < w ? Fw : Fv;
  double t = D + 0.5 * a;
  return t + 0.5 * b;
}

struct Top3 {  // graph_cuda_grad.hpp:68-79 (three smallest, 1e6 sentinels)
  double a = 1e6, b = 1e6, c = 1e6;
  inline void push(double v) {
    if (v < a) { c = b; b = a; a = v; }
    else if (v < b) { c = b; b = v; }
    else if (v < c) { c = v; }
  }
  inline double grad() const { return -(b + c) / 2; }  // graph_cuda_grad.hpp:81
};

// Hoare partition of (d, idx) around thr; qpart_ptr_B (graph_cuda_partition.hpp:3-25).
// The bounds tests come first here; the reference tests the value first and
// so reads one element past either end, which does not change the result.
int qpart_B(double* d, int32_t* idx, int size, double thr) {
  int i = 0, j = size - 1;
  while (i < j) {
    while (i <= size - 1 && d[i] <= thr) ++i;
    while (j >= 0 && d[j] > thr) --j;
    if (i < j) { std::swap(d[i], d[j]); std::swap(idx[i], idx[j]); }
  }
  return j + 1;
}

// qpart_ptr_EF (graph_cuda_partition.hpp:27-54): re-partition with thr += 1
// until at least 3 entries lie at or below it.  A re-partition whose
// threshold admits no new entry swaps nothing, so only thresholds that admit
// a new entry are partitioned; the thresholds follow the same += 1 sequence.
int qpart_EF(double* d, int32_t* idx, int size, double thr) {
  int p = qpart_B(d, idx, size, thr);
  if (size < 3) return p;  // the reference loops forever here (never reached)
  while (p < 3) {
    double m = d[p];
    for (int k = p + 1; k < size; ++k) m = std::min(m, d[k]);
    int64_t guard = 0;
    do { thr += 1.0; } while (!(m <= thr) && ++guard < (int64_t)1 << 40);
    p = qpart_B(d, idx, size, thr);
  }
  return p;
}

struct Solver {
  int n;
  const orc_opts& o;
  Pool& pool;
  std::vector<int32_t> rp, deg, act, clen, nbr, orig;
  std::vector<double> D, cost, F, Fn, g;
  std::vector<uint8_t> frozen;
  double dmin = 0, eps = 0;
  int64_t sweeps = 0;

  Solver(int n_, const int32_t* rowptr, const int32_t* nb, const double* c, const orc_opts& o_,
         Pool& p)
      : n(n_), o(o_), pool(p) {
    rp.assign(rowptr, rowptr + n + 1);
    int64_t m = rp[n];
    nbr.assign(nb, nb + m);
    cost.assign(c, c + m);
    orig.resize(m);
    std::iota(orig.begin(), orig.end(), 0);
    deg.resize(n);
    for (int v = 0; v < n; ++v) deg[v] = rp[v + 1] - rp[v];
    // graph_tour.hpp:94: D = Do / d_min.  Zero-cost edges (NINT rounding of a
    // distance < 0.5) keep D = 0; the reference aborts on them
    // (graph_tour.hpp:44-49), here d_min is the smallest positive cost.
    dmin = 0;
    for (int64_t k = 0; k < m; ++k)
      if (cost[k] > 0 && (dmin == 0 || cost[k] < dmin)) dmin = cost[k];
    D.resize(m);
    for (int64_t k = 0; k < m; ++k) D[k] = cost[k] / dmin;
    F.assign(n, 0.0);
    Fn.assign(n, 0.0);
    g.assign(n, 0.0);
    act.assign(n, 0);
    clen.assign(n, 0);
    frozen.assign(n, 0);
    eps = o.err_grad * o.err_comb_mult;  // graph_cuda_comb.hpp:37
  }

  inline double d_at(int v, int64_t k) const { return dval(D[k], F[v], F[nbr[k]], v, nbr[k]); }

  inline double grad_node(int v, int len) const {
    Top3 t;
    int64_t b = rp[v];
    for (int k = 0; k < len; ++k) t.push(d_at(v, b + k));
    return t.grad();
  }

  // cuda_get_Nodes_grad_R over the first len[v] slots; returns max |g|
  // (graph_cuda_maxmin.hpp:29-37 returns the signed value; callers use fabs).
  double grad_all(const std::vector<int32_t>& len) {
    std::vector<double> part(pool.size(), 0.0);
    pool.run([&](int t, int nt) {
      int64_t lo = (int64_t)n * t / nt, hi = (int64_t)n * (t + 1) / nt;
      double m = 0;
      for (int64_t v = lo; v < hi; ++v) {
        g[v] = grad_node((int)v, len[v]);
        m = std::max(m, std::fabs(g[v]));
      }
      part[t] = m;
    });
    return *std::max_element(part.begin(), part.end());
  }

  // One sweep: grad_R -> update_F[_BLOCK] -> recompute (graph_cuda.cu:293-304,
  // 278-290).  Returns max |g| of the gradient used.
  double sweep(const std::vector<int32_t>& len, bool use_frozen) {
    std::vector<double> part(pool.size(), 0.0);
    pool.run([&](int t, int nt) {
      int64_t lo = (int64_t)n * t / nt, hi = (int64_t)n * (t + 1) / nt;
      double m = 0;
      for (int64_t v = lo; v < hi; ++v) {
        double gv = grad_node((int)v, len[v]);
        g[v] = gv;
        m = std::max(m, std::fabs(gv));
        Fn[v] = (use_frozen && frozen[v]) ? F[v] : F[v] + gv;
      }
      part[t] = m;
    });
    F.swap(Fn);
    ++sweeps;
    return *std::max_element(part.begin(), part.end());
  }

  // Partition node v's first `size` slots, permuting the slots in place.
  int partition(int v, int size, double thr, int kind /*0 B, 1 EF*/) {
    int64_t b = rp[v];
    double dl[512];
    int32_t il[512];
    std::vector<double> dv;
    std::vector<int32_t> iv;
    double* dp = dl;
    int32_t* ip = il;
    if (size > 512) { dv.resize(size); iv.resize(size); dp = dv.data(); ip = iv.data(); }
    for (int k = 0; k < size; ++k) { dp[k] = d_at(v, b + k); ip[k] = k; }
    int p = kind ? qpart_EF(dp, ip, size, thr) : qpart_B(dp, ip, size, thr);
    // apply the permutation to (nbr, D, cost, orig)
    int32_t tn[512]; double tD[512], tc[512]; int32_t to[512];
    std::vector<int32_t> vn, vo; std::vector<double> vD, vc;
    int32_t* pn = tn; int32_t* po = to; double* pD = tD; double* pc = tc;
    if (size > 512) { vn.resize(size); vo.resize(size); vD.resize(size); vc.resize(size);
      pn = vn.data(); po = vo.data(); pD = vD.data(); pc = vc.data(); }
    for (int k = 0; k < size; ++k) {
      pn[k] = nbr[b + ip[k]]; po[k] = orig[b + ip[k]]; pD[k] = D[b + ip[k]]; pc[k] = cost[b + ip[k]];
    }
    for (int k = 0; k < size; ++k) { nbr[b + k] = pn[k]; orig[b + k] = po[k]; D[b + k] = pD[k]; cost[b + k] = pc[k]; }
    return p;
  }

  // cuda_partition_v4 (graph_cuda_partition.hpp:89-115)
  void partition_v4(int kk) {
    pool.parallel_for(n, [&](int64_t lo, int64_t hi) {
      for (int64_t v = lo; v < hi; ++v) {
        if (deg[v] < kk) { act[v] = deg[v]; continue; }
        double thr = d_at((int)v, rp[v]);
        for (int i = 1; i < kk; ++i) thr = std::max(thr, d_at((int)v, rp[v] + i)) ;
        act[v] = partition((int)v, deg[v], thr, 1);
      }
    });
  }
  // cuda_partition_v2 (graph_cuda_partition.hpp:57-69), threshold 1.0
  void partition_v2(double thr) {
    pool.parallel_for(n, [&](int64_t lo, int64_t hi) {
      for (int64_t v = lo; v < hi; ++v) {
        if (deg[v] < 10) act[v] = deg[v];
        else act[v] = partition((int)v, deg[v], thr, 1);
      }
    });
  }

  // block_GLOB (graph_cuda.cu:363-413)
  int block_glob() {
    for (int v = 0; v < n; ++v) act[v] = std::min(deg[v], o.glob_k);  // setvalue2V
    int go = 0;
    for (; go < o.glob_rounds; ++go) {
      for (int s = 0; s < o.glob_sweeps; ++s) sweep(act, false);  // block_GLOB_N
      partition_v4(o.glob_k);
      double M = grad_all(act);
      if (M < o.glob_exit) { ++go; break; }
    }
    return go;
  }
  // check_OPT (graph_cuda.cu:254-271) + update_EF (graph_cuda.cu:229-252):
  // the active list of v is the prefix of act[v] slots of its permuted list.
  double check_opt() {
    partition_v2(1.0);
    return grad_all(act);
  }
  // block_EF (graph_cuda.cu:308-338)
  void block_ef(int N, bool use_frozen) {
    double MA = grad_all(act);
    int cnt = 0;
    double MB = MA;
    for (int go = 0; go < N; ++go) {
      MB = sweep(act, use_frozen);
      if (++cnt >= o.check_every) {
        double M = o.early_exit ? MB : MA;
        if (M < o.err_grad) break;
        cnt = 0;
      }
    }
  }
  // run_start_cuda (graph_cuda.cu:163-199)
  int run_start(double& M) {
    M = check_opt();
    int factor = o.factor0, go = 1;
    for (; go <= o.max_outer; ++go) {
      block_ef(factor, false);
      M = check_opt();
      if (factor < o.factor_max) factor += o.factor_step;
      if (M < o.err_grad) break;
    }
    return go;
  }
  // run_fix_cuda (graph_cuda.cu:204-226)
  double run_fix() {
    double M = check_opt();
    for (int go = 1; go <= o.fix_rounds; ++go) {
      block_ef(o.fix_factor, false);
      M = check_opt();
      if (M < o.err_grad) break;
    }
    return M;
  }

  // ---- classes and tight components (graph_cuda_comb.hpp:296-356) ----
  std::vector<int32_t> A, T, root;
  void classify() {
    // partition_v3: the comb scan length is the count of d <= 0.25 within the
    // active prefix (graph_cuda_comb.hpp:301)
    A.assign(n, 0); T.assign(n, 0);
    pool.parallel_for(n, [&](int64_t lo, int64_t hi) {
      for (int64_t v = lo; v < hi; ++v) {
        clen[v] = partition((int)v, act[v], 0.25, 0);
        int a = 0, t = 0;
        for (int k = 0; k < clen[v]; ++k) {
          double d = d_at((int)v, rp[v] + k);
          if (d < -eps) ++a;
          else if (d <= eps) ++t;
        }
        A[v] = a; T[v] = t;
      }
    });
    // connected components of the tight subgraph; root = smallest node id
    // (get_cuda_comb_size emits only from the component's minimum id)
    std::vector<int32_t> par(n);
    std::iota(par.begin(), par.end(), 0);
    auto find = [&](int32_t a) {
      while (par[a] != a) { par[a] = par[par[a]]; a = par[a]; }
      return a;
    };
    for (int v = 0; v < n; ++v)
      for (int k = 0; k < clen[v]; ++k) {
        int64_t s = rp[v] + k;
        double d = d_at(v, s);
        if (d >= -eps && d <= eps) {
          int32_t a = find(v), b = find(nbr[s]);
          if (a != b) { if (a < b) par[b] = a; else par[a] = b; }
        }
      }
    root.assign(n, -1);
    for (int v = 0; v < n; ++v)
      if (T[v] > 0) root[v] = find(v);
  }
};

}  // namespace

// ===========================================================================
// extern "C" API
// ===========================================================================
extern "C" {

void orc_default_opts(orc_opts* o) {
  o->err_grad = 1e-7;
  o->err_comb_mult = 1e2;
  o->glob_k = 5;
  o->glob_rounds = 200;
  o->glob_sweeps = 50;
  o->glob_exit = 0.01;
  o->max_outer = 10000;
  o->factor0 = 100;
  o->factor_step = 300;
  o->factor_max = 5000;
  o->check_every = 50;
  o->early_exit = 0;
  o->repair = 1;
  o->shift0 = 0.5;
  o->shift_max = 32;
  o->fix_rounds = 100;
  o->fix_factor = 500;
  o->frozen_sweeps = 20;
  o->stop_phase = 0;
  o->exact = 1;
  o->exact_alpha = 4;
  o->exact_eps0 = 0.02;
  o->exact_tol = 1e-10;
  o->exact_max_rounds = 1000000;
  o->exact_max_phases = 40;
  o->exact_reserved = 0;
}

double orc_cost(double xu, double yu, double xw, double yw, int mode) {
  return cost_of(xu, yu, xw, yw, mode);
}

int orc_knn_directed(int n, const double* x, const double* y, int K, int mode, int brute,
                     int nthreads, int32_t* out) {
  if (n < 2 || K < 1) return -1;
  Pool pool(nthreads);
  knn_directed(n, x, y, K, mode, brute != 0, pool, out);
  return 0;
}

void* orc_graph_build(int n, const double* x, const double* y, int K, int mode, int brute,
                      int nthreads) {
  if (n < 2 || K < 1) return nullptr;
  Pool pool(nthreads);
  Graph* G = new Graph;
  build_graph(*G, n, x, y, K, mode, brute != 0, pool);
  return G;
}
int64_t orc_graph_nslots(void* h) { return static_cast<Graph*>(h)->rowptr.back(); }
void orc_graph_export(void* h, int32_t* rowptr, int32_t* nbr, double* cost) {
  Graph* G = static_cast<Graph*>(h);
  std::memcpy(rowptr, G->rowptr.data(), sizeof(int32_t) * (G->n + 1));
  std::memcpy(nbr, G->nbr.data(), sizeof(int32_t) * G->nbr.size());
  std::memcpy(cost, G->cost.data(), sizeof(double) * G->cost.size());
}
void orc_graph_free(void* h) { delete static_cast<Graph*>(h); }

void orc_sweep(int n, const int32_t* rowptr, const int32_t* nbr, const double* D,
               const int32_t* len, const double* F_in, double* F_out, double* g_out) {
  for (int v = 0; v < n; ++v) {
    Top3 t;
    for (int k = 0; k < len[v]; ++k) {
      int64_t s = rowptr[v] + k;
      t.push(dval(D[s], F_in[v], F_in[nbr[s]], v, nbr[s]));
    }
    double gv = t.grad();
    if (g_out) g_out[v] = gv;
    F_out[v] = F_in[v] + gv;
  }
}

int orc_qpart(double* d, int32_t* idx, int size, double thr, int raise) {
  return raise ? qpart_EF(d, idx, size, thr) : qpart_B(d, idx, size, thr);
}

// Ax_b::parse_Ax_b + Ax_b::solve (Ax_b.cpp:19-60, 85-158): steepest descent
// with exact line search on the row-normalised residual, x0 = 0, at most 999
// iterations, stop when max |r_j| / |A_j| < 1e-8 (Ax_b.h:39).
int orc_axb(int m, const int32_t* row_ptr, const int32_t* col, const int32_t* b, int nv,
            double* x_out, double* max_err) {
  std::vector<double> x(nv, 0.0), xg(nv), dj(m), bg(m), Aj(m);
  for (int j = 0; j < m; ++j) Aj[j] = std::sqrt((double)(row_ptr[j + 1] - row_ptr[j]));
  int go_max = 0;
  double dmax = 0;
  for (int go = 1; go < 1000; ++go) {
    std::fill(xg.begin(), xg.end(), 0.0);
    dmax = 0;
    for (int j = 0; j < m; ++j) {
      double v = 0;
      for (int k = row_ptr[j]; k < row_ptr[j + 1]; ++k) v += x[col[k]];
      dj[j] = (v - b[j]) / Aj[j];
      if (std::fabs(dj[j]) > dmax) dmax = std::fabs(dj[j]);
      for (int k = row_ptr[j]; k < row_ptr[j + 1]; ++k) xg[col[k]] += dj[j] / Aj[j];
    }
    if (dmax < 1e-8) { go_max = go; break; }
    for (int j = 0; j < m; ++j) {
      double v = 0;
      for (int k = row_ptr[j]; k < row_ptr[j + 1]; ++k) v += xg[col[k]];
      bg[j] = v;
    }
    double gg = 0, gd = 0;
    for (int j = 0; j < m; ++j) { gg += bg[j] * bg[j]; gd += dj[j] * bg[j]; }
    // Deviation: the reference divides 0/0 here (an inconsistent system at its
    // least-squares point), fills x with NaN and then reports convergence
    // because fabs(NaN) > d_max is false.  Here gg == 0 ends the solve as
    // not converged.
    if (gg == 0) break;
    double alfa = -gd / gg;
    for (int i = 0; i < nv; ++i) x[i] += xg[i] * alfa;
  }
  for (int i = 0; i < nv; ++i) x_out[i] = x[i];
  if (max_err) *max_err = dmax;
  return go_max;
}

int orc_solve(int n, const int32_t* rowptr, const int32_t* nbr, const double* cost,
              const orc_opts* opts, int nthreads, double* F_out, double* x_out,
              int32_t* comb_root, orc_stats* st) {
  return orc_solve_warm(n, rowptr, nbr, cost, opts, nthreads, nullptr, F_out, x_out, comb_root, st);
}

int orc_solve_warm(int n, const int32_t* rowptr, const int32_t* nbr, const double* cost,
                   const orc_opts* opts, int nthreads, const double* F_init, double* F_out, double* x_out,
                   int32_t* comb_root, orc_stats* st) {
  orc_opts o;
  if (opts) o = *opts; else orc_default_opts(&o);
  Pool pool(nthreads);
  Solver S(n, rowptr, nbr, cost, o, pool);
  orc_stats z{};
  z.d_min = S.dmin;
  if (!(S.dmin > 0)) { if (st) *st = z; return -2; }

  // graph::RUN (graph_cuda.cu:47-155); a warm start (new) runs the same
  // schedule from the given potentials instead of zero.  (Skipping block_GLOB
  // as well was tried: from a perturbed optimum the Jacobi descent then failed
  // to reach max|g| < 1e-7 on some instances, running to the outer-round cap.)
  if (F_init) {
    for (int v = 0; v < n; ++v) S.F[v] = F_init[v];
  }
  z.glob_rounds = S.block_glob();
  double M = 0;
  if (o.stop_phase != 1) z.outer_rounds = S.run_start(M);
  z.max_grad = M;

  auto finish_F = [&]() { if (F_out) std::memcpy(F_out, S.F.data(), sizeof(double) * n); };
  if (o.stop_phase == 1 || o.stop_phase == 2) {
    finish_F();
    z.sweeps = S.sweeps;
    if (st) *st = z;
    return 0;
  }

  const double eps = S.eps;
  int64_t m = S.rp[n];
  std::vector<double> xs(m, 0.0);  // x per working slot
  int irr_b = 0;
  std::vector<int32_t> irregular_b_roots;

  // host_combs_check (graph_cuda_comb.hpp:2-28): regular combs get 1/2 on their
  // tight edges (comb_thread.cpp:180-195); the others go to least squares
  // (comb_thread.cpp:272-304).  Returns the number of IRREGULAR_B combs.
  auto combs_check = [&]() {
    S.classify();
    std::vector<int32_t> roots;
    for (int v = 0; v < n; ++v) if (S.root[v] == v) roots.push_back(v);
    z.n_combs = (int32_t)roots.size();
    std::vector<uint8_t> regular(n, 1);
    for (int v = 0; v < n; ++v)
      if (S.root[v] >= 0 && 4 - 2 * S.A[v] - S.T[v] != 0) regular[S.root[v]] = 0;
    std::fill(xs.begin(), xs.end(), 0.0);
    for (int v = 0; v < n; ++v)
      for (int k = 0; k < S.clen[v]; ++k) {
        int64_t s = S.rp[v] + k;
        double d = S.d_at(v, s);
        if (d < -eps) xs[s] = 1.0;  // node::get_eA_X (node.cpp:61-71)
        else if (d <= eps && regular[S.root[v]]) xs[s] = 0.5;
      }
    // irregular combs: least squares  sum_{e in T(v)} x_e = 2 - A(v)
    std::vector<std::vector<int32_t>> members(n);
    int na = 0;
    irregular_b_roots.clear();
    for (int v = 0; v < n; ++v)
      if (S.root[v] >= 0 && !regular[S.root[v]]) members[S.root[v]].push_back(v);
    for (int r = 0; r < n; ++r) {
      if (members[r].empty()) continue;
      ++na;
      // with exact recovery on, irregular combs go to the exact stage: their
      // least squares (and the dual repairs it would start) are not run
      if (o.exact >= 1) continue;
      // variables: tight edges keyed by (min id, max id), first appearance
      // over nodes ascending, each node's tight edges by neighbour id
      std::vector<std::pair<int64_t, int32_t>> var;  // key -> var id
      auto key_of = [&](int a, int b) { return (int64_t)std::min(a, b) * n + std::max(a, b); };
      std::vector<int32_t> rptr{0}, colv, bv;
      for (int v : members[r]) {
        std::vector<std::pair<int32_t, int64_t>> tl;  // (nbr, slot)
        for (int k = 0; k < S.clen[v]; ++k) {
          int64_t s = S.rp[v] + k;
          double d = S.d_at(v, s);
          if (d >= -eps && d <= eps) tl.push_back({S.nbr[s], s});
        }
        std::sort(tl.begin(), tl.end());
        for (auto& t : tl) {
          int64_t kk = key_of(v, t.first);
          int32_t id = -1;
          for (auto& pv : var) if (pv.first == kk) { id = pv.second; break; }
          if (id < 0) { id = (int32_t)var.size(); var.push_back({kk, id}); }
          colv.push_back(id);
        }
        rptr.push_back((int32_t)colv.size());
        bv.push_back(2 - S.A[v]);
      }
      std::vector<double> xv(var.size());
      double err = 0;
      int ok = orc_axb((int)bv.size(), rptr.data(), colv.data(), bv.data(), (int)var.size(),
                       xv.data(), &err);
      if (!ok) irregular_b_roots.push_back(r);
      // XB is stored as float (F2M.h:202)
      for (int v : members[r])
        for (int k = 0; k < S.clen[v]; ++k) {
          int64_t s = S.rp[v] + k;
          double d = S.d_at(v, s);
          if (d >= -eps && d <= eps) {
            int64_t kk = key_of(v, S.nbr[s]);
            for (auto& pv : var) if (pv.first == kk) { xs[s] = (double)(float)xv[pv.second]; break; }
          }
        }
    }
    z.n_irregular_a = na;
    return (int)irregular_b_roots.size();
  };

  irr_b = combs_check();
  z.n_irregular_b = irr_b;

  // dual perturbation of IRREGULAR_B combs (graph_cuda.cu:69-126,
  // graph_cuda_comb.hpp:365-455).  Type 1 shifts only the first violating node;
  // "first" is the smallest node id here (the reference uses BFS order).
  if (irr_b > 0 && o.repair) {
    double shiftV = o.shift0;
    auto attempt = [&](double sv, int type) {
      ++z.n_repairs;
      std::fill(S.frozen.begin(), S.frozen.end(), 0);
      std::vector<std::vector<int32_t>> members(n);
      for (int v = 0; v < n; ++v) if (S.root[v] >= 0) members[S.root[v]].push_back(v);
      for (int r : irregular_b_roots)
        for (int v : members[r]) {
          int s = 4 - 2 * S.A[v] - S.T[v];
          if (s != 0) {
            S.F[v] += sv * (s > 0 ? 1 : -1);
            S.frozen[v] = 1;
            if (type == 1) break;
          }
        }
      for (int k = 0; k < o.frozen_sweeps; ++k) S.sweep(S.act, true);
      std::fill(S.frozen.begin(), S.frozen.end(), 0);
      z.max_grad = S.run_fix();
      return combs_check();
    };
    for (;;) {
      if ((irr_b = attempt(shiftV, 0)) == 0) break;
      if ((irr_b = attempt(shiftV, 1)) == 0) break;
      shiftV = -shiftV;
      if ((irr_b = attempt(shiftV, 0)) == 0) break;
      shiftV = -shiftV;
      if ((irr_b = attempt(shiftV, 1)) == 0) break;
      shiftV = -shiftV;
      shiftV *= 2;
      if (std::fabs(shiftV) > o.shift_max) break;
    }
  }
  z.status = irr_b == 0 ? 0 : 1;

  // outputs in canonical slot order
  std::vector<int32_t> rank_of(m);  // canonical slot -> rank within its node
  for (int v = 0; v < n; ++v)
    for (int32_t k = rowptr[v]; k < rowptr[v + 1]; ++k) rank_of[k] = k - rowptr[v];
  std::vector<double> xc(m, 0.0);
  std::vector<uint8_t> inX(m, 0);
  for (int64_t s = 0; s < m; ++s) {
    xc[S.orig[s]] = xs[s];
  }
  for (int v = 0; v < n; ++v)
    for (int k = 0; k < S.clen[v]; ++k) {
      int64_t s = S.rp[v] + k;
      if (S.d_at(v, s) <= eps) inX[S.orig[s]] = 1;
    }
  // exact recovery (f2m_exact_oracle.cpp) where the reference's recovery is
  // not exact: status 1, or some tight component irregular (its x came from
  // the float least squares, possibly outside [0,1])
  const bool run_exact = o.exact >= 2 || (o.exact == 1 && (z.status != 0 || z.n_irregular_a > 0));
  std::vector<int32_t> hroot;
  if (run_exact) {
    std::vector<double> Fx(n);
    hroot.assign(n, -1);
    const int rc = orc_exact(n, rowptr, nbr, cost, S.dmin, S.F.data(), &o, Fx.data(), xc.data(), hroot.data(), &z);
    if (rc >= 0) {
      z.status = rc == 0 ? 0 : 1;
      S.F = Fx;
      for (int64_t k = 0; k < m; ++k) inX[k] = xc[k] > 0.0;
      z.n_combs = 0;
      for (int v = 0; v < n; ++v) z.n_combs += hroot[v] == v;
    } else {
      hroot.clear();
      z.status = 1;
    }
  }
  if (x_out) std::memcpy(x_out, xc.data(), sizeof(double) * m);
  // objective sum over undirected edges in canonical order (graph_tour.hpp:165-184)
  double obj = 0;
  int maxrank = 0;
  for (int v = 0; v < n; ++v)
    for (int32_t k = rowptr[v]; k < rowptr[v + 1]; ++k) {
      int w = nbr[k];
      if (w <= v) continue;
      obj += xc[k] * cost[k];
      if (inX[k]) {
        int r2 = -1;
        for (int32_t q = rowptr[w]; q < rowptr[w + 1]; ++q) if (nbr[q] == v) { r2 = q - rowptr[w]; break; }
        int r = std::min(rank_of[k], r2 < 0 ? rank_of[k] : r2);
        maxrank = std::max(maxrank, r);
      }
    }
  z.objective = obj;
  z.max_rank = maxrank;
  // dual bound: d_min * (-sum F + sum_e min(0, d_e))
  double sF = 0, sneg = 0;
  for (int v = 0; v < n; ++v) sF += S.F[v];
  for (int v = 0; v < n; ++v)
    for (int32_t k = S.rp[v]; k < S.rp[v + 1]; ++k) {
      if (S.nbr[k] <= v) continue;
      sneg += std::min(0.0, S.d_at(v, k));
    }
  z.dual = S.dmin * (-sF + sneg);
  z.sweeps = S.sweeps;
  finish_F();
  if (comb_root) std::memcpy(comb_root, hroot.empty() ? S.root.data() : hroot.data(), sizeof(int32_t) * n);
  if (st) *st = z;
  return 0;
}

void orc_bound_backward(int n, const int32_t* rowptr, const int32_t* nbr, const double* x_slot,
                        const double* px, const double* py, double* gx, double* gy) {
  for (int u = 0; u < n; ++u) {
    double ax = 0, ay = 0;
    for (int32_t k = rowptr[u]; k < rowptr[u + 1]; ++k) {
      if (x_slot[k] == 0) continue;
      int w = nbr[k];
      double dx = px[u] - px[w], dy = py[u] - py[w];
      double r = std::sqrt(dx * dx + dy * dy);
      if (r > 0) { ax += x_slot[k] * dx / r; ay += x_slot[k] * dy / r; }
    }
    gx[u] = ax;
    gy[u] = ay;
  }
}

}  // extern "C"

extern "C" {

int64_t orc_knn_certify(int n, const int32_t* rowptr, const double* cost, int K, double dmin, const double* F) {
  int64_t bad = 0;
  for (int v = 0; v < n; ++v) {
    const int32_t deg = rowptr[v + 1] - rowptr[v];
    if (deg < K) continue;  // v keeps every other point: nothing left out
    const double r = cost[rowptr[v] + K - 1];
    if (r < -(dmin * F[v])) ++bad;
  }
  return bad;
}

}  // extern "C"
