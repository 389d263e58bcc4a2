// f2m_exact_oracle.cpp -- CPU restatement of the exact primal recovery.
//
// TEST INFRASTRUCTURE ONLY (see f2m_oracle.h).
//
// This stage has no counterpart in the reference: it replaces the reference's
// least-squares recovery and dual-perturbation loop (comb_thread.cpp:272-304,
// Ax_b.cpp:139-158, graph_cuda.cu:69-126) where those fail.  Its parity pins
// are the LP optima of HiGHS (tests/golden/golden.json), primal feasibility
// and the duality gap certificate it computes itself.  The GPU stage
// (f2m_amd/csrc/f2m_exact.hip) executes the same synchronous schedule and is
// checked against this file bit for bit.
//
// Method.  The fractional 2-matching LP  min sum c_e x_e, sum_{e at v} x_e = 2,
// 0 <= x <= 1  equals half the min-cost transportation problem on the
// bipartite double cover: left copy L_v (supply 2), right copy R_w (demand 2),
// arc L_v -> R_w of capacity 1 and cost c_vw for every edge in both
// orientations; x_e = (f_vw + f_wv) / 2.  The transportation problem is
// solved by cost-scaling push-relabel (Goldberg-Tarjan) with eps-optimality
// rc(v,w) = c + pL[v] - pR[w] >= -eps on residual arcs.  Pushes alternate
// between all left nodes with excess and all right nodes with excess
// (Jacobi steps); since no arc joins two nodes of one side, a step equals
// some sequential order of valid push/relabel operations.  After its pushes a
// node also relabels when no admissible arc is left (a valid relabel of an
// inactive node), which makes a push act as an auction bid with the full
// margin to the next-best arc and cuts price wars among tied arcs by an order
// of magnitude on lattice instances.  Prices start from
// the dual descent's potentials (pL = d_min F/2, pR = -pL, so rc = d_min d_e)
// and eps shrinks by alpha per phase until the duality gap
//     P - D,  P = sum c_e x_e,  D = 2 sum y_v - sum_e max(0, y_u + y_w - c_e),
//     y_v = (pR[v] - pL[v]) / 2   (any y gives a lower bound D <= opt)
// is within the tolerance (below 1/2 for integer costs: opt is then exact).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "f2m_oracle.h"

namespace {

// Fixed-order sum shared with the GPU: 1024 contiguous chunks summed
// sequentially, then the 1024 partials sequentially.
double det_sum(const std::vector<double>& a) {
  const int64_t n = (int64_t)a.size();
  const int64_t T = 1024;
  const int64_t chunk = (n + T - 1) / T;
  double tot = 0.0;
  for (int64_t t = 0; t < T; ++t) {
    double s = 0.0;
    const int64_t b = t * chunk, e = std::min(n, b + chunk);
    for (int64_t i = b; i < e; ++i) s += a[i];
    tot += s;
  }
  return tot;
}

struct Exact {
  int n;
  int64_t m;
  const int32_t* rp;
  const int32_t* nbr;
  const double* c;
  std::vector<int32_t> mir;
  std::vector<uint8_t> f;
  std::vector<double> pL, pR;
  std::vector<int32_t> eL, eR;
  double eps = 0;

  Exact(int n_, const int32_t* rp_, const int32_t* nbr_, const double* c_)
      : n(n_), m(rp_[n_]), rp(rp_), nbr(nbr_), c(c_), mir(m, -1), f(m, 0), pL(n), pR(n), eL(n), eR(n) {}

  // mirror slot of (v -> w) in w's row; false if missing or asymmetric cost
  bool mirror() {
    for (int v = 0; v < n; ++v)
      for (int32_t s = rp[v]; s < rp[v + 1]; ++s) {
        const int w = nbr[s];
        int32_t q = rp[w];
        while (q < rp[w + 1] && nbr[q] != v) ++q;
        if (q == rp[w + 1] || c[q] != c[s]) return false;
        mir[s] = q;
      }
    return true;
  }
  double rc(int v, int32_t s) const { return (c[s] + pL[v]) - pR[nbr[s]]; }

  void init(const double* F, double dmin) {
    for (int v = 0; v < n; ++v) {
      pL[v] = 0.5 * (dmin * F[v]);
      pR[v] = -pL[v];
    }
    for (int v = 0; v < n; ++v)
      for (int32_t s = rp[v]; s < rp[v + 1]; ++s) f[s] = rc(v, s) < 0.0 ? 1 : 0;
  }
  // restore eps-optimality after eps shrank: saturate / empty violating arcs
  void saturate() {
    for (int v = 0; v < n; ++v)
      for (int32_t s = rp[v]; s < rp[v + 1]; ++s) {
        const double r = rc(v, s);
        if (!f[s] && r < -eps) f[s] = 1;
        else if (f[s] && r > eps) f[s] = 0;
      }
  }
  void excess() {
    for (int v = 0; v < n; ++v) {
      int out = 0, in = 0;
      for (int32_t s = rp[v]; s < rp[v + 1]; ++s) {
        out += f[s];
        in += f[mir[s]];
      }
      eL[v] = 2 - out;
      eR[v] = in - 2;
    }
  }
  // discharge L_v: push along the residual arcs of largest pR[w] - c (ties:
  // lowest slot); relabel pL = (pR[w] - c) - eps when that arc is not admissible
  void l_node(int v) {
    int e = eL[v];
    if (e <= 0) return;
    double p = pL[v];
    while (e > 0) {
      int32_t bs = -1;
      double bv = 0;
      for (int32_t s = rp[v]; s < rp[v + 1]; ++s) {
        if (f[s]) continue;
        const double val = pR[nbr[s]] - c[s];
        if (bs < 0 || val > bv) { bv = val; bs = s; }
      }
      if (bs < 0) break;
      if (!((c[bs] + p) - pR[nbr[bs]] < 0.0)) p = bv - eps;
      f[bs] = 1;
      if (++eR[nbr[bs]] == 1) actR.push_back(nbr[bs]);
      --e;
    }
    {  // look-ahead relabel: bid down to the next residual arc
      int32_t bs = -1;
      double bv = 0;
      for (int32_t s = rp[v]; s < rp[v + 1]; ++s) {
        if (f[s]) continue;
        const double val = pR[nbr[s]] - c[s];
        if (bs < 0 || val > bv) { bv = val; bs = s; }
      }
      if (bs >= 0 && !((c[bs] + p) - pR[nbr[bs]] < 0.0)) p = bv - eps;
    }
    pL[v] = p;
    eL[v] = e;
  }
  // discharge R_w: return the in-flow arcs of largest c + pL[v] (ties: lowest
  // slot of w's row); relabel pR = (c + pL[v]) - eps when not admissible
  void r_node(int w) {
    int e = eR[w];
    if (e <= 0) return;
    double p = pR[w];
    while (e > 0) {
      int32_t bs = -1;
      double bt = 0;
      for (int32_t s = rp[w]; s < rp[w + 1]; ++s) {
        if (!f[mir[s]]) continue;
        const double t = c[s] + pL[nbr[s]];
        if (bs < 0 || t > bt) { bt = t; bs = s; }
      }
      if (bs < 0) break;
      if (!(bt - p > 0.0)) p = bt - eps;
      f[mir[bs]] = 0;
      if (++eL[nbr[bs]] == 1) nxt.push_back(nbr[bs]);
      --e;
    }
    {  // look-ahead relabel: price up to the costliest kept unit
      int32_t bs = -1;
      double bt = 0;
      for (int32_t s = rp[w]; s < rp[w + 1]; ++s) {
        if (!f[mir[s]]) continue;
        const double t = c[s] + pL[nbr[s]];
        if (bs < 0 || t > bt) { bt = t; bs = s; }
      }
      if (bs >= 0 && !(bt - p > 0.0)) p = bt - eps;
    }
    pR[w] = p;
    eR[w] = e;
  }
  // Active nodes: every node with positive excess.  A step discharges each of
  // them completely, so the next step's active set is exactly the nodes whose
  // excess crossed from <= 0 to > 0 during the other side's step.  Processing
  // order inside a step does not matter (no arc joins two nodes of one side).
  std::vector<int32_t> actL, actR, nxt;
  void seed_active() {
    actL.clear();
    actR.clear();
    for (int v = 0; v < n; ++v) if (eL[v] > 0) actL.push_back(v);
    for (int v = 0; v < n; ++v) if (eR[v] > 0) actR.push_back(v);
  }
  // one L step and one R step; returns the number of active nodes seen
  int64_t round() {
    int64_t cnt = (int64_t)actL.size();
    for (int v : actL) l_node(v);
    actL.clear();
    cnt += (int64_t)actR.size();
    for (int w : actR) r_node(w);
    actR.clear();
    std::swap(actL, nxt);
    nxt.clear();
    return cnt;
  }
  void certificate(double& P, double& D) const {
    std::vector<double> pv(n), dv(n);
    for (int v = 0; v < n; ++v) {
      const double yv = 0.5 * (pR[v] - pL[v]);
      double a = 0.0, b = 0.0;
      for (int32_t s = rp[v]; s < rp[v + 1]; ++s) {
        const int w = nbr[s];
        if (w <= v) continue;
        const double x = 0.5 * (double)(f[s] + f[mir[s]]);
        a += c[s] * x;
        const double yw = 0.5 * (pR[w] - pL[w]);
        b += std::max(0.0, (yv + yw) - c[s]);
      }
      pv[v] = a;
      dv[v] = 2.0 * yv - b;
    }
    P = det_sum(pv);
    D = det_sum(dv);
  }
};

}  // namespace

extern "C" {

int orc_exact(int n, const int32_t* rowptr, const int32_t* nbr, const double* cost, double dmin,
              const double* F_in, const orc_opts* opts, double* F_out, double* x_out, int32_t* half_root,
              orc_stats* st) {
  orc_opts o;
  if (opts) o = *opts; else orc_default_opts(&o);
  if (st) st->exact_ran = 1;
  Exact X(n, rowptr, nbr, cost);
  if (!X.mirror()) return -1;
  for (int v = 0; v < n; ++v)
    if (rowptr[v + 1] - rowptr[v] < 2) return -2;
  bool integral = true;
  for (int64_t s = 0; s < X.m; ++s)
    if (cost[s] != std::floor(cost[s])) { integral = false; break; }
  X.init(F_in, dmin);
  X.eps = o.exact_eps0 * dmin;
  int phases = 0;
  int64_t rounds = 0;
  double P = 0, D = 0;
  bool certified = false, stuck = false;
  for (;;) {
    if (phases > 0) X.saturate();
    X.excess();
    X.seed_active();
    for (;;) {
      if (rounds >= o.exact_max_rounds) { stuck = true; break; }
      if (X.round() == 0) break;
      ++rounds;
    }
    ++phases;
    if (stuck) break;
    X.certificate(P, D);
    const double tol = std::max(integral ? 0.4 : 0.0, o.exact_tol * std::fabs(P));
    if (P - D <= tol) { certified = true; break; }
    X.eps = X.eps / (double)o.exact_alpha;
    if (phases >= o.exact_max_phases) break;
  }
  if (stuck) X.certificate(P, D);
  // outputs: x per canonical slot, potentials, 1/2-edge components
  for (int v = 0; v < n; ++v)
    for (int32_t s = rowptr[v]; s < rowptr[v + 1]; ++s)
      if (x_out) x_out[s] = 0.5 * (double)(X.f[s] + X.f[X.mir[s]]);
  if (F_out)
    for (int v = 0; v < n; ++v) F_out[v] = (X.pL[v] - X.pR[v]) / dmin;
  if (half_root) {
    std::vector<int32_t> par(n);
    for (int v = 0; v < n; ++v) par[v] = v;
    auto find = [&](int a) { while (par[a] != a) a = par[a] = par[par[a]]; return a; };
    std::vector<uint8_t> has(n, 0);
    for (int v = 0; v < n; ++v)
      for (int32_t s = rowptr[v]; s < rowptr[v + 1]; ++s)
        if (X.f[s] + X.f[X.mir[s]] == 1) {
          has[v] = 1;
          int a = find(v), b = find(nbr[s]);
          if (a != b) { if (a < b) std::swap(a, b); par[a] = b; }
        }
    for (int v = 0; v < n; ++v) half_root[v] = has[v] ? find(v) : -1;
  }
  if (st) {
    st->exact_ran = 1;
    st->exact_phases = phases;
    st->exact_rounds = rounds;
    st->exact_gap = P - D;
    st->exact_certified = certified ? 1 : 0;
  }
  return certified ? 0 : 1;
}

}  // extern "C"
