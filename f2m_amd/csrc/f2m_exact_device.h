// f2m_exact_device.h -- device routines of the exact primal recovery, shared
// by the single-graph stage (f2m_exact.hip, kernel per step) and the batched
// solve (f2m_batch.hip, one workgroup per instance).  Both execute the same
// synchronous schedule with the same arithmetic, so their results are equal
// bit for bit (and equal to the oracle's restatement).
#pragma once
#include <climits>

#include "f2m_device.h"

namespace f2m {

constexpr int XG = 16;        // lanes per node in the discharge steps
constexpr int DET_T = 1024;   // chunks of the fixed-order sums

// Per-slot and per-node state of one graph (or one instance, with local ids
// and slot indices).  Row bounds are passed separately.
struct XState {
  const int32_t* nbr;
  const double* c;
  const int32_t* mir;
  uint8_t* f;
  double* pL;
  double* pR;
  int32_t* eL;
  int32_t* eR;
};

// (key, slot) arg-max over the XG lanes of a node group: larger key wins,
// equal keys go to the lower slot.  A total order, so any reduction tree
// returns the same pair.
__device__ __forceinline__ void xg_best(double& k, int& s) {
#pragma unroll
  for (int off = XG / 2; off > 0; off >>= 1) {
    const double k2 = __shfl_xor(k, off, XG);
    const int s2 = __shfl_xor(s, off, XG);
    if (k2 > k || (k2 == k && s2 < s)) { k = k2; s = s2; }
  }
}

// best residual arc of L_v: max pR[w] - c over slots with f == 0
__device__ __forceinline__ void x_l_best(const XState& X, int b, int e, int lane, double& bv, int& bs) {
  bv = -INFINITY;
  bs = INT_MAX;
  for (int s = b + lane; s < e; s += XG) {
    if (X.f[s]) continue;
    const double val = X.pR[X.nbr[s]] - X.c[s];
    if (val > bv || (val == bv && s < bs)) { bv = val; bs = s; }
  }
  xg_best(bv, bs);
}
// costliest in-flow arc of R_w: max c + pL[v] over slots whose mirror has f == 1
__device__ __forceinline__ void x_r_best(const XState& X, int b, int e, int lane, double& bt, int& bs) {
  bt = -INFINITY;
  bs = INT_MAX;
  for (int s = b + lane; s < e; s += XG) {
    if (!X.f[X.mir[s]]) continue;
    const double t = X.c[s] + X.pL[X.nbr[s]];
    if (t > bt || (t == bt && s < bs)) { bt = t; bs = s; }
  }
  xg_best(bt, bs);
}

// Discharge L_v (slots [b, e)) by its XG lanes: push along the residual arcs
// of largest pR[w] - c, relabelling pL = (pR[w] - c) - eps when the arc is not
// admissible, then relabel once more if no admissible arc is left.  on_cross(w)
// runs (in the pushing lane) when eR[w] crosses 0 -> 1.
template <class Cross>
__device__ __forceinline__ void x_discharge_l(const XState& X, int v, int b, int e, int lane, double eps,
                                              Cross on_cross) {
  int ex = X.eL[v];
  if (ex <= 0) return;
  double p = X.pL[v];
  while (ex > 0) {
    double bv;
    int bs;
    x_l_best(X, b, e, lane, bv, bs);
    if (bs == INT_MAX) break;
    const int w = X.nbr[bs];
    if (!((X.c[bs] + p) - X.pR[w] < 0.0)) p = bv - eps;
    if ((bs - b) % XG == lane) {
      X.f[bs] = 1;
      if (atomicAdd(&X.eR[w], 1) == 0) on_cross(w);
    }
    --ex;
  }
  {
    double bv;
    int bs;
    x_l_best(X, b, e, lane, bv, bs);
    if (bs != INT_MAX && !((X.c[bs] + p) - X.pR[X.nbr[bs]] < 0.0)) p = bv - eps;
  }
  if (lane == 0) {
    X.pL[v] = p;
    X.eL[v] = ex;
  }
}

// Discharge R_w: return the in-flow arcs of largest c + pL[v], relabelling
// pR = (c + pL[v]) - eps when not admissible, then relabel once more if no
// admissible arc is left.  on_cross(v) runs when eL[v] crosses 0 -> 1.
template <class Cross>
__device__ __forceinline__ void x_discharge_r(const XState& X, int w, int b, int e, int lane, double eps,
                                              Cross on_cross) {
  int ex = X.eR[w];
  if (ex <= 0) return;
  double p = X.pR[w];
  while (ex > 0) {
    double bt;
    int bs;
    x_r_best(X, b, e, lane, bt, bs);
    if (bs == INT_MAX) break;
    if (!(bt - p > 0.0)) p = bt - eps;
    if ((bs - b) % XG == lane) {
      const int v = X.nbr[bs];
      X.f[X.mir[bs]] = 0;
      if (atomicAdd(&X.eL[v], 1) == 0) on_cross(v);
    }
    --ex;
  }
  {
    double bt;
    int bs;
    x_r_best(X, b, e, lane, bt, bs);
    if (bs != INT_MAX && !(bt - p > 0.0)) p = bt - eps;
  }
  if (lane == 0) {
    X.pR[w] = p;
    X.eR[w] = ex;
  }
}

// mirror slot of every slot of v (rows [rp(u), rp(u+1))); returns false on a
// missing or asymmetric mirror (then mir[s] = s).  Rows sorted by (cost, id)
// are searched by bisection; a row out of that order falls back to a scan.
template <class RP>
__device__ __forceinline__ bool x_mirror_slot(int v, int s, RP rp, const int32_t* __restrict__ nbr,
                                              const double* __restrict__ c, int32_t* __restrict__ mir,
                                              bool& nonint) {
  const int w = nbr[s];
  const double cs = c[s];
  const int qb = (int)rp(w), qe = (int)rp(w + 1);
  int lo = qb, hi = qe;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (c[mid] < cs || (c[mid] == cs && nbr[mid] < v)) lo = mid + 1;
    else hi = mid;
  }
  int q = lo;
  if (!(q < qe && nbr[q] == v)) {
    q = qb;
    while (q < qe && nbr[q] != v) ++q;
  }
  if (cs != floor(cs)) nonint = true;
  if (q == qe || c[q] != cs) {
    mir[s] = s;
    return false;
  }
  mir[s] = q;
  return true;
}
template <class RP>
__device__ __forceinline__ bool x_mirror_node(int v, RP rp, const int32_t* __restrict__ nbr,
                                              const double* __restrict__ c, int32_t* __restrict__ mir,
                                              bool& nonint) {
  bool ok = true;
  const int b = (int)rp(v), e = (int)rp(v + 1);
  for (int s = b; s < e; ++s) ok = x_mirror_slot(v, s, rp, nbr, c, mir, nonint) && ok;
  return ok;
}

// f = [rc < 0] (init) or the eps-violating arcs saturated / emptied.  With
// INC the excesses follow the flips (eL[v] by this caller, eR[w] atomically),
// which equals recomputing them when every excess was 0 before.
template <bool SAT, bool INC = false>
__device__ __forceinline__ void x_flow_node(const XState& X, int v, int b, int e, double eps) {
  const double pv = X.pL[v];
  int dl = 0;
  for (int s = b; s < e; ++s) {
    const double r = (X.c[s] + pv) - X.pR[X.nbr[s]];
    if (!SAT) {
      X.f[s] = r < 0.0 ? 1 : 0;
    } else {
      const uint8_t fs = X.f[s];
      if (!fs && r < -eps) {
        X.f[s] = 1;
        if (INC) { --dl; atomicAdd(&X.eR[X.nbr[s]], 1); }
      } else if (fs && r > eps) {
        X.f[s] = 0;
        if (INC) { ++dl; atomicAdd(&X.eR[X.nbr[s]], -1); }
      }
    }
  }
  if (INC && dl) X.eL[v] += dl;
}

__device__ __forceinline__ void x_excess_node(const XState& X, int v, int b, int e, int& el, int& er) {
  int out = 0, in = 0;
  for (int s = b; s < e; ++s) {
    out += X.f[s];
    in += X.f[X.mir[s]];
  }
  el = 2 - out;
  er = in - 2;
  X.eL[v] = el;
  X.eR[v] = er;
}

// terms of P = sum c x and D = 2 sum y - sum max(0, y_u + y_w - c), y = (pR - pL)/2
__device__ __forceinline__ void x_cert_node(const XState& X, int v, int b, int e, double& pv, double& dv) {
  const double yv = 0.5 * (X.pR[v] - X.pL[v]);
  double a = 0.0, bsum = 0.0;
  for (int s = b; s < e; ++s) {
    const int w = X.nbr[s];
    if (w <= v) continue;
    const double x = 0.5 * (double)(X.f[s] + X.f[X.mir[s]]);
    a += X.c[s] * x;
    const double yw = 0.5 * (X.pR[w] - X.pL[w]);
    bsum += fmax(0.0, (yv + yw) - X.c[s]);
  }
  pv = a;
  dv = 2.0 * yv - bsum;
}

}  // namespace f2m
