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

// The same discharges with the node's row gathered once into registers
// (lane `lane` holds slots b + lane + XG j, j < R; rows of at most XG R
// slots).  Within one step nothing a discharge reads changes except what
// that discharge writes: pR moves in R steps only and pL in L steps only;
// the flow of arc v -> w is set by L_v's discharge (L step) and cleared by
// R_w's (R step), so each arc has one writer per step.  The per-lane best
// uses the same (key, slot) order as x_l_best / x_r_best and the tests
// the same operands, so the outcome is the same bit for bit; the gathers
// go from one per push (plus one for the last relabel) to one.
template <int R, class Cross>
__device__ __forceinline__ void x_discharge_l_reg(const XState& X, int v, int b, int e, int lane, double eps,
                                                  Cross on_cross) {
  int ex = X.eL[v];
  if (ex <= 0) return;
  double p = X.pL[v];
  double key[R], cs[R], pw[R];
  int wn[R];
  unsigned ok = 0;  // bit j: slot j is a residual arc (f == 0)
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int s = b + lane + XG * j;
    key[j] = 0.0;
    cs[j] = 0.0;
    pw[j] = 0.0;
    wn[j] = 0;
    if (s < e) {
      wn[j] = X.nbr[s];
      cs[j] = X.c[s];
      if (!X.f[s]) ok |= 1u << j;
    }
  }
#pragma unroll
  for (int j = 0; j < R; ++j) {
    if (b + lane + XG * j < e) {
      pw[j] = X.pR[wn[j]];
      key[j] = pw[j] - cs[j];
    }
  }
  // best residual arc, and whether it is admissible at the current p
  auto best = [&](double& bv, int& bs, bool& adm) {
    bv = -INFINITY;
    bs = INT_MAX;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int s = b + lane + XG * j;
      if (((ok >> j) & 1u) && (key[j] > bv || (key[j] == bv && s < bs))) {
        bv = key[j];
        bs = s;
      }
    }
    xg_best(bv, bs);
    int a = 0;
    if (bs != INT_MAX && (bs - b) % XG == lane) {
      const int jb = (bs - b) / XG;
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (j == jb) a = (cs[j] + p) - pw[j] < 0.0 ? 1 : 0;
    }
    adm = __shfl(a, (bs - b) % XG, XG) != 0;
  };
  // the excess atomics' old values are tested after the pushes, so that
  // they are in flight together (each slot is pushed at most once)
  unsigned pushed = 0;
  int old[R];
#pragma unroll
  for (int j = 0; j < R; ++j) old[j] = 1;
  while (ex > 0) {
    double bv;
    int bs;
    bool adm;
    best(bv, bs, adm);
    if (bs == INT_MAX) break;
    if (!adm) p = bv - eps;
    if ((bs - b) % XG == lane) {
      const int jb = (bs - b) / XG;
      ok &= ~(1u << jb);
      pushed |= 1u << jb;
      X.f[bs] = 1;
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (j == jb) old[j] = atomicAdd(&X.eR[wn[j]], 1);
    }
    --ex;
  }
  {
    double bv;
    int bs;
    bool adm;
    best(bv, bs, adm);
    if (bs != INT_MAX && !adm) p = bv - eps;
  }
  if (lane == 0) {
    X.pL[v] = p;
    X.eL[v] = ex;
  }
#pragma unroll
  for (int j = 0; j < R; ++j)
    if (((pushed >> j) & 1u) && old[j] == 0) on_cross(wn[j]);
}

template <int R, class Cross>
__device__ __forceinline__ void x_discharge_r_reg(const XState& X, int w, int b, int e, int lane, double eps,
                                                  Cross on_cross) {
  int ex = X.eR[w];
  if (ex <= 0) return;
  double p = X.pR[w];
  double key[R];
  int vn[R], mr[R];
  unsigned ok = 0;  // bit j: the mirror of slot j carries flow (f == 1)
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int s = b + lane + XG * j;
    key[j] = 0.0;
    vn[j] = 0;
    mr[j] = 0;
    if (s < e) {
      vn[j] = X.nbr[s];
      mr[j] = X.mir[s];
    }
  }
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int s = b + lane + XG * j;
    if (s < e) {
      key[j] = X.c[s] + X.pL[vn[j]];
      if (X.f[mr[j]]) ok |= 1u << j;
    }
  }
  auto best = [&](double& bt, int& bs) {
    bt = -INFINITY;
    bs = INT_MAX;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int s = b + lane + XG * j;
      if (((ok >> j) & 1u) && (key[j] > bt || (key[j] == bt && s < bs))) {
        bt = key[j];
        bs = s;
      }
    }
    xg_best(bt, bs);
  };
  unsigned pushed = 0;
  int old[R];
#pragma unroll
  for (int j = 0; j < R; ++j) old[j] = 1;
  while (ex > 0) {
    double bt;
    int bs;
    best(bt, bs);
    if (bs == INT_MAX) break;
    if (!(bt - p > 0.0)) p = bt - eps;
    if ((bs - b) % XG == lane) {
      const int jb = (bs - b) / XG;
      ok &= ~(1u << jb);
      pushed |= 1u << jb;
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (j == jb) {
          X.f[mr[j]] = 0;
          old[j] = atomicAdd(&X.eL[vn[j]], 1);
        }
    }
    --ex;
  }
  {
    double bt;
    int bs;
    best(bt, bs);
    if (bs != INT_MAX && !(bt - p > 0.0)) p = bt - eps;
  }
  if (lane == 0) {
    X.pR[w] = p;
    X.eR[w] = ex;
  }
#pragma unroll
  for (int j = 0; j < R; ++j)
    if (((pushed >> j) & 1u) && old[j] == 0) on_cross(vn[j]);
}

// R > 0: the register form for rows of at most XG R slots, else the gathers
template <int R, class Cross>
__device__ __forceinline__ void x_discharge_l_any(const XState& X, int v, int b, int e, int lane, double eps,
                                                  Cross on_cross) {
  if (R > 0 && e - b <= XG * R) x_discharge_l_reg<(R > 0 ? R : 1)>(X, v, b, e, lane, eps, on_cross);
  else x_discharge_l(X, v, b, e, lane, eps, on_cross);
}
template <int R, class Cross>
__device__ __forceinline__ void x_discharge_r_any(const XState& X, int w, int b, int e, int lane, double eps,
                                                  Cross on_cross) {
  if (R > 0 && e - b <= XG * R) x_discharge_r_reg<(R > 0 ? R : 1)>(X, w, b, e, lane, eps, on_cross);
  else x_discharge_r(X, w, b, e, lane, eps, on_cross);
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
