// f2m_device.h -- device routines shared by the single-graph solve
// (f2m_solve.hip) and the batched persistent solve (f2m_batch.hip).
#pragma once
#include "f2m_internal.h"

namespace f2m {

// Active-list slot: the reference's EF array (graph_cuda.cu:229-252,
// get_cuda_EF graph_cuda_grad.hpp:96-111) holds pointers to the active edges;
// here the (D, nbr) pair itself is copied, 16 bytes, one vector load.
struct ESlot {
  double D;
  int32_t nbr;
  int32_t pad;
};

// Threads that run the least squares of one component: fixed, so that its
// reductions add in the same order whatever the launch's block size.
constexpr int AXB_T = 256;

__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block reductions over blockDim.x threads; every thread gets the result.
// sh: at least blockDim.x/64 doubles of shared memory.  Barriers inside.
__device__ __forceinline__ double block_sum_all(double v, double* sh) {
  v = wave_sum_d(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
  return s;
}
__device__ __forceinline__ double block_max_all(double v, double* sh) {
  v = wave_max_d(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = sh[0];
  for (int k = 1; k < (int)(blockDim.x >> 6); ++k) s = fmax(s, sh[k]);
  return s;
}

// ---- per-node steps (one thread per node), shared by both solve paths ------
// cuda_partition_v4 (graph_cuda_partition.hpp:89-115): threshold = max d of
// the first kk slots, partition the whole list (>= 3 below).
__device__ inline int part_v4_node(int v, int64_t b, int dg, int kk, int32_t* __restrict__ nbr,
                                   double* __restrict__ D, int32_t* __restrict__ orig,
                                   const double* __restrict__ F) {
  if (dg < kk) return dg;
  const double Fv = F[v];
  double thr;
  if (kk >= 1 && kk <= 8) {  // the first kk slots loaded together
    int w[8];
    double dd[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = i < kk ? i : 0;
      w[i] = nbr[b + k];
      dd[i] = D[b + k];
    }
    thr = dval(dd[0], Fv, F[w[0]], v, w[0]);
#pragma unroll
    for (int i = 1; i < 8; ++i) {
      if (i < kk) {
        const double d = dval(dd[i], Fv, F[w[i]], v, w[i]);
        thr = d > thr ? d : thr;
      }
    }
  } else {
    thr = dval(D[b], Fv, F[nbr[b]], v, nbr[b]);
    for (int i = 1; i < kk; ++i) {
      const int w = nbr[b + i];
      const double d = dval(D[b + i], Fv, F[w], v, w);
      thr = d > thr ? d : thr;
    }
  }
  return qpart_EF(v, b, dg, thr, nbr, D, orig, F);
}
// cuda_partition_v2 (graph_cuda_partition.hpp:57-69): lists shorter than 10
// stay whole.
__device__ inline int part_v2_node(int v, int64_t b, int dg, double thr, int32_t* __restrict__ nbr,
                                   double* __restrict__ D, int32_t* __restrict__ orig,
                                   const double* __restrict__ F) {
  return dg < 10 ? dg : qpart_EF(v, b, dg, thr, nbr, D, orig, F);
}
// The same partition with four lanes per node: lane q computes the d of
// slots q, q + 4, ... (a row's slots are loaded four at a time, consecutive),
// the threshold is the max of the first kk d's in slot order (kk <= 8; they
// sit in the lanes' first two chunks), the predicate bits are or-ed across
// the four lanes, and the lanes carry out qpart_mask's swaps for them
// between them (qpart_mask_lanes): the bits, swaps and pivot of
// part_v4_node (p >= kk >= 3, so qpart_EF's raise loop never runs).  A
// row's first 32 slots are read together.  Lists above 64 slots, kk < 3 or kk > 8 take
// part_v4_node on the first lane.  The
// four lanes of a node call together (shuffles of width 4); the result is
// valid on lane q = 0.  Lists of fewer than kk slots are the caller's.
constexpr int PV4_LANES = 4;
__device__ __attribute__((noinline)) int part_v4_long(int v, int64_t b, int dg, int kk, int32_t* __restrict__ nbr,
                                                      double* __restrict__ D, int32_t* __restrict__ orig,
                                                      const double* __restrict__ F) {
  return part_v4_node(v, b, dg, kk, nbr, D, orig, F);
}
__device__ __forceinline__ int part_v4_lanes(int v, int q, int64_t b, int dg, int kk, int32_t* __restrict__ nbr,
                                             double* __restrict__ D, int32_t* __restrict__ orig,
                                             const double* __restrict__ F) {
  if (dg > 64 || kk < 3 || kk > 8) return q == 0 ? part_v4_long(v, b, dg, kk, nbr, D, orig, F) : 0;
  const double Fv = F[v];
  // the row's first 32 slots read together (lane q: slots q, q + 4, ..., q + 28)
  constexpr int U = 8;
  double d[U];
  {
    int w[U];
    double dd[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = q + 4 * u < dg ? q + 4 * u : dg - 1;  // past the list: repeats the last slot, unused
      w[u] = nbr[b + k];
      dd[u] = D[b + k];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) d[u] = dval(dd[u], Fv, F[w[u]], v, w[u]);
  }
  double thr = 0.0;  // the max of the first kk d's in slot order: slots q (d[0]) and q + 4 (d[1])
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (i < kk) {
      const double di = __shfl(i < 4 ? d[0] : d[1], i & 3, PV4_LANES);
      thr = (i == 0 || di > thr) ? di : thr;
    }
  unsigned long long m = 0;
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (q + 4 * u < dg && d[u] <= thr) m |= 1ull << (q + 4 * u);
  for (int k0 = 4 * U; k0 < dg; k0 += 16) {  // the rest, 16 slots (four per lane) per step
    int w[4];
    double dd[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + q + 4 * u < dg ? k0 + q + 4 * u : dg - 1;
      w[u] = nbr[b + k];
      dd[u] = D[b + k];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + q + 4 * u;
      if (k < dg && dval(dd[u], Fv, F[w[u]], v, w[u]) <= thr) m |= 1ull << k;
    }
  }
  unsigned lo = (unsigned)m, hi = (unsigned)(m >> 32);
  lo |= __shfl_xor(lo, 1, PV4_LANES);
  hi |= __shfl_xor(hi, 1, PV4_LANES);
  lo |= __shfl_xor(lo, 2, PV4_LANES);
  hi |= __shfl_xor(hi, 2, PV4_LANES);
  return qpart_mask_lanes<PV4_LANES, 4>(b, dg, ((unsigned long long)hi << 32) | lo, q, nbr, D, orig);
}

// check_opt's partition (part_v2_node: lists of 10 or more slots at thr 1.0)
// with four lanes per node as s_part_v4_q.  Each pass: the four lanes read
// the row in its current order (four slots at a time), or their predicate
// bits and take the least d above thr (a min: the same in any order); the
// lanes swap (qpart_mask_lanes).  While fewer than 3 slots lie at or below
// thr, thr is raised (raise_thr, the same on every lane) and the permuted row
// is partitioned again, as qpart_EF does.  Lists above 64 slots take
// part_v2_node on the first lane.  Called
// as part_v4_lanes; lists of fewer than 10 slots are the caller's.
__device__ __attribute__((noinline)) int part_v2_long(int v, int64_t b, int dg, double thr,
                                                      int32_t* __restrict__ nbr, double* __restrict__ D,
                                                      int32_t* __restrict__ orig, const double* __restrict__ F) {
  return part_v2_node(v, b, dg, thr, nbr, D, orig, F);
}
__device__ __forceinline__ int part_v2_lanes(int v, int q, int64_t b, int dg, double thr, int32_t* __restrict__ nbr,
                                             double* __restrict__ D, int32_t* __restrict__ orig,
                                             const double* __restrict__ F) {
  if (dg > 64) return q == 0 ? part_v2_long(v, b, dg, thr, nbr, D, orig, F) : 0;
  const double Fv = F[v];
  for (int pass = 0;; ++pass) {
    unsigned long long m = 0;
    double ab = 1.0 / 0.0;
    {  // the row's first 32 slots read together (lane q: slots q, q + 4, ..., q + 28)
      constexpr int U = 8;
      int w[U];
      double dd[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = q + 4 * u < dg ? q + 4 * u : dg - 1;  // past the list: the last slot, unused
        w[u] = nbr[b + k];
        dd[u] = D[b + k];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = q + 4 * u;
        if (k < dg) {
          const double d = dval(dd[u], Fv, F[w[u]], v, w[u]);
          if (d <= thr) m |= 1ull << k;
          else ab = d < ab ? d : ab;
        }
      }
    }
    for (int k0 = 32; k0 < dg; k0 += 16) {  // the rest, 16 slots (four per lane) per step
      int w[4];
      double dd[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + q + 4 * u < dg ? k0 + q + 4 * u : dg - 1;  // past the list: the last slot, unused
        w[u] = nbr[b + k];
        dd[u] = D[b + k];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + q + 4 * u;
        if (k < dg) {
          const double d = dval(dd[u], Fv, F[w[u]], v, w[u]);
          if (d <= thr) m |= 1ull << k;
          else ab = d < ab ? d : ab;
        }
      }
    }
    unsigned lo = (unsigned)m, hi = (unsigned)(m >> 32);
#pragma unroll
    for (int o = 1; o < PV4_LANES; o <<= 1) {
      lo |= __shfl_xor(lo, o, PV4_LANES);
      hi |= __shfl_xor(hi, o, PV4_LANES);
      const double a2 = __shfl_xor(ab, o, PV4_LANES);
      ab = a2 < ab ? a2 : ab;
    }
    const int p = qpart_mask_lanes<PV4_LANES, 4>(b, dg, ((unsigned long long)hi << 32) | lo, q, nbr, D, orig);
    if (p >= 3) return p;
    __threadfence_block();  // the lanes' swaps, before they read the row again
    thr = raise_thr(thr, ab);
  }
}

// cuda_partition_v3 at 0.25 over the active prefix (graph_cuda_comb.hpp:301)
// and the class counts over that prefix (graph_cuda_grad.hpp:18-39):
// A = #{d < -eps} (x = 1), T = #{-eps <= d <= eps} (tight).
__device__ inline void classify_node(int v, int64_t b, int act, double eps, int32_t* __restrict__ nbr,
                                     double* __restrict__ D, int32_t* __restrict__ orig,
                                     const double* __restrict__ F, int& cl, int& A, int& T) {
  cl = qpart_B(v, b, act, 0.25, nbr, D, orig, F);
  const double Fv = F[v];
  int a = 0, t = 0;
  for (int k = 0; k < cl; ++k) {
    const int w = nbr[b + k];
    const double d = dval(D[b + k], Fv, F[w], v, w);
    if (d < -eps) ++a;
    else if (d <= eps) ++t;
  }
  A = a;
  T = t;
}
// x per working slot: 1 on d < -eps (node::get_eA_X, node.cpp:61-71), 1/2 on
// the tight slots of a regular component (comb::solve_simple,
// comb_thread.cpp:180-195), 0 elsewhere (irregular components are filled by
// the least squares).  flag marks the slots of the reference's X map.
__device__ inline void primal_node(int v, int64_t b, int dg, int cl, double eps, const int32_t* __restrict__ nbr,
                                   const double* __restrict__ D, const double* __restrict__ F, bool regular,
                                   double* __restrict__ xw, uint8_t* __restrict__ flag) {
  const double Fv = F[v];
  for (int k = 0; k < dg; ++k) {
    double x = 0.0;
    uint8_t f = 0;
    if (k < cl) {
      const int w = nbr[b + k];
      const double d = dval(D[b + k], Fv, F[w], v, w);
      if (d < -eps) { x = 1.0; f = 1; }
      else if (d <= eps) { x = regular ? 0.5 : 0.0; f = 1; }
    }
    xw[b + k] = x;
    flag[b + k] = f;
  }
}

// Union-find with par[x] <= x: hooking always puts the larger root under the
// smaller, so each root is the smallest id of its component.
__device__ __forceinline__ int uf_find(int* par, int x) {
  volatile int* vp = par;
  int p = vp[x];
  while (p != x) {
    const int gp = vp[p];
    if (gp != p) vp[x] = gp;  // path halving; any ancestor is a valid parent
    x = p;
    p = vp[x];
  }
  return x;
}
__device__ __forceinline__ void uf_unite(int* par, int a, int c) {
  for (;;) {
    a = uf_find(par, a);
    c = uf_find(par, c);
    if (a == c) return;
    if (a < c) { int t = a; a = c; c = t; }
    if (atomicCAS(&par[a], a, c) == a) return;
  }
}

// Ax_b::solve (Ax_b.cpp:139-158) on one irregular tight component, executed
// by the whole block.  nodes: the component's nodes in ascending id order.
// Equation of node v: sum over its tight slots s of x_s = 2 - A(v), row
// norm sqrt(T(v)).  Variables are tight edges; the two slots of an edge hold
// the same value because each update uses the same expression in the
// canonical (lo, hi) order.  x0 = 0; steepest descent with exact line
// search; stop when max |r_j| < 1e-8; at most 999 iterations; x stored
// rounded to float (XB is float, F2M.h:202).  Returns go_max (0 = failed).
// rp64/base: slot offsets of node v are [rp(v), rp(v)+clen[v]).
// Only threads [0, AXB_T) take nodes; the others join the barriers and add
// zeros to the reductions, which leaves the sums and maxima unchanged.
template <class RP>
__device__ int axb_comb(const int32_t* __restrict__ nodes, int m, RP rp, const int32_t* __restrict__ clen,
                        const int32_t* __restrict__ A, const int32_t* __restrict__ T, double eps,
                        const int32_t* __restrict__ nbr, const double* __restrict__ D,
                        const double* __restrict__ F, double* __restrict__ xw, double* __restrict__ xg,
                        double* __restrict__ rn, double* sh) {
  auto tight = [&](int v, int64_t s, double Fv) {
    const int w = nbr[s];
    const double d = dval(D[s], Fv, F[w], v, w);
    return d >= -eps && d <= eps;
  };
  for (int i = threadIdx.x < AXB_T ? (int)threadIdx.x : m; i < m; i += AXB_T) {
    const int v = nodes[i];
    const int64_t b = rp(v);
    const double Fv = F[v];
    for (int k = 0; k < clen[v]; ++k)
      if (tight(v, b + k, Fv)) xw[b + k] = 0.0;
  }
  __syncthreads();
  int go_max = 0;
  for (int go = 1; go < 1000; ++go) {
    double dmax = 0.0;
    for (int i = threadIdx.x < AXB_T ? (int)threadIdx.x : m; i < m; i += AXB_T) {
      const int v = nodes[i];
      const int64_t b = rp(v);
      const double Fv = F[v];
      double s = 0.0;
      for (int k = 0; k < clen[v]; ++k)
        if (tight(v, b + k, Fv)) s += xw[b + k];
      const double r = (s - (double)(2 - A[v])) / sqrt((double)T[v]);
      rn[v] = r;
      dmax = fmax(dmax, fabs(r));
    }
    dmax = block_max_all(dmax, sh);  // barriers publish rn
    if (dmax < 1e-8) { go_max = go; break; }
    for (int i = threadIdx.x < AXB_T ? (int)threadIdx.x : m; i < m; i += AXB_T) {
      const int v = nodes[i];
      const int64_t b = rp(v);
      const double Fv = F[v];
      for (int k = 0; k < clen[v]; ++k) {
        if (!tight(v, b + k, Fv)) continue;
        const int w = nbr[b + k];
        const int lo = v < w ? v : w, hi = v < w ? w : v;
        xg[b + k] = rn[lo] / sqrt((double)T[lo]) + rn[hi] / sqrt((double)T[hi]);
      }
    }
    __syncthreads();
    double gg = 0.0, gd = 0.0;
    for (int i = threadIdx.x < AXB_T ? (int)threadIdx.x : m; i < m; i += AXB_T) {
      const int v = nodes[i];
      const int64_t b = rp(v);
      const double Fv = F[v];
      double bg = 0.0;
      for (int k = 0; k < clen[v]; ++k)
        if (tight(v, b + k, Fv)) bg += xg[b + k];
      gg += bg * bg;
      gd += rn[v] * bg;
    }
    gg = block_sum_all(gg, sh);
    gd = block_sum_all(gd, sh);
    if (gg == 0.0) break;  // the reference would take 0/0 = NaN here; treated as no convergence
    const double alfa = -gd / gg;
    for (int i = threadIdx.x < AXB_T ? (int)threadIdx.x : m; i < m; i += AXB_T) {
      const int v = nodes[i];
      const int64_t b = rp(v);
      const double Fv = F[v];
      for (int k = 0; k < clen[v]; ++k)
        if (tight(v, b + k, Fv)) xw[b + k] += xg[b + k] * alfa;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x < AXB_T ? (int)threadIdx.x : m; i < m; i += AXB_T) {
    const int v = nodes[i];
    const int64_t b = rp(v);
    const double Fv = F[v];
    for (int k = 0; k < clen[v]; ++k)
      if (tight(v, b + k, Fv)) xw[b + k] = (double)(float)xw[b + k];
  }
  __syncthreads();
  return go_max;
}

}  // namespace f2m
