// f2m_exact.hip -- exact primal recovery (no reference counterpart).
//
// Replaces the reference's least-squares recovery and dual-perturbation loop
// (comb_thread.cpp:272-304, Ax_b.cpp:139-158, graph_cuda.cu:69-126) where they
// run (irregular tight components: float least squares, often x outside
// [0,1] on tie-heavy instances) or fail (status 1), or always on request.
//
// The fractional 2-matching LP equals half the min-cost transportation
// problem on the bipartite double cover: L_v (supply 2) -> R_w (demand 2),
// arc capacity 1, cost c_vw, both orientations of every edge;
// x_e = (f_vw + f_wv) / 2.  It is solved by cost-scaling push-relabel with
// eps-optimality  rc(v,w) = (c + pL[v]) - pR[w] >= -eps  on residual arcs.
//
// Schedule (synchronous, hence deterministic and equal bit for bit to the
// oracle's restatement): one round = an L step over every left node with
// positive excess, then an R step over every right node with positive
// excess.  No arc joins two nodes of one side, so a step is a valid sequence
// of push/relabel operations in any order.  A node discharges completely:
//   L_v: repeatedly the residual arc of largest pR[w] - c (ties: lowest slot);
//        relabel pL = (pR[w] - c) - eps when it is not admissible; push 1.
//   R_w: repeatedly the in-flow arc of largest c + pL[v] (ties: lowest slot);
//        relabel pR = (c + pL[v]) - eps when not admissible; return 1.
// then relabels once more when no admissible arc is left (a push becomes an
// auction bid with the full margin to the next-best arc).
// Phases: prices from the dual descent (pL = d_min F/2, pR = -pL, rc = d_min
// d_e), f = [rc < 0]; eps = eps0 d_min, divided by alpha per phase, with the
// violating arcs saturated or emptied at each new eps.  After each phase the
// duality gap P - D of x and y = (pR - pL)/2 is summed in a fixed order; the
// stage stops when it is within tolerance (< 1/2 for integer costs, where
// the LP optimum is a multiple of 1/2 and P is then exactly optimal).
//
// Layout.  Per slot (canonical CSR): flow byte f, mirror slot index.  Per
// node: pL, pR (double), excess eL, eR (int32).  Active lists: actL (two
// buffers, by round parity), actR, with counters in a ring of 4 rounds;
// a node joins the other side's list when its excess crosses 0 -> 1.
// Discharge steps: 16 lanes per node, grid-stride over the active list.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>

#include "f2m_exact_device.h"

namespace f2m {
namespace {

constexpr int XBS = 256;    // threads per block
constexpr int XGRID = 512;  // blocks of a discharge step
constexpr int RING = 4;     // counter ring, rounds mod 4

struct XDev {
  XState S;
  int n;
  const int32_t* rp;
  int32_t* actL0;
  int32_t* actL1;
  int32_t* actR;
  int32_t* cnt;                  // [2*RING]: nL, nR of round r at 2*(r%RING)
  unsigned long long* rounds;    // rounds with an active node
};

struct RP32 {
  const int32_t* rp;
  __device__ __forceinline__ int64_t operator()(int v) const { return rp[v]; }
};

// XM_LANES lanes per node, lane l searching the mirrors of slots l, l + 16,
// ... (each slot's search is independent; a thread per node ran ~23 binary
// searches in a row: 158 us at n = 100k)
constexpr int XM_LANES = 16;
__global__ void x_mirror(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ nbr,
                         const double* __restrict__ c, int32_t* __restrict__ mir, int32_t* __restrict__ flag) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = (int)(t / XM_LANES), l = (int)(t % XM_LANES);
  if (v >= n) return;
  const int b = rp[v], e = rp[v + 1];
  if (l == 0 && e - b < 2) atomicOr(&flag[1], 1);
  bool ok = true, nonint = false;
  for (int s = b + l; s < e; s += XM_LANES) ok = x_mirror_slot(v, s, RP32{rp}, nbr, c, mir, nonint) && ok;
  if (!ok) atomicOr(&flag[0], 1);
  if (nonint) atomicOr(&flag[2], 1);
}

__global__ void x_prices(int n, const double* __restrict__ F, double dmin, double* __restrict__ pL,
                         double* __restrict__ pR) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const double p = 0.5 * (dmin * F[v]);
  pL[v] = p;
  pR[v] = -p;
}

// The initial flows f = [rc < 0] with XF_LANES lanes per node (lane l:
// slots l, l + 8, ...; the arithmetic of x_flow_node<false>), and the
// excesses from them (lane sums of the flows out and in, then the first lane)
constexpr int XF_LANES = 8;
__global__ void x_flow_init_lanes(XDev X) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = (int)(t / XF_LANES), l = (int)(t % XF_LANES);
  if (v >= X.n) return;  // the lanes of a node leave together
  const XState& S = X.S;
  const int b = X.rp[v], e = X.rp[v + 1];
  const double pv = S.pL[v];
  for (int s = b + l; s < e; s += XF_LANES) {
    const double r = (S.c[s] + pv) - S.pR[S.nbr[s]];
    S.f[s] = r < 0.0 ? 1 : 0;
  }
}
__global__ void x_count_lanes(XDev X) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = (int)(t / XF_LANES), l = (int)(t % XF_LANES);
  if (v >= X.n) return;
  const XState& S = X.S;
  const int b = X.rp[v], e = X.rp[v + 1];
  int out = 0, in = 0;
  for (int s = b + l; s < e; s += XF_LANES) {
    out += S.f[s];
    in += S.f[S.mir[s]];
  }
#pragma unroll
  for (int o = 1; o < XF_LANES; o <<= 1) {
    out += __shfl_xor(out, o, XF_LANES);
    in += __shfl_xor(in, o, XF_LANES);
  }
  if (l == 0) {
    S.eL[v] = 2 - out;
    S.eR[v] = in - 2;
  }
}

// Later phases: the arcs violating eps-optimality at the new eps saturated
// or emptied (x_flow_node<true, true>) with XF_LANES lanes per node: lane l
// flips the arcs of slots l, l + 8, ... (each arc's test and flip are its
// own; eR moves by atomics as there) and the lanes' changes of eL[v],
// integers, are summed with shuffles before the first lane applies them.
__global__ void x_flow_sat_lanes(XDev X, double eps) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = (int)(t / XF_LANES), l = (int)(t % XF_LANES);
  if (v >= X.n) return;  // the lanes of a node leave together
  const XState& S = X.S;
  const int b = X.rp[v], e = X.rp[v + 1];
  const double pv = S.pL[v];
  int dl = 0;
  for (int s = b + l; s < e; s += XF_LANES) {
    const double r = (S.c[s] + pv) - S.pR[S.nbr[s]];
    const uint8_t fs = S.f[s];
    if (!fs && r < -eps) {
      S.f[s] = 1;
      --dl;
      atomicAdd(&S.eR[S.nbr[s]], 1);
    } else if (fs && r > eps) {
      S.f[s] = 0;
      ++dl;
      atomicAdd(&S.eR[S.nbr[s]], -1);
    }
  }
#pragma unroll
  for (int o = 1; o < XF_LANES; o <<= 1) dl += __shfl_xor(dl, o, XF_LANES);
  if (l == 0 && dl) S.eL[v] += dl;
}

// The stage's counters for the host (XPOST_W words from X.cnt: the ring,
// flags, rounds, the last x_rounds_wg's round count, the certificate sums),
// stored by one lane to pinned host memory with system scope, then seq with
// a release: a host that reads seq there reads the values too.
constexpr int XPOST_W = 12;
__global__ void x_post(const unsigned long long* __restrict__ src, unsigned long long* __restrict__ host,
                       unsigned long long seq) {
  if (threadIdx.x == 0) {
    unsigned long long v[XPOST_W];
#pragma unroll
    for (int k = 0; k < XPOST_W; ++k) v[k] = src[k];
#pragma unroll
    for (int k = 0; k < XPOST_W; ++k) __hip_atomic_store(&host[k], v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&host[XPOST_W], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__global__ void x_zero_ring(int32_t* cnt) {
  if (threadIdx.x < 2 * RING) cnt[threadIdx.x] = 0;
}

// Entries of the L and R lists of a workgroup's threads: each wave's lanes
// with a flag take consecutive slots of a shared-memory count, then one
// atomic per workgroup and list reserves them in the global list (one per
// wave had ~1,500 waves' atomics queue on two addresses at n = 100k).
// Every thread of the workgroup calls it; returns the global indices (-1:
// no entry).
__device__ __forceinline__ int2 block_append2(int32_t* cl, int32_t* cr, bool pl, bool pr) {
  __shared__ int s_n[2], s_base[2];
  if (threadIdx.x < 2) s_n[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long ml = __ballot(pl), mr = __ballot(pr);
  int ol = 0, orr = 0;
  if (lane == 0) {
    if (ml) ol = atomicAdd(&s_n[0], __popcll(ml));
    if (mr) orr = atomicAdd(&s_n[1], __popcll(mr));
  }
  ol = __shfl(ol, 0) + __popcll(ml & below);
  orr = __shfl(orr, 0) + __popcll(mr & below);
  __syncthreads();
  if (threadIdx.x == 0) s_base[0] = s_n[0] ? atomicAdd(cl, s_n[0]) : 0;
  if (threadIdx.x == 1) s_base[1] = s_n[1] ? atomicAdd(cr, s_n[1]) : 0;
  __syncthreads();
  return make_int2(pl ? s_base[0] + ol : -1, pr ? s_base[1] + orr : -1);
}

// the active lists of round r from the excesses (counted by x_count_lanes
// in the first phase, then maintained); every thread of the launch runs
// the appends (whole workgroups), so no early return
constexpr int XE_T = 1024;
__global__ void __launch_bounds__(XE_T) x_excess(XDev X, int64_t r) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  int el = 0, er = 0;
  if (v < X.n) {
    el = X.S.eL[v];
    er = X.S.eR[v];
  }
  const int k = (int)(r % RING);
  const int2 ix = block_append2(&X.cnt[2 * k], &X.cnt[2 * k + 1], el > 0, er > 0);
  if (ix.x >= 0) ((r & 1) ? X.actL1 : X.actL0)[ix.x] = v;
  if (ix.y >= 0) X.actR[ix.y] = v;
}

// A workgroup's appends to a global list: collected in shared memory, then
// one atomic on the list's count per workgroup (a step's appends otherwise
// all go to one address); past XQ entries they go to the list directly.
// The lists' order does not matter to the steps that read them.
constexpr int XQ = 512;
struct BlockQueue {
  int32_t* q;  // [XQ] shared
  int* n;      // shared count
  __device__ __forceinline__ void push(int32_t* list, int32_t* cnt, int v) {
    const int k = atomicAdd(n, 1);
    if (k < XQ) q[k] = v;
    else list[atomicAdd(cnt, 1)] = v;
  }
  // every thread of the workgroup calls it, after its pushes
  __device__ __forceinline__ void flush(int32_t* list, int32_t* cnt, int* base) {
    __syncthreads();
    const int m = min(*n, XQ);
    if (threadIdx.x == 0 && m > 0) *base = atomicAdd(cnt, m);
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += blockDim.x) list[*base + i] = q[i];
  }
};

// L step of round r: every listed left node discharges; right nodes whose
// excess crosses 0 -> 1 join the R list of round r.  XR > 0: rows of at
// most XG XR slots discharge from registers (x_discharge_l_reg), and the
// appends go through a BlockQueue.
template <int XR>
__global__ void __launch_bounds__(XBS) x_lstep(XDev X, int64_t r, double eps) {
  __shared__ int32_t sq[XQ];
  __shared__ int sn, sbase;
  const int k = (int)(r % RING), k1 = (int)((r + 1) % RING);
  if (blockIdx.x == 0 && threadIdx.x == 0) X.cnt[2 * k1] = 0;  // nL of round r+1, filled by R(r)
  const int na = X.cnt[2 * k];
  const int32_t* list = (r & 1) ? X.actL1 : X.actL0;
  const int lane = threadIdx.x % XG;
  const int groups = gridDim.x * (XBS / XG);
  int32_t* cr = &X.cnt[2 * k + 1];
  int32_t* lr = X.actR;
  BlockQueue bq{sq, &sn};
  if (XR > 0) {
    if (threadIdx.x == 0) sn = 0;
    __syncthreads();
  }
  for (int i = blockIdx.x * (XBS / XG) + threadIdx.x / XG; i < na; i += groups) {
    const int v = list[i];
    x_discharge_l_any<XR>(X.S, v, X.rp[v], X.rp[v + 1], lane, eps, [&](int w) {
      if (XR > 0) bq.push(lr, cr, w);
      else lr[atomicAdd(cr, 1)] = w;
    });
  }
  if (XR > 0) bq.flush(lr, cr, &sbase);
}

// R step of round r: every listed right node discharges; left nodes whose
// excess crosses 0 -> 1 join the L list of round r+1
template <int XR>
__global__ void __launch_bounds__(XBS) x_rstep(XDev X, int64_t r, double eps) {
  __shared__ int32_t sq[XQ];
  __shared__ int sn, sbase;
  const int k = (int)(r % RING), k1 = (int)((r + 1) % RING);
  const int na = X.cnt[2 * k + 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    X.cnt[2 * k1 + 1] = 0;  // nR of round r+1, filled by L(r+1)
    if (X.cnt[2 * k] + na > 0) atomicAdd(X.rounds, 1ull);
  }
  int32_t* next = (r & 1) ? X.actL0 : X.actL1;
  int32_t* cl = &X.cnt[2 * k1];
  const int lane = threadIdx.x % XG;
  const int groups = gridDim.x * (XBS / XG);
  BlockQueue bq{sq, &sn};
  if (XR > 0) {
    if (threadIdx.x == 0) sn = 0;
    __syncthreads();
  }
  for (int i = blockIdx.x * (XBS / XG) + threadIdx.x / XG; i < na; i += groups) {
    const int w = X.actR[i];
    x_discharge_r_any<XR>(X.S, w, X.rp[w], X.rp[w + 1], lane, eps, [&](int v) {
      if (XR > 0) bq.push(next, cl, v);
      else next[atomicAdd(cl, 1)] = v;
    });
  }
  if (XR > 0) bq.flush(next, cl, &sbase);
}

// Rounds r0, r0+1, ... of a phase inside one workgroup, used once the active
// lists have shrunk (a round of x_lstep + x_rstep is then mostly launch
// overhead): the same steps, barriers for the launch boundaries, the round's
// counters in shared memory, the lists in the same global arrays.  Stops
// after the first round without an active node, after max_rounds rounds with
// one, or when the next L list reaches `small` nodes (back to the grid
// steps); runs no round when round r0's L list already has `small`.  On exit the ring holds the counts of the last round and of the
// next one (as the grid steps leave it), X.rounds is advanced and *ran =
// rounds run.  The steps' outcome does not depend on the order in which a
// step's nodes are processed, so the bits equal the grid steps'.
constexpr int XWG = 1024;
template <int XR>
__global__ void __launch_bounds__(XWG) x_rounds_wg(XDev X, int64_t r0, double eps, int64_t max_rounds, int small,
                                                   int32_t* ran) {
  __shared__ int sc[3];  // nL(r), nR(r), nL(r+1)
  const int lane = threadIdx.x % XG, gid = threadIdx.x / XG;
  constexpr int NG = XWG / XG;
  if (threadIdx.x == 0) {
    sc[0] = X.cnt[2 * (int)(r0 % RING)];
    sc[1] = X.cnt[2 * (int)(r0 % RING) + 1];
    sc[2] = 0;
  }
  __syncthreads();
  if (sc[0] >= small) {  // lists still long: back to the grid steps, ring untouched
    if (threadIdx.x == 0) *ran = 0;
    return;
  }
  int64_t r = r0, counted = 0;
  for (;;) {
    const int nl = sc[0];
    const int32_t* list = (r & 1) ? X.actL1 : X.actL0;
    for (int i = gid; i < nl; i += NG) {
      const int v = list[i];
      x_discharge_l_any<XR>(X.S, v, X.rp[v], X.rp[v + 1], lane, eps,
                            [&](int w) { X.actR[atomicAdd(&sc[1], 1)] = w; });
    }
    __syncthreads();
    const int nr = sc[1];
    int32_t* next = (r & 1) ? X.actL0 : X.actL1;
    for (int i = gid; i < nr; i += NG) {
      const int w = X.actR[i];
      x_discharge_r_any<XR>(X.S, w, X.rp[w], X.rp[w + 1], lane, eps, [&](int v) { next[atomicAdd(&sc[2], 1)] = v; });
    }
    __syncthreads();
    const int nl1 = sc[2];
    const bool active = nl + nr > 0;
    if (active) ++counted;
    ++r;
    if (!active || counted >= max_rounds || nl1 >= small) {
      if (threadIdx.x == 0) {
        const int kl = (int)((r - 1) % RING), kn = (int)(r % RING);
        X.cnt[2 * kl] = nl;
        X.cnt[2 * kl + 1] = nr;
        X.cnt[2 * kn] = nl1;
        X.cnt[2 * kn + 1] = 0;
        atomicAdd(X.rounds, (unsigned long long)counted);
        *ran = (int32_t)(r - r0);
      }
      break;
    }
    __syncthreads();  // every thread has read the counters
    if (threadIdx.x == 0) {
      sc[0] = nl1;
      sc[1] = 0;
      sc[2] = 0;
    }
    __syncthreads();
  }
}

// per-node terms of the duality gap certificate
// Certificate sums in a fixed order: DET_T contiguous chunks of nodes, each
// chunk's node terms (x_cert_node: slots in order) summed in node order from
// 0, then the chunk sums in chunk order -- the oracle's order, so the same
// bits; both the primal P and the dual bound D.  One workgroup per chunk:
// XC_LANES lanes per node compute the terms of slots q, q + XC_LANES, ...
// together (their gathers in flight at once); the node's first lane adds
// them in slot order (a round of XC_LANES slots at a time, read with
// shuffles; a slot with w <= v adds +0.0, which leaves these sums of
// non-negative terms, never -0.0, unchanged), thread 0 the chunk's node
// values in node order.
constexpr int XC_LANES = 8, XC_T = 256, XC_G = XC_T / XC_LANES;
__global__ void __launch_bounds__(XC_T) x_cert_chunks(XDev X, double* __restrict__ part) {
  __shared__ double na[XC_G], nd[XC_G];
  const int tid = threadIdx.x, q = tid % XC_LANES, g = tid / XC_LANES;
  const int64_t chunk = ((int64_t)X.n + DET_T - 1) / DET_T;
  const int64_t b = (int64_t)blockIdx.x * chunk;
  const int64_t e = b + chunk < X.n ? b + chunk : X.n;
  const XState& S = X.S;
  double A = 0.0, Dc = 0.0;  // the chunk's sums (thread 0)
  for (int64_t v0 = b; v0 < e; v0 += XC_G) {
    const int64_t v = v0 + g;
    const bool ok = v < e;
    const int s0 = ok ? X.rp[v] : 0, s1 = ok ? X.rp[v + 1] : 0;
    const double yv = ok ? 0.5 * (S.pR[v] - S.pL[v]) : 0.0;
    double a = 0.0, bsum = 0.0;
    const int rounds = (s1 - s0 + XC_LANES - 1) / XC_LANES;
    // every lane of the wave runs the wave's most rounds (the shuffles need all lanes)
    int rmax = rounds;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) rmax = max(rmax, __shfl_xor(rmax, off, 64));
    for (int r = 0; r < rmax; ++r) {
      const int s = s0 + r * XC_LANES + q;
      double ta = 0.0, tb = 0.0;
      if (s < s1) {
        const int w = S.nbr[s];
        if (w > v) {
          const double x = 0.5 * (double)(S.f[s] + S.f[S.mir[s]]);
          ta = S.c[s] * x;
          const double yw = 0.5 * (S.pR[w] - S.pL[w]);
          tb = fmax(0.0, (yv + yw) - S.c[s]);
        }
      }
      const int base = (threadIdx.x & 63) - q;  // the node's first lane within the wave
#pragma unroll
      for (int k = 0; k < XC_LANES; ++k) {
        const double xa = __shfl(ta, base + k, 64);
        const double xb = __shfl(tb, base + k, 64);
        a += xa;
        bsum += xb;
      }
    }
    if (q == 0) {
      na[g] = a;
      nd[g] = 2.0 * yv - bsum;
    }
    __syncthreads();
    if (tid == 0) {
      const int cnt = (int)(e - v0 < XC_G ? e - v0 : XC_G);
      for (int k = 0; k < cnt; ++k) {
        A += na[k];
        Dc += nd[k];
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    part[blockIdx.x] = A;
    part[DET_T + blockIdx.x] = Dc;
  }
}
// P = sum of the chunk sums in chunk order, D likewise: the sums are read
// into shared memory with coalesced loads, then added in order by thread 0
__global__ void __launch_bounds__(256) x_det_final(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sp[2 * DET_T];
  for (int k = threadIdx.x; k < 2 * DET_T; k += 256) sp[k] = part[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    double P = 0.0, D = 0.0;
    for (int k = 0; k < DET_T; ++k) {
      P += sp[k];
      D += sp[DET_T + k];
    }
    out[0] = P;
    out[1] = D;
  }
}

// outputs: x and its X-flag per canonical slot, F = (pL - pR) / d_min
// (a thread per slot, and per node for F)
__global__ void x_out(XDev X, int64_t m, double dmin, double* __restrict__ xc, uint8_t* __restrict__ fc,
                      double* __restrict__ F) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < m) {
    const double x = 0.5 * (double)(X.S.f[s] + X.S.f[X.S.mir[s]]);
    xc[s] = x;
    fc[s] = x > 0.0 ? 1 : 0;
  }
  if (s < X.n) F[s] = (X.S.pL[s] - X.S.pR[s]) / dmin;
}
__global__ void x_uf_init(int n, int32_t* par) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) par[v] = v;
}
// XF_LANES lanes per node (lane l: slots l, l + 8, ...); the unions link
// the larger root under the smaller, so the roots do not depend on their order
__global__ void x_half_union(XDev X, int32_t* par) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = (int)(t / XF_LANES), l = (int)(t % XF_LANES);
  if (v >= X.n) return;
  for (int s = X.rp[v] + l; s < X.rp[v + 1]; s += XF_LANES) {
    const int w = X.S.nbr[s];
    if (w > v && X.S.f[s] + X.S.f[X.S.mir[s]] == 1) uf_unite(par, v, w);
  }
}
__global__ void __launch_bounds__(1024) x_half_root(XDev X, int32_t* par, int32_t* __restrict__ root,
                                                    int32_t* __restrict__ nroots) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int v = (int)(t / XF_LANES), l = (int)(t % XF_LANES);
  bool has = false;
  if (v < X.n)
    for (int s = X.rp[v] + l; s < X.rp[v + 1]; s += XF_LANES) has = has || X.S.f[s] + X.S.f[X.S.mir[s]] == 1;
  // any lane of the node
  const unsigned long long any = __ballot(has);
  const int g0 = (threadIdx.x & 63) & ~(XF_LANES - 1);
  has = ((any >> g0) & ((1ull << XF_LANES) - 1ull)) != 0;
  int r = -1;
  if (v < X.n && l == 0) {
    r = has ? uf_find(par, v) : -1;
    root[v] = r;
  }
  const unsigned long long nr = __ballot(v < X.n && l == 0 && r == v);
  if ((threadIdx.x & 63) == 0 && nr) atomicAdd(&cnt, __popcll(nr));
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(nroots, cnt);  // one per workgroup
}

__global__ void x_zero_i32(int32_t* p, int k) {
  if ((int)threadIdx.x < k) p[threadIdx.x] = 0;
}

}  // namespace

f2m_status exact_stage(f2m_ctx* c, SolveWs& w, int n, int64_t m, const int32_t* rp, const int32_t* nbr,
                       const double* cost, double dmin, const f2m_opts& o, double* F, double* xc, uint8_t* fc,
                       int32_t* root, f2m_stats& z, int& rc) {
  F2M_HIP(c, w.xmir.reserve(m * 4));
  F2M_HIP(c, w.xf.reserve(m));
  F2M_HIP(c, w.xp.reserve((int64_t)n * 8 * 4 + 2 * DET_T * 8));  // pL, pR, terms (2n), chunk sums
  F2M_HIP(c, w.xe.reserve((int64_t)n * 4 * 2));
  F2M_HIP(c, w.xact.reserve((int64_t)n * 4 * 3));
  F2M_HIP(c, w.xcnt.reserve(64 * 8));
  XDev X;
  X.n = n;
  X.rp = rp;
  X.S.nbr = nbr;
  X.S.c = cost;
  X.S.mir = w.xmir.as<int32_t>();
  X.S.f = w.xf.as<uint8_t>();
  X.S.pL = w.xp.as<double>();
  X.S.pR = X.S.pL + n;
  double* pv = X.S.pR + n;
  double* dv = pv + n;
  double* part = dv + n;
  X.S.eL = w.xe.as<int32_t>();
  X.S.eR = X.S.eL + n;
  X.actL0 = w.xact.as<int32_t>();
  X.actL1 = X.actL0 + n;
  X.actR = X.actL1 + n;
  X.cnt = w.xcnt.as<int32_t>();                  // [0, 8) ring
  int32_t* flag = X.cnt + 8;                     // [8, 12) flags, [12] root count
  X.rounds = reinterpret_cast<unsigned long long*>(X.cnt + 16);
  double* sums = reinterpret_cast<double*>(X.cnt + 20);
  const dim3 gn(grid_for(n, XBS)), bs(XBS);
  hipStream_t st = c->stream;
  // the host's copy of X.cnt [0, 24): ring, flags, rounds, ran, sums
  struct Post {
    int32_t cnt[8];
    int32_t flag[8];
    unsigned long long rounds;
    int32_t ran;  // rounds run by the last x_rounds_wg
    int32_t pad;
    double P, D;
  };
  static_assert(sizeof(Post) == XPOST_W * 8, "x_post copies the Post words");
  unsigned long long* hpost = static_cast<unsigned long long*>(c->pinned) + 3072;
  Post hv;
  // waits for the kernels so far and reads X.cnt's words (x_post; a stream
  // synchronisation after 50 ms, which also reports any error)
  auto post = [&]() -> f2m_status {
    const unsigned long long want = ++c->read_seq;
    hipLaunchKernelGGL(x_post, dim3(1), dim3(64), 0, st, reinterpret_cast<const unsigned long long*>(X.cnt), hpost,
                       want);
    F2M_LAUNCH_CHECK(c);
    bool seen = false;
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned it = 1;; ++it) {
      if (__atomic_load_n(&hpost[XPOST_W], __ATOMIC_ACQUIRE) == want) {
        seen = true;
        break;
      }
      if ((it & 1023u) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) break;
    }
    if (!seen) F2M_HIP(c, hipStreamSynchronize(st));
    unsigned long long wv[XPOST_W];
    for (int k = 0; k < XPOST_W; ++k) wv[k] = __atomic_load_n(&hpost[k], __ATOMIC_ACQUIRE);
    std::memcpy(&hv, wv, sizeof(hv));
    return F2M_OK;
  };
  hipLaunchKernelGGL(x_zero_i32, dim3(1), dim3(64), 0, st, X.cnt, 20);
  hipLaunchKernelGGL(x_mirror, dim3(grid_for((int64_t)n * XM_LANES, bs.x)), bs, 0, st, n, rp, nbr, cost,
                     w.xmir.as<int32_t>(), flag);
  F2M_LAUNCH_CHECK(c);
  {
    f2m_status s = post();
    if (s) return s;
  }
  z.exact_ran = 1;
  if (hv.flag[0]) { rc = -1; return F2M_OK; }
  if (hv.flag[1]) { rc = -2; return F2M_OK; }
  const bool integral = hv.flag[2] == 0;
  hipLaunchKernelGGL(x_prices, gn, bs, 0, st, n, F, dmin, X.S.pL, X.S.pR);
  const dim3 gl(grid_for((int64_t)n * XF_LANES, bs.x));
  hipLaunchKernelGGL(x_flow_init_lanes, gl, bs, 0, st, X);
  F2M_LAUNCH_CHECK(c);
  double eps = o.exact_eps0 * dmin;
  int64_t r = 0;          // rounds launched (ring / parity index)
  int64_t done = 0;       // rounds with an active node
  int phases = 0;
  bool stuck = false, certified = false;
  double P = 0, D = 0;
  int32_t* ran = reinterpret_cast<int32_t*>(X.rounds + 1);
  // lists below this many nodes run their rounds inside one workgroup
  const int small = env_int("F2M_EXACT_WG", 512);
  // rows of at most 64 slots discharge from registers (0: gather per push)
  const bool reg = env_int("F2M_EXACT_REG", 1) != 0;
  // a phase's first pass: fuse_k grid rounds, the in-workgroup rounds and
  // the certificate, launched together with one wait (0: a wait after each)
  const int fuse_k = small > 0 ? env_int("F2M_EXACT_FUSE", 6) : 0;
  bool fuse = fuse_k > 0;
  auto grid_rounds = [&](int64_t kk) {
    for (int64_t i = 0; i < kk; ++i, ++r) {
      if (reg) {
        hipLaunchKernelGGL(x_lstep<4>, dim3(XGRID), dim3(XBS), 0, st, X, r, eps);
        hipLaunchKernelGGL(x_rstep<4>, dim3(XGRID), dim3(XBS), 0, st, X, r, eps);
      } else {
        hipLaunchKernelGGL(x_lstep<0>, dim3(XGRID), dim3(XBS), 0, st, X, r, eps);
        hipLaunchKernelGGL(x_rstep<0>, dim3(XGRID), dim3(XBS), 0, st, X, r, eps);
      }
    }
  };
  // at most 65536 rounds per launch, so that no launch runs for seconds
  auto wg_rounds = [&](int64_t maxr) {
    if (reg) hipLaunchKernelGGL(x_rounds_wg<4>, dim3(1), dim3(XWG), 0, st, X, r, eps,
                                std::min<int64_t>(maxr, 65536), small, ran);
    else hipLaunchKernelGGL(x_rounds_wg<0>, dim3(1), dim3(XWG), 0, st, X, r, eps,
                            std::min<int64_t>(maxr, 65536), small, ran);
  };
  auto cert_launch = [&]() {
    hipLaunchKernelGGL(x_cert_chunks, dim3(DET_T), dim3(XC_T), 0, st, X, part);
    hipLaunchKernelGGL(x_det_final, dim3(1), dim3(256), 0, st, part, sums);
  };
  auto cert = [&]() -> f2m_status {
    cert_launch();
    F2M_LAUNCH_CHECK(c);
    f2m_status s = post();
    if (s) return s;
    P = hv.P;
    D = hv.D;
    return F2M_OK;
  };
  for (;;) {
    if (phases > 0)
      hipLaunchKernelGGL(x_flow_sat_lanes, dim3(grid_for((int64_t)n * XF_LANES, bs.x)), bs, 0, st, X, eps);
    hipLaunchKernelGGL(x_zero_ring, dim3(1), dim3(64), 0, st, X.cnt);
    const dim3 ge(grid_for(n, XE_T)), be(XE_T);
    if (phases == 0) hipLaunchKernelGGL(x_count_lanes, gl, bs, 0, st, X);
    hipLaunchKernelGGL(x_excess, ge, be, 0, st, X, r);
    F2M_LAUNCH_CHECK(c);
    int chunk = 2;
    bool in_wg = false, first = true, have_cert = false;
    for (;;) {
      const int64_t left = o.exact_max_rounds - done;
      if (left <= 0) { stuck = true; break; }
      if (fuse && first && left > fuse_k) {
        // the usual course of a phase in one go; the in-workgroup rounds
        // return at once if the lists are still long, and the certificate
        // is used only if the phase has ended
        first = false;
        grid_rounds(fuse_k);
        wg_rounds(left - fuse_k);
        cert_launch();
        F2M_LAUNCH_CHECK(c);
        f2m_status s = post();
        if (s) return s;
        r += hv.ran;
        done = (int64_t)hv.rounds;
        const int kl = (int)((r - 1) % RING);
        if (hv.cnt[2 * kl] + hv.cnt[2 * kl + 1] == 0) {
          P = hv.P;
          D = hv.D;
          have_cert = true;
          break;
        }
        fuse = false;  // this graph's phases outlast the fused pass: a wait per launch group from here
        in_wg = hv.cnt[2 * (int)(r % RING)] < small;
        continue;
      }
      first = false;
      if (in_wg) {
        wg_rounds(left);
        F2M_LAUNCH_CHECK(c);
        f2m_status s = post();
        if (s) return s;
        r += hv.ran;
        done = (int64_t)hv.rounds;
        const int kl = (int)((r - 1) % RING);
        if (hv.cnt[2 * kl] + hv.cnt[2 * kl + 1] == 0) break;
        in_wg = hv.cnt[2 * (int)(r % RING)] < small;  // else the lists grew: grid steps
        chunk = 2;
        continue;
      }
      grid_rounds(std::min<int64_t>(chunk, left));
      F2M_LAUNCH_CHECK(c);
      f2m_status s = post();
      if (s) return s;
      done = (int64_t)hv.rounds;
      const int kl = (int)((r - 1) % RING);
      if (hv.cnt[2 * kl] + hv.cnt[2 * kl + 1] == 0) break;
      chunk = std::min(chunk * 2, 64);
      in_wg = small > 0 && hv.cnt[2 * (int)(r % RING)] < small;
    }
    ++phases;
    if (stuck) break;
    if (!have_cert) {
      f2m_status s = cert();
      if (s) return s;
    }
    const double tol = std::max(integral ? 0.4 : 0.0, o.exact_tol * std::fabs(P));
    if (P - D <= tol) { certified = true; break; }
    eps = eps / (double)o.exact_alpha;
    if (phases >= o.exact_max_phases) break;
  }
  if (stuck) {
    f2m_status s = cert();
    if (s) return s;
  }
  hipLaunchKernelGGL(x_out, dim3(grid_for(std::max<int64_t>(m, n), XBS)), bs, 0, st, X, m, dmin, xc, fc, F);
  int32_t* par = X.actL0;  // free now
  hipLaunchKernelGGL(x_uf_init, gn, bs, 0, st, n, par);
  hipLaunchKernelGGL(x_half_union, gl, bs, 0, st, X, par);
  hipLaunchKernelGGL(x_zero_i32, dim3(1), dim3(64), 0, st, flag + 4, 1);
  hipLaunchKernelGGL(x_half_root, dim3(grid_for((int64_t)n * XF_LANES, 1024)), dim3(1024), 0, st, X, par, root,
                     flag + 4);
  F2M_LAUNCH_CHECK(c);
  {
    f2m_status s = post();
    if (s) return s;
  }
  z.n_combs = hv.flag[4];
  z.exact_phases = phases;
  z.exact_rounds = done;
  z.exact_gap = P - D;
  z.exact_certified = certified ? 1 : 0;
  rc = certified ? 0 : 1;
  return F2M_OK;
}

}  // namespace f2m
