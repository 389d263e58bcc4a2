// f2m_internal.h -- shared declarations of the f2m_hip library (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

#include "f2m.h"

namespace f2m {

// Grow-only device buffer.  Reallocation frees first, so callers must not hold
// raw pointers across a reserve() of the same buffer.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t reserve(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&p, bytes < 256 ? 256 : bytes);
    if (e == hipSuccess) cap = bytes < 256 ? 256 : bytes;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// Solve workspace for one graph (single instance or a packed batch).
struct SolveWs {
  DevBuf nbr, D, orig;          // working slots (permuted in place), per slot
  DevBuf F0, F1, g;             // potentials (double-buffered), gradient
  DevBuf act, clen, deg, A, T;  // per node
  DevBuf root, irr, frozen;     // per node
  DevBuf xw, xc;                // x per working slot / per canonical slot
  DevBuf red;                   // reduction scratch (doubles / max bits)
  DevBuf axb;                   // least-squares scratch
  DevBuf sortk, sortv, sortk2, sortv2, sorttmp;  // comb node sort
  DevBuf ef, efoff;             // compacted active lists (16-byte slots) + offsets
  DevBuf psl, povo, ptmp;       // padded active lists (v*8 + k, overflow after n*8) + offsets
  DevBuf xmir, xf, xp, xe, xact, xcnt;  // exact recovery: mirror slots, flows, prices, excesses, lists
  DevBuf thdr, tnode, tslot, tovf;      // spatial tiles for multi-sweep launches (f2m_solve.hip)
  int cur = 0;                  // which F buffer holds the current potentials
  int64_t n = 0, m = 0;
};

}  // namespace f2m

struct f2m_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  // graph storage (f2m_knn_build / f2m_graph_from_csr)
  f2m::DevBuf g_rowptr, g_nbr, g_cost;
  // kNN scratch
  f2m::DevBuf k_x, k_y, k_cell, k_cstart, k_order, k_sx, k_sy, k_knn, k_kcost, k_rev, k_cur,
      k_deg, k_red, k_scan, k_geom;
  // symmetrisation scratch: reverse-only flags per directed edge, unsorted slots (owner, nbr, cost)
  f2m::DevBuf k_flag, k_town, k_tnbr, k_tcost;
  // single-pass scan: per-tile status words, tagged with the call's epoch
  f2m::DevBuf k_lbst;
  uint32_t k_lb_epoch = 0;
  // sequence number of the plain solve's max read-backs (pinned word 2050)
  unsigned long long read_seq = 0;
  // staging of host arrays
  f2m::DevBuf stage[6];
  // spatial ordering scratch (f2m_order.hip)
  f2m::DevBuf o_buf[5];
  f2m::SolveWs ws;
  // batched storage
  f2m::DevBuf b_node_off, b_rowptr, b_nbr, b_cost, b_order, b_stats, b_cur, b_knn, b_kc, b_ef, b_efoff, b_finit;
  f2m::DevBuf b_fout;                              // potentials of the last batch, gathered per instance
  f2m::SolveWs bws;
  void* b_plan = nullptr;                          // captured launch plan (f2m_batch.hip)
  struct TuneEntry {
    int64_t max_n;  // instances with n <= max_n ...
    int32_t lpn;    // ... use this many lanes per node
    int32_t bs;     // ... and workgroups of this many threads
    int32_t lds;    // ... and this shared-memory placement (-1: default, else as F2M_BATCH_LDS)
    int32_t pad;
  };
  std::vector<TuneEntry> b_tune;
  f2m::DevBuf b_clk;                               // per-instance ticks while autotuning
  bool b_clk_on = false;

  int64_t b_n_total = 0, b_nslots = 0;
  int32_t b_n_inst = 0;
  // pinned host scalars for small read-backs
  void* pinned = nullptr;
  // last solved graph (for f2m_result_copy)
  int64_t last_n = 0, last_m = 0;
  double last_dmin = 0;    // d_min of the last single-graph solve
  int32_t last_k = 0;      // k of its graph (0: caller CSR)
  // instrumentation
  int32_t prof_every = 0;
  std::vector<hipEvent_t> prof_ev;  // pool of start/end pairs
  f2m_profile prof{};
};

namespace f2m {

f2m_status fail(f2m_ctx* c, f2m_status s, const std::string& msg);
// Same, assigning into the context's message buffer (reserved at create), so
// a message shorter than that buffer allocates nothing.
f2m_status fail(f2m_ctx* c, f2m_status s, const char* msg);
// Looks up this device's address of the allocation-free entry point's scratch
// array (f2m_noalloc.hip); call with `device` current.  Allocates nothing.
hipError_t noalloc_init(int device);
f2m_status hip_fail(f2m_ctx* c, hipError_t e, const char* what, const char* file, int line);

// Device pointer for p: p itself if device-accessible, else a staged copy in
// ctx->stage[slot].  Returns nullptr and sets the error on failure.
const void* stage_in(f2m_ctx* c, const void* p, size_t bytes, int slot);
bool is_device_ptr(const void* p);
// Copy bytes from device src to dst (device or host), stream-ordered + synced.
hipError_t copy_out(f2m_ctx* c, void* dst, const void* src, size_t bytes);

// Exclusive scan of int32/int64 counts (n entries) into out (n+1 entries,
// out[n] = total); total_host may be null.  tmp is scratch.
hipError_t scan_i32_to_i32(f2m_ctx* c, const int32_t* in, int32_t* out, int64_t n, int32_t* total_host);
hipError_t scan_i32_to_i64(f2m_ctx* c, const int32_t* in, int64_t* out, int64_t n, int64_t* total_host);
// Exclusive scan of max(in[i] - sub, 0) into out (n+1 entries, out[n] =
// total) in one launch (tiles chained by decoupled look-back); the totals fit
// an int32.  With set_ptr, the launch also stores set_val there (stream order
// for a flag the next kernel reads).
hipError_t scan_excess_i32(f2m_ctx* c, const int32_t* in, int32_t sub, int32_t* out, int64_t n,
                           int32_t* set_ptr = nullptr, int32_t set_val = 0);

// Destroys the captured batched-launch plan of a context (f2m_batch.hip).
void release_batch_plan(f2m_ctx* c);

// Exact primal recovery (f2m_exact.hip).  From potentials F [dev,
// normalised] on the canonical CSR, writes x, its
// X-flags and the certifying F; root = components of the x = 1/2 edges;
// fills the exact_* stats and n_combs.  rc: 0 certified, 1 not certified,
// -1 asymmetric graph, -2 a node of degree below 2 (outputs untouched).
f2m_status exact_stage(f2m_ctx* c, SolveWs& w, int n, int64_t m, const int32_t* rp, const int32_t* nbr,
                       const double* cost, double dmin, const f2m_opts& o, double* F, double* xc, uint8_t* fc,
                       int32_t* root, f2m_stats& z, int& rc);

inline int grid_for(int64_t threads, int block) { return (int)((threads + block - 1) / block); }

// Integer tuning switch from the environment (dflt when unset).
inline int env_int(const char* name, int dflt) {
  const char* s = std::getenv(name);
  return s ? std::atoi(s) : dflt;
}

}  // namespace f2m

#define F2M_HIP(c, expr)                                                   \
  do {                                                                     \
    hipError_t e__ = (expr);                                               \
    if (e__ != hipSuccess) return f2m::hip_fail((c), e__, #expr, __FILE__, __LINE__); \
  } while (0)

#define F2M_LAUNCH_CHECK(c) F2M_HIP((c), hipGetLastError())

// ---------------------------------------------------------------------------
// Device helpers shared by the solve, batch and comb kernels.
// ---------------------------------------------------------------------------
namespace f2m {

// Reduced cost of slot (v -> w): d = (D + 0.5*F[lo]) + 0.5*F[hi], lo = min id
// (graph_cuda_grad.hpp:9; Edges_v1 < Edges_v2, graph_cuda_init.hpp:25-27).
// Compiled with -ffp-contract=off; 0.5*F is exact, so the order of the two
// roundings is the only thing that matters and it is fixed here.
__device__ __forceinline__ double dval(double D, double Fv, double Fw, int v, int w) {
  double a = v < w ? Fv : Fw;
  double b = v < w ? Fw : Fv;
  double t = D + 0.5 * a;
  return t + 0.5 * b;
}

// Reduced cost of slot (v -> w) from the halves hv = 0.5*F[v], hw = 0.5*F[w]:
// the operations of dval (0.5*F is exact), so the same bits.
__device__ __forceinline__ double dval_h(double D, double hv, double hw, int v, int w) {
  const bool lo = v < w;
  return (D + (lo ? hv : hw)) + (lo ? hw : hv);
}

// Three smallest values, sentinels 1e6 (graph_cuda_grad.hpp:68-79).  The
// reference inserts with a branch chain (v < a: shift both; else v < b; else
// v < c); the min/max network below keeps the same three values: min and
// max return one of their operands, and ties between equal values only
// differ by the sign of a zero.  A reduced cost is never -0.0 (D >= +0, and a
// round-to-nearest sum that is zero with a +0 or nonzero operand is +0) nor
// NaN, so the bits equal the branch chain's.  5 v_min/v_max_f64 per value
// instead of 3 compares and up to 6 selects of double pairs.
struct Top3 {
  double a, b, c;
  __device__ __forceinline__ Top3() : a(1e6), b(1e6), c(1e6) {}
  __device__ __forceinline__ void push(double v) {
    const double ha = __builtin_fmax(a, v);
    const double hb = __builtin_fmax(b, v);
    a = __builtin_fmin(a, v);
    c = __builtin_fmin(c, hb);
    b = __builtin_fmin(b, ha);
  }
  __device__ __forceinline__ double grad() const { return -(b + c) / 2; }
};

// A fresh Top3 after pushing x0..x3: the three smallest of x0..x3 and the
// sentinels.  A selection network gives the three smallest of four values
// (pairs sorted; the least of the lows, then the larger low and the smaller
// high in order): 9 min/max, and 3 against the sentinel, instead of 20 for
// four pushes.  min and max return an operand, so the bits are the pushes'.
__device__ __forceinline__ Top3 top3_of4(double x0, double x1, double x2, double x3) {
  const double l1 = __builtin_fmin(x0, x1), h1 = __builtin_fmax(x0, x1);
  const double l2 = __builtin_fmin(x2, x3), h2 = __builtin_fmax(x2, x3);
  const double m1 = __builtin_fmax(l1, l2), m2 = __builtin_fmin(h1, h2);
  Top3 t;
  t.a = __builtin_fmin(__builtin_fmin(l1, l2), t.a);
  t.b = __builtin_fmin(__builtin_fmin(m1, m2), t.b);
  t.c = __builtin_fmin(__builtin_fmax(m1, m2), t.c);
  return t;
}

// Hoare partition of node v's slots [base, base+size) around thr, permuting
// (nbr, D, orig) in place (qpart_ptr_B, graph_cuda_partition.hpp:3-25).  This
// is the step-by-step form; lists of up to 64 slots take qpart_mask below,
// which makes the same swaps.
__device__ inline int qpart_B_loop(int v, int64_t base, int size, double thr, int* __restrict__ nbr,
                                   double* __restrict__ D, int* __restrict__ orig,
                                   const double* __restrict__ F) {
  const double Fv = F[v];
  int i = 0, j = size - 1;
  while (i < j) {
    while (i <= size - 1) {
      int w = nbr[base + i];
      if (!(dval(D[base + i], Fv, F[w], v, w) <= thr)) break;
      ++i;
    }
    while (j >= 0) {
      int w = nbr[base + j];
      if (!(dval(D[base + j], Fv, F[w], v, w) > thr)) break;
      --j;
    }
    if (i < j) {
      int64_t si = base + i, sj = base + j;
      int tn = nbr[si]; nbr[si] = nbr[sj]; nbr[sj] = tn;
      double tD = D[si]; D[si] = D[sj]; D[sj] = tD;
      int to = orig[si]; orig[si] = orig[sj]; orig[sj] = to;
    }
  }
  return j + 1;
}

// The predicate of a partition, for a list of up to 64 slots in its current
// order: bit k of the result = (d of slot k <= thr); *above = the least d
// above thr (+inf if none).  The slots are read in groups of 8 whose loads
// do not wait on each other.
__device__ inline unsigned long long part_mask(int v, int64_t base, int size, double thr,
                                               const int* __restrict__ nbr, const double* __restrict__ D,
                                               const double* __restrict__ F, double* above) {
  const double Fv = F[v];
  unsigned long long m = 0;
  double ab = 1.0 / 0.0;
  for (int k0 = 0; k0 < size; k0 += 8) {
    int w[8];
    double dd[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = k0 + q < size ? k0 + q : size - 1;  // past the list: repeats the last slot, unused
      w[q] = nbr[base + k];
      dd[q] = D[base + k];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (k0 + q < size) {
        const double d = dval(dd[q], Fv, F[w[q]], v, w[q]);
        if (d <= thr) m |= 1ull << (k0 + q);
        else ab = d < ab ? d : ab;
      }
    }
  }
  *above = ab;
  return m;
}

// The Hoare partition of qpart_B_loop driven by the predicate bits m of the
// slots in their current order.  Its scans (i: the next slot from i not at or
// below the threshold, j: the previous slot from j at or below it, swapped
// while i < j) pair the k-th such slot from the left with the k-th at-or-below
// slot from the right, for as long as the first lies left of the second, and
// move past both.  With p = popcount(m), those pairs are exactly the slots
// above the threshold in [0, p) in ascending order against the slots at or
// below it in [p, size) in descending order, every slot swapped at most once,
// and the step-by-step form returns j + 1 = p.  So the swaps are computed from
// the bits first and carried out in groups of four whose loads are issued
// together: the same swaps, with one memory round trip per group instead of
// one per swap.  Not inlined: inside the batched solve's large function the
// group's registers were spilled, with reloads placed ahead of an exec
// restore (the build's ISA scan rejects that).
__device__ __attribute__((noinline)) int qpart_mask(int64_t base, int size, unsigned long long m, int* __restrict__ nbr,
                                 double* __restrict__ D, int* __restrict__ orig) {
  const unsigned long long full = size >= 64 ? ~0ull : ((1ull << size) - 1ull);
  m &= full;
  const int p = __popcll(m);
  const unsigned long long left = p >= 64 ? ~0ull : ((1ull << p) - 1ull);
  unsigned long long A = ~m & left;  // above the threshold, left of the pivot (ascending)
  unsigned long long B = m & ~left;  // at or below it, right of the pivot (descending); |A| = |B|
  constexpr int G = 4;
  while (A) {
    int si[G], sj[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
      si[q] = -1;
      sj[q] = -1;
      if (A) {
        si[q] = __builtin_ctzll(A);
        sj[q] = 63 - __builtin_clzll(B);
        A &= A - 1;
        B &= ~(1ull << sj[q]);
      }
    }
    int tn[G], to[G], un[G], uo[G];
    double tD[G], uD[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
      if (si[q] >= 0) {
        const int64_t a = base + si[q], b = base + sj[q];
        tn[q] = nbr[a];
        to[q] = orig[a];
        tD[q] = D[a];
        un[q] = nbr[b];
        uo[q] = orig[b];
        uD[q] = D[b];
      }
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
      if (si[q] >= 0) {
        const int64_t a = base + si[q], b = base + sj[q];
        nbr[a] = un[q];
        orig[a] = uo[q];
        D[a] = uD[q];
        nbr[b] = tn[q];
        orig[b] = to[q];
        D[b] = tD[q];
      }
    }
  }
  return p;
}

// The swaps of qpart_mask shared by the NL lanes of a node (lane q of
// 0..NL-1; every lane holds the same bits m): the same pairs, lane q
// carrying out pairs q, q + NL, q + 2 NL, ... in groups of G whose loads are
// issued together.  A slot is in at most one pair, so the lanes' swaps are
// independent.  Returns p = popcount(m & the list) on every lane.
template <int NL, int G>
__device__ __forceinline__ int qpart_mask_lanes(int64_t base, int size, unsigned long long m, int q,
                                                int* __restrict__ nbr, double* __restrict__ D,
                                                int* __restrict__ orig) {
  const unsigned long long full = size >= 64 ? ~0ull : ((1ull << size) - 1ull);
  m &= full;
  const int p = __popcll(m);
  const unsigned long long left = p >= 64 ? ~0ull : ((1ull << p) - 1ull);
  unsigned long long A = ~m & left;  // above the threshold, left of the pivot (ascending)
  unsigned long long B = m & ~left;  // at or below it, right of the pivot (descending); |A| = |B|
  auto drop = [&]() {                // the next pair, not this lane's
    A &= A - 1;
    B &= ~(1ull << (63 - __builtin_clzll(B)));
  };
  for (int i = 0; i < q && A; ++i) drop();
  while (A) {
    int si[G], sj[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      si[g] = -1;
      sj[g] = -1;
      if (A) {
        si[g] = __builtin_ctzll(A);
        sj[g] = 63 - __builtin_clzll(B);
        drop();
        for (int i = 1; i < NL && A; ++i) drop();
      }
    }
    int tn[G], to[G], un[G], uo[G];
    double tD[G], uD[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (si[g] >= 0) {
        const int64_t a = base + si[g], b = base + sj[g];
        tn[g] = nbr[a];
        to[g] = orig[a];
        tD[g] = D[a];
        un[g] = nbr[b];
        uo[g] = orig[b];
        uD[g] = D[b];
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (si[g] >= 0) {
        const int64_t a = base + si[g], b = base + sj[g];
        nbr[a] = un[g];
        orig[a] = uo[g];
        D[a] = uD[g];
        nbr[b] = tn[g];
        orig[b] = to[g];
        D[b] = tD[g];
      }
    }
  }
  return p;
}

__device__ inline int qpart_B(int v, int64_t base, int size, double thr, int* __restrict__ nbr,
                              double* __restrict__ D, int* __restrict__ orig,
                              const double* __restrict__ F) {
  if (size > 64) return qpart_B_loop(v, base, size, thr, nbr, D, orig, F);
  double above;
  const unsigned long long m = part_mask(v, base, size, thr, nbr, D, F, &above);
  return qpart_mask(base, size, m, nbr, D, orig);
}

// The threshold after qpart_ptr_EF's `do thr += 1 while (!(above <= thr)
// && ++guard < 2^40)` (graph_cuda_partition.hpp:27-54).  From an integral
// thr (|thr| < 2^52) every step is exact and thr + t is an integer, so
// above <= thr + t holds exactly from t = ceil(above) - thr on: the loop ends
// at thr + 1, at ceil(above), or at the guard's thr + 2^40.  Any other thr
// (or a NaN above) runs the loop itself.  part_v2's thresholds start at 1.0.
// Not inlined: a rare path, kept out of the callers' register allocation.
__device__ __attribute__((noinline)) double raise_thr(double thr, double above) {
  if (thr == floor(thr) && fabs(thr) < 4503599627370496.0 && above == above) {
    const double t1 = thr + 1.0;
    if (above <= t1) return t1;
    const double cmax = thr + 1099511627776.0;
    const double c = ceil(above);
    return c < cmax ? c : cmax;
  }
  long long guard = 0;
  do { thr += 1.0; } while (!(above <= thr) && ++guard < (1ll << 40));
  return thr;
}

// qpart_ptr_EF (graph_cuda_partition.hpp:27-54): raise thr by 1 until >= 3
// entries lie at or below it.  Re-partitions that admit no new entry swap
// nothing, so only thresholds admitting a new entry are partitioned.  After a
// partition at thr the slots from the pivot on are exactly those above thr,
// so their least d is the least d above thr of the whole list.
__device__ inline int qpart_EF(int v, int64_t base, int size, double thr, int* __restrict__ nbr,
                               double* __restrict__ D, int* __restrict__ orig,
                               const double* __restrict__ F) {
  if (size > 64) {
    int p = qpart_B_loop(v, base, size, thr, nbr, D, orig, F);
    if (size < 3) return p;
    const double Fv = F[v];
    while (p < 3) {
      double m = 1.0 / 0.0;
      for (int k = p; k < size; ++k) {
        int w = nbr[base + k];
        double d = dval(D[base + k], Fv, F[w], v, w);
        m = d < m ? d : m;
      }
      thr = raise_thr(thr, m);
      p = qpart_B_loop(v, base, size, thr, nbr, D, orig, F);
    }
    return p;
  }
  double above;
  unsigned long long mask = part_mask(v, base, size, thr, nbr, D, F, &above);
  int p = qpart_mask(base, size, mask, nbr, D, orig);
  if (size < 3) return p;
  while (p < 3) {
    thr = raise_thr(thr, above);
    mask = part_mask(v, base, size, thr, nbr, D, F, &above);  // the slots in their new order
    p = qpart_mask(base, size, mask, nbr, D, orig);
  }
  return p;
}

// |x| of a double as an order-preserving 64-bit key (for atomicMax).
__device__ __forceinline__ unsigned long long abs_bits(double x) {
  return (unsigned long long)__double_as_longlong(fabs(x));
}

}  // namespace f2m
