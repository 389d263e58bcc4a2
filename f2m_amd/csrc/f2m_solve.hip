// f2m_solve.hip -- dual descent, partitions, max reductions, tight components,
// primal recovery and objective for one graph (replaces graph::RUN and the
// graph_c GPU methods, graph_cuda.cu:47-430 and graph_cuda_*.hpp).
//
// Layout.  The canonical CSR (rowptr, nbr, cost) is never modified.  The
// solver works on a copy of the slots, (nbr, D = cost/d_min, orig slot),
// which the partitions permute in place within each node's segment, exactly
// as the reference permutes its per-node pointer lists
// (graph_cuda_partition.hpp).  A node's active list (the reference's EF
// array, graph_cuda.cu:229-252) is then simply the first act[v] slots of its
// segment: no compaction, no pointer chasing, and a sweep reads one
// contiguous run of (nbr, D) per node.
//
// Sweep (graph_cuda_grad.hpp:1-84 fused): one kernel computes, per node, the
// three smallest reduced costs of its active slots from the current F, the
// gradient g = -(t2+t3)/2 and F' = F + g into the other F buffer, and the
// block max |g| into a device slot.  No E-sized d array is stored: d is
// recomputed from F on the fly with the reference's exact expression.
// LPN lanes cooperate on a node and merge their top-3 lists with xor
// shuffles; the merge is a selection, so every LPN gives identical bits.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <hip/hip_ext.h>

#include "f2m_device.h"

namespace f2m {
namespace {

constexpr int BS = 256;
constexpr int RED_BLOCKS = 512;
// Algorithmic bytes of one sweep (instrumentation): per active slot the
// neighbour id and the cost (4 + 4 B), per node the list offset, F in and
// F out (4 + 8 + 8 B): B_sweep = 8 * (active slots) + 20 * n.
constexpr double kAlgSlotBytes = 8.0;
constexpr double kAlgNodeBytes = 20.0;
constexpr unsigned long long INF_BITS = 0x7FF0000000000000ull;

__global__ void s_zero_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void s_zero_i32(int32_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void s_zero_u8(uint8_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void s_set_i32(int32_t* p, int32_t v) {
  if (threadIdx.x == 0) *p = v;
}
// One wave: the max of the 32 strided max slots of a region (|g| bits, >= 0)
// and, with fit, the tiles' fit counter, stored to pinned host memory
// (host[0], host[1]), then seq to host[2] with a system-scope release, so a
// host that reads seq there reads the values too; the slots are cleared for
// the region's next use, the fit counter reset to INT_MAX for the next tile
// build's atomicMin.
__global__ void s_read_max(unsigned long long* __restrict__ region, int32_t* __restrict__ fit,
                           unsigned long long* __restrict__ host, unsigned long long seq) {
  const int k = threadIdx.x;
  unsigned long long b = k < 32 ? region[16 * k] : 0ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(b, off, 64);
    b = o > b ? o : b;
  }
  if (k < 32) region[16 * k] = 0ull;
  if (k == 0) {  // system scope: written through to host memory
    const int32_t f = fit ? *fit : 0;
    if (fit) *fit = INT_MAX;
    __hip_atomic_store(&host[0], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&host[1], (unsigned long long)(uint32_t)f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&host[2], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
__global__ void s_zero_f64(double* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
// Block-wide max of a non-negative value into *slot (as bits, atomicMax).
__device__ __forceinline__ void block_max_to(double v, unsigned long long* slot) {
  __shared__ double wm[BS / 64];
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wm[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wm[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = fmax(m, wm[k]);
    atomicMax(slot, (unsigned long long)__double_as_longlong(m));
  }
}

// smallest positive cost, as bits (positive doubles order like uint64); one
// atomic per workgroup (one per wave contended on the single word: 51 us at
// 2.3M slots)
__global__ void __launch_bounds__(BS) s_min_pos(const double* __restrict__ c, int64_t m, unsigned long long* out) {
  __shared__ unsigned long long wb[BS / 64];
  unsigned long long best = INF_BITS;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    double v = c[i];
    if (v > 0) {
      unsigned long long b = (unsigned long long)__double_as_longlong(v);
      best = b < best ? b : best;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = (unsigned long long)__shfl_xor((long long)best, off, 64);
    best = o < best ? o : best;
  }
  if ((threadIdx.x & 63) == 0) wb[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < BS / 64; ++k) best = wb[k] < best ? wb[k] : best;
    atomicMin(out, best);
  }
}

// working slots: D = cost / d_min (graph_tour.hpp:94; IEEE division)
__global__ void s_init_slots(int64_t m, const int32_t* __restrict__ nbr, const double* __restrict__ cost,
                             const unsigned long long* __restrict__ dmin_bits, int32_t* __restrict__ wnbr,
                             double* __restrict__ wD, int32_t* __restrict__ worig) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  const double dmin = __longlong_as_double((long long)*dmin_bits);
  wnbr[s] = nbr[s];
  wD[s] = cost[s] / dmin;
  worig[s] = (int32_t)s;
}
__global__ void s_init_nodes(int n, const int32_t* __restrict__ rp, int32_t* __restrict__ deg,
                             double* __restrict__ F, uint8_t* __restrict__ frozen) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  deg[v] = rp[v + 1] - rp[v];
  F[v] = 0.0;
  frozen[v] = 0;
}

// setvalue2V (graph_cuda.cu:354-361): act = min(deg, k)
__global__ void s_act_min(int n, const int32_t* __restrict__ deg, int32_t* __restrict__ act, int k) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) act[v] = deg[v] > k ? k : deg[v];
}

// Compact the first act[v] working slots of every node into ef[efoff[v]..].
__global__ void s_ef_fill(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ act,
                          const int32_t* __restrict__ efoff, const int32_t* __restrict__ nbr,
                          const double* __restrict__ D, ESlot* __restrict__ ef) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int b = rp[v], o = efoff[v], L = act[v];
  for (int k = 0; k < L; ++k) ef[o + k] = ESlot{D[b + k], nbr[b + k], 0};
}

// Fused sweep / gradient over the compacted active lists.  WRITE_F=false
// computes g and max|g| only (cuda_get_Nodes_grad_R + max_node_grad);
// WRITE_F=true also applies F' = F + g (cuda_update_Nodes_F[_BLOCK]); the
// reduced costs of the next sweep are recomputed from F' (cuda_recompute_Edges).
// The block max goes to one of 32 slots 128 bytes apart (blockIdx % 32) so the
// atomics of different blocks do not queue on one address.
template <int LPN, bool FROZEN, bool WRITE_F>
__global__ void __launch_bounds__(BS) s_sweep(int n, const int32_t* __restrict__ efoff,
                                              const ESlot* __restrict__ ef, const double* __restrict__ Fin,
                                              double* __restrict__ Fout, double* __restrict__ gout,
                                              const uint8_t* __restrict__ frozen,
                                              unsigned long long* __restrict__ maxslots) {
  const int gt = blockIdx.x * BS + threadIdx.x;
  const int v = gt / LPN, lane = gt % LPN;
  Top3 t;
  double Fv = 0.0;
  const bool ok = v < n;
  if (ok) {
    const int b = efoff[v], e = efoff[v + 1];
    Fv = Fin[v];
    for (int k = b + lane; k < e; k += LPN) {
      const ESlot sl = ef[k];
      t.push(dval(sl.D, Fv, Fin[sl.nbr], v, sl.nbr));
    }
  }
#pragma unroll
  for (int off = 1; off < LPN; off <<= 1) {
    const double a2 = __shfl_xor(t.a, off, 64);
    const double b2 = __shfl_xor(t.b, off, 64);
    const double c2 = __shfl_xor(t.c, off, 64);
    t.push(a2);
    t.push(b2);
    t.push(c2);
  }
  double gabs = 0.0;
  if (ok && lane == 0) {
    const double g = t.grad();
    if (gout) gout[v] = g;
    if (WRITE_F) Fout[v] = (FROZEN && frozen[v]) ? Fv : Fv + g;
    gabs = fabs(g);
  }
  if (maxslots) block_max_to(gabs, maxslots + (blockIdx.x & 31) * 16);
}

// profiling: acc += (active slots of the current lists) x sweeps
__global__ void s_acc_slots(const int32_t* __restrict__ total, unsigned long long* acc, unsigned long long sweeps) {
  if (threadIdx.x == 0) *acc += (unsigned long long)(*total) * sweeps;
}

// nodes whose K-th nearest cost r_v is below -d_min F_v: an edge left out of
// the kNN graph costs at least max(r_u, r_w), so with no such node every
// left-out edge has a nonnegative reduced cost (y = -d_min F / 2)
__global__ void s_knn_certify(int n, const int32_t* __restrict__ rp, const double* __restrict__ cost, int K,
                              double dmin, const double* __restrict__ F, unsigned long long* cnt) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  int bad = 0;
  if (v < n && rp[v + 1] - rp[v] >= K) bad = cost[rp[v] + K - 1] < -(dmin * F[v]) ? 1 : 0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) bad += __shfl_xor(bad, off, 64);
  if ((threadIdx.x & 63) == 0 && bad) atomicAdd(cnt, (unsigned long long)bad);
}

// ---- padded layout ----------------------------------------------------------
// The compacted sweep above reads three dependent levels: list offset ->
// slot -> neighbour potential.  The padded layout keeps each node's first PP
// active slots at a fixed stride (v*PP + k; longer lists continue in an
// overflow region after n*PP), so the slot loads need no offset and are
// issued unconditionally together with act[v] and F[v]: two levels (slots,
// then neighbour potentials) for every node with at most PP active slots.
// (Pushing potentials into per-slot copies instead, to save the second
// level, was measured slower: the scattered copy stores cost more than the
// gather they replace.)  Stride and lanes per node are selectable
// (F2M_PAD="stride,lanes[,g]"; g: guarded slot loads); default 4,2, the
// fastest measured on the n = 100k Gaussian-mixture graph (5.6 us per sweep
// against 6.2 us for 8,4, 5.8 us for 4,4 and 2,2, 6.5 us for 2,1).
constexpr int PP_MAX = 8;  // largest padded stride (workspace sizing)

// Four lanes per node, lane q copying slots q, q + 4, ...: the padded slots
// of consecutive nodes are consecutive, so the stores of a wave coalesce.
// ovo == nullptr: only the first pp slots of each list (what the tile build
// reads from this copy); set_ptr: *set_ptr = set_val (the tiles' fit counter
// before a build).
constexpr int PF_LANES = 4;
__global__ void s_pad_fill(int n, int pp, const int32_t* __restrict__ rp, const int32_t* __restrict__ act,
                           const int32_t* __restrict__ ovo, const int32_t* __restrict__ nbr,
                           const double* __restrict__ D, ESlot* __restrict__ sl, int32_t* __restrict__ set_ptr,
                           int32_t set_val) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0 && set_ptr) *set_ptr = set_val;
  const int v = (int)(t / PF_LANES), q = (int)(t % PF_LANES);
  if (v >= n) return;
  const int b = rp[v], L = ovo ? act[v] : min(act[v], pp), oo = ovo ? ovo[v] : 0;
  for (int k = q; k < L; k += PF_LANES) {
    const int64_t idx = k < pp ? (int64_t)v * pp + k : (int64_t)n * pp + oo + (k - pp);
    sl[idx] = ESlot{D[b + k], nbr[b + k], 0};
  }
}

// PPT padded slots per node, LPN lanes per node (PPT / LPN slots each); with
// SPEC the padded slots are loaded before act[v] is known, else only those
// below act[v].  Each lane group handles NPG nodes, v = g + j * (groups in
// the grid), their loads issued together so their latencies overlap (measured
// slower for NPG 2 and 4 at n = 100k: 5.5-8.0 us against 4.5 us for one node,
// so only NPG = 1 is launched).  Every variant selects the same three values
// per node: equal bits.
template <int PPT, int LPN, bool SPEC, int NPG, bool FROZEN, bool WRITE_F>
__global__ void __launch_bounds__(BS) s_sweep_pad(int n, const int32_t* __restrict__ act,
                                                  const int32_t* __restrict__ ovo, const ESlot* __restrict__ sl,
                                                  const double* __restrict__ Fin, double* __restrict__ Fout,
                                                  const uint8_t* __restrict__ frozen,
                                                  unsigned long long* __restrict__ maxslots) {
  constexpr int R = PPT / LPN;
  const int gt = blockIdx.x * BS + threadIdx.x;
  const int g0 = gt / LPN, lane = gt % LPN;
  const int ng = gridDim.x * (BS / LPN);
  Top3 t[NPG];
  double Fv[NPG];
  int L[NPG], oo[NPG];
  ESlot sr[NPG][R];
#pragma unroll
  for (int j = 0; j < NPG; ++j) {
    const int v = g0 + j * ng;
    Fv[j] = 0.0;
    L[j] = 0;
    oo[j] = 0;
    if (v < n) {
      L[j] = act[v];
      oo[j] = ovo[v];
      Fv[j] = Fin[v];
      const int64_t p = (int64_t)v * PPT + lane;
#pragma unroll
      for (int q = 0; q < R; ++q)
        if (SPEC || lane + q * LPN < L[j]) sr[j][q] = sl[p + q * LPN];
    }
  }
#pragma unroll
  for (int j = 0; j < NPG; ++j) {
    const int v = g0 + j * ng;
#pragma unroll
    for (int q = 0; q < R; ++q)
      if (lane + q * LPN < L[j]) t[j].push(dval(sr[j][q].D, Fv[j], Fin[sr[j][q].nbr], v, sr[j][q].nbr));
    for (int k = lane + PPT; k < L[j]; k += LPN) {
      const ESlot sk = sl[(int64_t)n * PPT + oo[j] + (k - PPT)];
      t[j].push(dval(sk.D, Fv[j], Fin[sk.nbr], v, sk.nbr));
    }
  }
  double gabs = 0.0;
#pragma unroll
  for (int j = 0; j < NPG; ++j) {
#pragma unroll
    for (int off = 1; off < LPN; off <<= 1) {
      const double a2 = __shfl_xor(t[j].a, off, 64);
      const double b2 = __shfl_xor(t[j].b, off, 64);
      const double c2 = __shfl_xor(t[j].c, off, 64);
      t[j].push(a2);
      t[j].push(b2);
      t[j].push(c2);
    }
    const int v = g0 + j * ng;
    if (v < n && lane == 0) {
      const double g = t[j].grad();
      if (WRITE_F) Fout[v] = (FROZEN && frozen[v]) ? Fv[j] : Fv[j] + g;
      gabs = fmax(gabs, fabs(g));
    }
  }
  if (maxslots) block_max_to(gabs, maxslots + (blockIdx.x & 31) * 16);
}

// ---- spatial tiles: several Jacobi sweeps per launch -------------------------
// A sweep launch costs ~4.5 us at n = 100k however little it moves (two
// dependent HBM levels plus the wave ramp), so the schedule's sweeps are run S
// at a time.  Tile t is the id range [t*ts, t*ts + ts) (ts <= TS, chosen per
// graph, see Runner::setup): with Morton-numbered nodes a compact
// region.  Its local set is the tile plus a halo: nodes reachable from the
// tile over active slots in at most depth edges, numbered breadth first
// (local index = id - t*ts inside the tile).  A launch of S
// sweeps (S <= depth + 1) gathers the potentials of the local set's active
// neighbours from HBM once, computes sweep 1 for every local node of depth
// <= S-1, sweep 2 from shared memory for depth <= S-2, ..., sweep S for the
// tile, and writes only the tile's final potentials.  The halo is recomputed
// by every tile that needs it: the arithmetic per node and sweep is the one
// of s_sweep_pad, so the potentials are bit-identical to S single sweeps.
constexpr int TS_W = 512;    // most nodes per tile of the 768-thread shape (1536 local nodes)
constexpr int TS = 256;      // most nodes per tile (the default is chosen per graph, <= TS).  On the
                             // n = 100k Gaussian-mixture solve 128-node tiles for the block_GLOB lists
                             // keep deeper halos within TL (5 sweeps per launch instead of 3) but the
                             // doubled halo work made the solve slower (42.7 against 38.7 ms)
constexpr int TPP_MAX = 6;   // padded slots per local node: 6 while every list holds >= 5, else 4
constexpr int THDR = 16;     // header ints per tile
constexpr int TMAXS = 12;    // most sweeps per launch
constexpr int THDR_OV = 14, THDR_FIT = 15;  // header ints: overflow slots, fit (levels use 0..TMAXS-1)
static_assert(TMAXS <= THDR_OV, "level counts stay below the overflow and fit ints");
// sweeps per launch unless F2M_FUSE says otherwise.  The n = 100k Gaussian-mixture solve at
// widths 5 / 6 / 7 / 8: 28.8 / 28.4 / 27.95 / 28.0 ms after the cheaper list rebuilds; with
// 384-thread tiles grouped per XCD widths 6 / 7 / 8: 25.15 / 24.86 / 24.62 ms; with one
// 768-thread tile per CU widths 7 / 8 / 9 / 10 / 12: 22.8 / 22.55 / 22.42 / 22.33 / 22.58 ms.
constexpr int DEFAULT_FUSE = 10;
// Tile workgroup shapes: W threads, local-node capacity TL = 2W (tile +
// halo), a halo hash of THSZ >= 2*TL entries, TOV overflow slots.  W = 384 by
// default.  768 (F2M_TILE_W_GLOB) lets the block_GLOB lists, whose halos grow
// faster, fit five levels instead of two at n = 100k (6 sweeps per launch instead
// of 3), but the larger recomputed halos made the solve slower (39.1 against
// 37.2 ms): a launch's time follows its node updates more than its count.
template <int W>
struct TCfg {
  static constexpr int TW = W;
  static constexpr int TL = 2 * W;
  static constexpr int THB = W <= 384 ? 11 : 13;  // log2 of the halo hash size (at most 3/16 full at 768)
  static constexpr int THSZ = 1 << THB;
  static constexpr int TOV = W <= 384 ? 1024 : 2048;
  static_assert((TOV & (TOV - 1)) == 0 && TOV >= TW, "TOV is a power of two, one speculative record per thread");
  static_assert(TL < 32768 && THSZ >= 2 * TL, "local indices fit a short; hash has room");
};
constexpr int TW_EF = 384, TW_GLOB = 768;
// XCDs of an MI355X: the dispatcher hands workgroups to them round robin.
// Consecutive tiles on one XCD share halo potentials through its L2
// (tile_of_block): the n = 100k Gaussian-mixture solve 25.5 -> 24.8 ms.
constexpr int kXcds = 8;
constexpr int TL_MAX = TCfg<TW_GLOB>::TL, TOV_MAX = TCfg<TW_GLOB>::TOV;

// local copy of an active slot: loc = local index of nbr (-1: not local)
struct TSlot {
  double D;
  int32_t nbr;
  int32_t loc;
};

template <int W>
__device__ __forceinline__ unsigned t_hash(int w) {
  return ((unsigned)w * 2654435761u) >> (32 - TCfg<W>::THB);
}
// local index clamped into the buffers (-1 only occurs where it is not read)
template <int W>
__device__ __forceinline__ unsigned t_loc(int loc) {
  return min((unsigned)loc, (unsigned)(TCfg<W>::TL - 1));
}

template <int W>
__device__ __forceinline__ void tile_max_to(double v, unsigned long long* slot) {
  constexpr int TW = W;
  __shared__ double wm[TW / 64];
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wm[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wm[0];
    for (int k = 1; k < TW / 64; ++k) m = fmax(m, wm[k]);
    atomicMax(slot, (unsigned long long)__double_as_longlong(m));
  }
}

// One workgroup per tile: breadth first from the tile over active slots, one
// level per pass, up to `depth` levels; fit = the deepest level kept (levels
// are kept while the local set fits TL and the hash has room).
// hdr[t*THDR + d] = local nodes of level <= d (d <= fit); hdr[t*THDR + THDR_OV] =
// overflow slots; hdr[t*THDR + THDR_FIT] = fit (-1 when the overflow slots do not
// fit).  Records are stored by position.  For lists longer than 4 padded
// slots (tpp > 4, the block_GLOB lists, a third of which overflow their
// padded slots) the local nodes of each level are ordered by decreasing
// list length, so that the sweeps' waves meet lists of similar lengths and
// loop over overflow slots together (n = 100k Gaussian mixture: 13.25 ->
// 12.71 us per 5-sweep launch; for the block_EF lists, which rarely
// overflow, the same order made the 10-sweep launch slower, 8.46 -> 8.74
// us, so they keep position = local index).  The record at position j says
// which local node it is.  tnode[t*TL + j] = {id,
// active slots, overflow offset, local index}; slots k < tpp of the node at
// position j at tslot[(t*TL + j)*tpp + k], the others at
// tovf[t*TOV + offset + k - tpp].  Positions of a level are the same range as
// its local indices, so a sweep's prefix of levels is a prefix of positions.  *fit_min = min over tiles of fit: a launch
// may run up to fit_min + 1 sweeps.
// Slot k of node v comes from the padded copy psl[v*pp + k] when psl is set
// and k < pp (readable together with the list length), else from the node's
// row, whose active slots come first: D[rp[v] + k], nbr[rp[v] + k].  Every
// block_GLOB list has more than pp slots, so a node needs its row offset
// before some of its slots anyway, and those builds run without the copy.
// The build runs T threads (TL by default: every level pass and the record
// pass are one pass over the local nodes) for the tiles of the W-thread sweep.
// With Fg set, the build also takes the gradient of the new lists at the
// potentials Fg for the tile's own nodes (every node is in one tile) and the
// max |g| into gmax (as s_sweep_pad<..., false> does for the convergence
// checks that follow a list rebuild): each lane pushes its slots' d into a
// Top3, the node's four are merged by shuffles.  Top3 keeps the same three
// values whatever the order, so g has the bits of the separate gradient pass
// whose launch this saves.
template <int W, int T = TCfg<W>::TL>
__global__ void __launch_bounds__(T) s_tile_build(int n, const int32_t* __restrict__ tst, int pp, int tpp, int depth,
                                                   const int32_t* __restrict__ act, const int32_t* __restrict__ rp,
                                                   const double* __restrict__ D, const int32_t* __restrict__ nbr,
                                                   const ESlot* __restrict__ psl,
                                                   int32_t* __restrict__ hdr, int4* __restrict__ tnode,
                                                   TSlot* __restrict__ tslot, TSlot* __restrict__ tovf,
                                                   int32_t* __restrict__ fit_min, const double* __restrict__ Fg,
                                                   unsigned long long* __restrict__ gmax) {
  constexpr int TW = T, TL = TCfg<W>::TL, THSZ = TCfg<W>::THSZ, TOV = TCfg<W>::TOV;
  static_assert(T % 64 == 0 && T <= 1024, "whole waves, one workgroup");
  // four lanes per local node in the level and record passes: lane q takes
  // slots q, q + 4, ... (a node's padded slots are read together)
  constexpr int LN = 4, NG = T / LN;
  __shared__ int hkey[THSZ];
  __shared__ short hval[THSZ];
  __shared__ int lg[TL];
  __shared__ int s_nl, s_full, s_ov;
  constexpr int LB = 32;                    // list-length bins of the position order
  __shared__ int s_lvl[TMAXS + 2];          // local nodes of level <= d
  constexpr int NK = ((TMAXS + 1) * LB + 63) / 64 * 64;
  __shared__ int s_key[NK];                 // counts, then offsets, of (level, length) keys
  __shared__ short s_pos[TL];               // local index at each position
  const int t = blockIdx.x, tid = threadIdx.x;
  const int q = tid & (LN - 1), g = tid / LN;
  const int t0 = tst[t], c0 = tst[t + 1] - t0;  // the tile: ids [t0, t0 + c0)
  const bool cp = psl != nullptr;  // the padded copy holds slots k < pp
  // slot k of node v, whose row starts at b
  auto slot_at = [&](int v, int b, int k) -> ESlot {
    if (cp && k < pp) return psl[(int64_t)v * pp + k];
    return ESlot{D[b + k], nbr[b + k], 0};
  };
  for (int i = tid; i < THSZ; i += TW) {
    hkey[i] = -1;
    hval[i] = -1;
  }
  for (int j = tid; j < c0; j += TW) lg[j] = t0 + j;
  if (tid == 0) {
    s_nl = c0;
    s_full = 0;
    s_ov = 0;
    hdr[t * THDR] = c0;
    s_lvl[0] = c0;
  }
  __syncthreads();
  const int lane = tid & 63;
  const unsigned long long below = (1ull << lane) - 1;  // lanes under this one
  int lo = 0, hi = c0, fit = 0;
  for (int d = 0; d < depth; ++d) {
    // passes over the frontier [lo, hi), NG nodes each, the same count in every
    // wave (ballots below need whole waves)
    for (int j0 = lo; j0 < hi; j0 += NG) {
      const int j = j0 + g;
      const bool live = j < hi;
      int v = 0, L = 0, rb = 0, wq = -1;
      if (live) {
        v = lg[j];
        L = act[v];
        rb = rp[v];
        // slot q of the padded copy read with the list length, not after it
        // (in bounds for any length): one dependent global read per level
        if (cp && q < pp) wq = psl[(int64_t)v * pp + q].nbr;
      }
      // claim a hash entry for neighbour w (has: the slot exists) outside the
      // tile and not seen yet; local indices for the wave's fresh nodes come
      // from one shared-memory atomic.  Called by every lane of the wave.
      auto claim = [&](bool has, int w) {
        bool fresh = false;
        unsigned hpos = 0;
        if (has && !(w >= t0 && w < t0 + c0)) {
          unsigned h = t_hash<W>(w);
          int probe = 0;
          for (; probe < THSZ; ++probe) {
            const int prev = atomicCAS(&hkey[h], -1, w);
            if (prev == -1) {
              fresh = true;
              hpos = h;
              break;
            }
            if (prev == w) break;
            h = (h + 1) & (THSZ - 1);
          }
          if (probe == THSZ) s_full = 1;
        }
        const unsigned long long m = __ballot(fresh);
        if (m) {
          const int first = __ffsll((long long)m) - 1;
          int base = 0;
          if (lane == first) base = atomicAdd(&s_nl, __popcll(m));
          base = __shfl(base, first);
          if (fresh) {
            const int idx = base + __popcll(m & below);
            if (idx < TL) {
              lg[idx] = w;
              hval[hpos] = (short)idx;
            } else {
              s_full = 1;
            }
          }
        }
      };
      auto nbr_at = [&](int k) { return cp && k < pp ? psl[(int64_t)v * pp + k].nbr : nbr[rb + k]; };
      // the node's slots q, q + LN, ... below KB LN read together, then claimed
      constexpr int KB = 3;
      int wk[KB];
#pragma unroll
      for (int i = 0; i < KB; ++i) {
        const int k = q + LN * i;
        wk[i] = -1;
        if (live && k < L) wk[i] = (cp && i == 0 && q < pp) ? wq : nbr_at(k);
      }
      bool more = true;
#pragma unroll
      for (int i = 0; i < KB; ++i) {
        const bool has = live && q + LN * i < L;
        if (more && !__any(has)) more = false;
        if (more) claim(has, wk[i]);
      }
      if (more) {
        for (int k = q + LN * KB;; k += LN) {
          const bool has = live && k < L;
          if (!__any(has)) break;
          claim(has, has ? nbr_at(k) : -1);
        }
      }
    }
    __syncthreads();
    const int nl = s_nl;
    const bool full = s_full != 0;
    __syncthreads();
    if (full) break;  // level d+1 incomplete: keep levels <= d
    lo = hi;
    hi = nl;
    fit = d + 1;
    if (tid == 0) {
      hdr[t * THDR + d + 1] = hi;
      s_lvl[d + 1] = hi;
    }
  }
  // position order: key (level, longest list first), a counting sort
  const bool by_len = tpp > 4;
  if (!by_len) {
    for (int j = tid; j < hi; j += TW) s_pos[j] = (short)j;
  } else {
  for (int i = tid; i < NK; i += TW) s_key[i] = 0;
  __syncthreads();
  constexpr int KP = (TL + TW - 1) / TW;
  int kj[KP];
#pragma unroll
  for (int r = 0; r < KP; ++r) {
    const int jj = tid + r * TW;
    kj[r] = -1;
    if (jj < hi) {
      int d = 0;
      while (jj >= s_lvl[d]) ++d;
      kj[r] = d * LB + (LB - 1 - min(act[lg[jj]], LB - 1));
      atomicAdd(&s_key[kj[r]], 1);
    }
  }
  __syncthreads();
  if (tid < 64) {  // exclusive scan of the counts by one wave, NK / 64 per lane
    constexpr int PL = NK / 64;
    static_assert(NK % 64 == 0, "whole lanes");
    int c[PL], sum = 0;
#pragma unroll
    for (int i = 0; i < PL; ++i) {
      c[i] = s_key[tid * PL + i];
      sum += c[i];
    }
    int incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(incl, off, 64);
      if (tid >= off) incl += y;
    }
    int base = incl - sum;
#pragma unroll
    for (int i = 0; i < PL; ++i) {
      s_key[tid * PL + i] = base;
      base += c[i];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < KP; ++r)
    if (kj[r] >= 0) s_pos[atomicAdd(&s_key[kj[r]], 1)] = (short)(tid + r * TW);
  }
  __syncthreads();
  // records: every pass's list lengths and row offsets are read before any
  // is used, then every pass's slots q and q + LN (rq, rq2) together
  constexpr int RP = (TL + NG - 1) / NG;
  int rv[RP], rL[RP], rb[RP];
  ESlot rq[RP], rq2[RP];
#pragma unroll
  for (int r = 0; r < RP; ++r) {
    const int j = g + r * NG;
    rv[r] = rL[r] = rb[r] = 0;
    rq[r] = rq2[r] = ESlot{0.0, 0, 0};
    if (j < hi) {
      const int v = lg[s_pos[j]];
      rv[r] = v;
      rL[r] = act[v];
      rb[r] = rp[v];
      if (cp && q < pp) rq[r] = psl[(int64_t)v * pp + q];
    }
  }
#pragma unroll
  for (int r = 0; r < RP; ++r) {
    if (!(cp && q < pp) && q < rL[r]) rq[r] = slot_at(rv[r], rb[r], q);
    if (q + LN < rL[r]) rq2[r] = slot_at(rv[r], rb[r], q + LN);
  }
#pragma unroll
  for (int r = 0; r < RP; ++r) {  // the four lanes of a node run together
    const int j = g + r * NG;
    // overflow offsets: a wave's exclusive prefix sum of its nodes' overflow
    // counts plus one shared-memory atomic per wave
    const int x = (j < hi && q == 0 && rL[r] > tpp) ? rL[r] - tpp : 0;
    int incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(incl, off);
      if (lane >= off) incl += y;
    }
    int base = 0;
    if (lane == 63 && incl > 0) base = atomicAdd(&s_ov, incl);
    base = __shfl(base, 63);
    if (j >= hi) continue;
    const int v = rv[r], L = rL[r], b = rb[r], jj = s_pos[j];
    const int ovo = __shfl(base + incl - x, 0, LN);
    if (q == 0) tnode[(int64_t)t * TL + j] = make_int4(v, L, ovo, jj);
    auto rec = [&](int k, const ESlot& e) {
      int loc = -1;
      if (e.nbr >= t0 && e.nbr < t0 + c0) {
        loc = e.nbr - t0;
      } else {
        unsigned h = t_hash<W>(e.nbr);
        for (int probe = 0; probe < THSZ; ++probe) {
          const int key = hkey[h];
          if (key == e.nbr) {
            loc = hval[h];
            break;
          }
          if (key == -1) break;
          h = (h + 1) & (THSZ - 1);
        }
        if (loc >= hi) loc = -1;  // from an incomplete level (neighbour of the deepest kept level)
      }
      const TSlot ts{e.D, e.nbr, loc};
      if (k < tpp) tslot[((int64_t)t * TL + j) * tpp + k] = ts;
      else if (ovo + (k - tpp) < TOV) tovf[(int64_t)t * TOV + ovo + (k - tpp)] = ts;
    };
    if (q < L) rec(q, rq[r]);
    if (q + LN < L) rec(q + LN, rq2[r]);
    for (int k = q + 2 * LN; k < L; k += LN) rec(k, slot_at(v, b, k));
    // padded slots past the list: D = +inf and the node itself, so that a
    // sweep computes them unmasked, as +inf (above every sentinel of Top3)
    for (int k = q; k < tpp; k += LN)
      if (k >= L) tslot[((int64_t)t * TL + j) * tpp + k] = TSlot{__builtin_inf(), v, jj};
  }
  if (Fg) {  // gradient of the tile's own nodes (positions below c0, the first RP0 passes)
    constexpr int RP0 = (TS_W + NG - 1) / NG;
    static_assert(RP0 <= RP, "the tile's nodes are among the recorded ones");
    double gm = 0.0;
#pragma unroll
    for (int r = 0; r < RP0; ++r) {
      const int j = g + r * NG;
      const bool own = j < c0;  // the same for a node's four lanes
      Top3 t3;
      int v = 0;
      if (own) {
        v = rv[r];
        const int L = rL[r], b = rb[r];
        const double Fv = Fg[v];
        if (q < L) t3.push(dval(rq[r].D, Fv, Fg[rq[r].nbr], v, rq[r].nbr));
        if (q + LN < L) t3.push(dval(rq2[r].D, Fv, Fg[rq2[r].nbr], v, rq2[r].nbr));
        for (int k = q + 2 * LN; k < L; k += LN) {
          const ESlot e = slot_at(v, b, k);
          t3.push(dval(e.D, Fv, Fg[e.nbr], v, e.nbr));
        }
      }
#pragma unroll
      for (int off = 1; off < LN; off <<= 1) {
        const double a2 = __shfl_xor(t3.a, off, LN), b2 = __shfl_xor(t3.b, off, LN), c2 = __shfl_xor(t3.c, off, LN);
        t3.push(a2);
        t3.push(b2);
        t3.push(c2);
      }
      if (own && q == 0) gm = fmax(gm, fabs(t3.grad()));
    }
    tile_max_to<T>(gm, gmax + (t & 31) * 16);
  }
  __syncthreads();
  if (tid == 0) {
    const int f = s_ov > TOV ? -1 : fit;
    hdr[t * THDR + THDR_OV] = min(s_ov, TOV);
    hdr[t * THDR + THDR_FIT] = f;
    atomicMin(fit_min, f);
  }
}

// Tile boundaries for the next build of a list kind, from this build's
// headers.  A launch lasts as long as its slowest tile, and a tile's time
// follows its work: the node updates of the launch's S sweeps (sum over the
// levels d < S of the local nodes of level <= d) plus ovw per overflow slot
// (an overflow slot adds a serial step to a sweep of its wave; on the n = 100k
// Gaussian mixture a fit of the tiles' times gives ~280 node updates per
// overflow slot for the block_EF lists, ~1.5 for the block_GLOB lists).  The
// new boundaries cut the cumulative work, taken as uniform over each tile's
// ids, into equal parts, halfway from the old boundaries.  Then every tile is
// clamped to between 1 and cap ids (a pass from the first boundary, then one
// from the last); if that still fails, the old boundaries stay.  (Falling
// back to the uniform split t*ts instead, as before, undid the balance
// whenever a light tile's share passed the cap: on the n = 100k Gaussian
// mixture the block_EF tiles went back to uniform after seven builds.)  One
// workgroup of BAL_T threads; T <= BAL_MAX tiles (not called above).
// Deterministic: the headers are set sizes.
constexpr int BAL_T = 1024, BAL_MAX = 4096;
__global__ void __launch_bounds__(BAL_T) s_tile_balance(int T, int n, int ts, int cap, int fuse, double ovw,
                                                       const int32_t* __restrict__ hdr,
                                                       const int32_t* __restrict__ fit_min, int32_t* __restrict__ tst) {
  __shared__ double P[BAL_MAX + 1];
  __shared__ int old[BAL_MAX + 1];
  __shared__ double part[BAL_T / 64];
  __shared__ int bad;
  const int tid = threadIdx.x;
  const int S = min(fuse, *fit_min + 1);
  for (int t = tid; t <= T; t += BAL_T) old[t] = tst[t];
  if (tid == 0) bad = 0;
  __syncthreads();
  // work per tile, then an exclusive prefix sum: thread i sums its chunk
  const int ch = (T + BAL_T - 1) / BAL_T, a = tid * ch, b = min(T, a + ch);
  double sum = 0.0;
  for (int t = a; t < b; ++t) {
    const int32_t* h = hdr + (int64_t)t * THDR;
    double wk = ovw * h[THDR_OV];
    for (int d = 0; d < S; ++d) wk += h[d];
    P[t] = wk;
    sum += wk;
  }
  // exclusive scan of the chunk sums: within each wave by shuffles, then the
  // wave totals (in order: the bits do not depend on timing)
  const int lane = tid & 63, wv = tid >> 6;
  double inc = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double y = __shfl_up(inc, off);
    if (lane >= off) inc += y;
  }
  if (lane == 63) part[wv] = inc;
  __syncthreads();
  if (tid == 0) {
    double run = 0.0;
    for (int i = 0; i < BAL_T / 64; ++i) {
      const double x = part[i];
      part[i] = run;
      run += x;
    }
  }
  __syncthreads();
  double acc = part[wv] + (inc - sum);
  for (int t = a; t < b; ++t) {
    const double x = P[t];
    P[t] = acc;
    acc += x;
  }
  if (b == T && a < b) P[T] = acc;
  __syncthreads();
  const double tot = P[T];
  // boundary k: the id where the cumulative work reaches k/T of the total
  int nb[(BAL_MAX + BAL_T - 1) / BAL_T + 1];
  int m = 0;
  for (int k = tid; k <= T; k += BAL_T, ++m) {
    int r = k == 0 ? 0 : n;
    if (k > 0 && k < T) {
      const double target = tot * k / T;
      int lo = 0, hi = T - 1;  // last u with P[u] <= target
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (P[mid] <= target) lo = mid;
        else hi = mid - 1;
      }
      const double wu = P[lo + 1] - P[lo];
      const double f = wu > 0.0 ? fmin(1.0, (target - P[lo]) / wu) : 0.0;
      r = old[lo] + (int)(f * (old[lo + 1] - old[lo]));
    }
    nb[m] = (int)(((int64_t)r + old[k]) / 2);
  }
  __syncthreads();  // every read of old is done
  m = 0;
  for (int k = tid; k <= T; k += BAL_T, ++m) old[k] = nb[m];
  __syncthreads();
  for (int k = tid; k < T; k += BAL_T) {
    const int sz = old[k + 1] - old[k];
    if (sz < 1 || sz > cap) bad = 1;
  }
  __syncthreads();
  if (bad) {  // tile sizes into [1, cap]: from the first boundary, then from the last
    if (tid == 0) {
      int prev = 0;
      old[0] = 0;
#pragma unroll 8
      for (int k = 1; k < T; ++k) {
        prev = min(max(old[k], prev + 1), prev + cap);
        old[k] = prev;
      }
      old[T] = prev = n;
#pragma unroll 8
      for (int k = T - 1; k >= 1; --k) {
        prev = min(max(old[k], prev - cap), prev - 1);
        old[k] = prev;
      }
      bad = 0;
    }
    __syncthreads();
    for (int k = tid; k < T; k += BAL_T) {
      const int sz = old[k + 1] - old[k];
      if (sz < 1 || sz > cap) bad = 1;
    }
    __syncthreads();
  }
  const bool ok = bad == 0 && old[0] == 0 && old[T] == n;
  if (ok)
    for (int k = tid; k <= T; k += BAL_T) tst[k] = old[k];
}
// uniform tile boundaries t*ts (the last tile ends at n)
__global__ void s_tile_uniform(int T, int n, int ts, int32_t* __restrict__ tst) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= T) tst[k] = (int)min((int64_t)k * ts, (int64_t)n);
}

// Top3 over the padded slots, every slot pushed, unmasked: a slot past the
// list holds D = +inf (s_tile_build), so it gives +inf, which leaves the
// three values unchanged (each is <= 1e6), and no lane diverges.
template <int TPP>
__device__ __forceinline__ void push_padded(Top3& t3, const TSlot* sl, const double* hn, double hv, int v) {
  static_assert(TPP >= 4, "the first four slots go through top3_of4");
  double d[TPP];
#pragma unroll
  for (int k = 0; k < TPP; ++k) d[k] = dval_h(sl[k].D, hv, hn[k], v, sl[k].nbr);
  t3 = top3_of4(d[0], d[1], d[2], d[3]);  // t3 is fresh: the same as pushing them
#pragma unroll
  for (int k = 4; k < TPP; ++k) t3.push(d[k]);
}

// S sweeps over the tiles (S <= fit_min + 1); Fin -> Fout, intermediate
// sweeps in shared memory.  maxslots: max |g| of the last sweep, as
// s_sweep_pad.  Two dependent HBM levels: the records (the first TW local
// nodes' before the header is read: the tile and most halos), then the
// potentials of every slot.  Overflow slots (beyond TPP, which is chosen
// above the lists' usual length) are staged in shared memory with their
// potentials, the first TW of them loaded with the records; the rest (a
// third level) only when a tile has more.  Sweeps 2..S read shared memory.
// Tile of workgroup b.  nxcd <= 1: tile b.  Else the workgroups that the
// dispatcher sends to one XCD (b = x, x + nxcd, ...: round robin over the
// XCDs) take consecutive tiles, so the tiles that share halo nodes (Morton
// neighbours) read their potentials through the same L2.  A bijection of
// [0, ntiles): XCD x gets q + (x < r) workgroups and tiles from x q + min(x, r).
__device__ __forceinline__ int tile_of_block(int b, int ntiles, int nxcd) {
  if (nxcd <= 1) return b;
  const int x = b % nxcd, i = b / nxcd, q = ntiles / nxcd, r = ntiles % nxcd;
  return x * q + min(x, r) + i;
}

template <int S, bool FROZEN, int TPP, int W, bool SELH = false>
__global__ void __launch_bounds__(W) s_sweep_tile(const int32_t* __restrict__ hdr, const int4* __restrict__ tnode,
                                                   const TSlot* __restrict__ tslot, const TSlot* __restrict__ tovf,
                                                   const double* __restrict__ Fin, double* __restrict__ Fout,
                                                   const uint8_t* __restrict__ frozen,
                                                   unsigned long long* __restrict__ maxslots, int nxcd) {
  constexpr int TW = W, TL = TCfg<W>::TL, TOV = TCfg<W>::TOV;
  constexpr int NP = TL / TW;
  __shared__ double fb[2][TL];
  __shared__ TSlot ov[TOV];
  __shared__ double ovF[TOV];
  const int t = tile_of_block(blockIdx.x, gridDim.x, nxcd), tid = threadIdx.x;
  const int32_t* h = hdr + t * THDR;
  const TSlot orec = tovf[(int64_t)t * TOV + tid];  // speculative: tid < TW <= TOV
  int cnt[S];  // local nodes of level <= d, read once (kept in registers across the barriers)
#pragma unroll
  for (int d = 0; d < S; ++d) cnt[d] = h[d];
  const int nloc = cnt[S - 1];
  const int nov = h[THDR_OV];
  int4 nd[NP];
  TSlot sl[NP][TPP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int j = tid + p * TW;
    nd[p] = make_int4(0, 0, 0, 0);
    if (p == 0 || j < nloc) {
      const int64_t r = (int64_t)t * TL + j;
      nd[p] = tnode[r];
#pragma unroll
      for (int k = 0; k < TPP; ++k) sl[p][k] = tslot[r * TPP + k];
    }
  }
  double fv[NP], fw[NP][TPP];
  bool fz[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int j = tid + p * TW;
    fz[p] = false;
    fv[p] = 0.0;
    if (j < nloc) {
      fv[p] = Fin[nd[p].x];
      if (FROZEN) fz[p] = frozen[nd[p].x] != 0;
#pragma unroll
      for (int k = 0; k < TPP; ++k) fw[p][k] = Fin[sl[p][k].nbr];  // padded slots: the node itself
    }
  }
  double of = 0.0;
  if (tid < nov) of = Fin[orec.nbr];
  for (int i = tid + TW; i < nov; i += TW) {
    const TSlot o = tovf[(int64_t)t * TOV + i];
    ov[i] = o;
    ovF[i] = Fin[o.nbr];
  }
  if (tid < nov) {
    ov[tid] = orec;
    ovF[tid] = of;
  }
  __syncthreads();
  // sweep 1: every local node of level <= S-1.  Shared memory holds the
  // halves 0.5*F of the local potentials (the operands of dval_h); a thread
  // keeps its nodes' own potentials in registers.
  double gabs = 0.0;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int j = tid + p * TW;
    if (j < nloc) {
      const int v = nd[p].x, L = nd[p].y;
      const double hv = 0.5 * fv[p];
      double hn[TPP];
#pragma unroll
      for (int k = 0; k < TPP; ++k) hn[k] = 0.5 * fw[p][k];
      Top3 t3;
      push_padded<TPP>(t3, sl[p], hn, hv, v);
#pragma unroll 1
      for (int k = TPP; k < L; ++k) {
        const int i = (nd[p].z + k - TPP) & (TOV - 1);
        t3.push(dval(ov[i].D, fv[p], ovF[i], v, ov[i].nbr));
      }
      const double g = t3.grad();
      const double f = (FROZEN && fz[p]) ? fv[p] : fv[p] + g;
      if (S == 1) {
        Fout[v] = f;
        gabs = fmax(gabs, fabs(g));
      } else {
        fv[p] = f;
        fb[0][nd[p].w] = 0.5 * f;
      }
    }
  }
  // byte offsets into a half buffer of each slot's two dval_h operands, the
  // lower id's half first: the operand order is fixed per slot, so sweeps
  // 2..S read the two halves in that order and add them without selects
  // (padded slots: the node itself twice, D = +inf).  The n = 100k
  // Gaussian-mixture solve: 21.73 -> 20.99 ms against selecting the halves
  // by value each sweep.
  // SELH: instead, oa = the neighbour's offset, ob = 1 when the owner has
  // the lower id, and the owner's half comes from its register, chosen by
  // selects.  Used for the block_GLOB lists, whose records are ordered by
  // list length, so the owners' halves sit at scattered local indices and
  // their shared-memory reads conflict: n = 100k Gaussian mixture 20.54 ->
  // 20.40 ms.  For the block_EF lists (owners' halves consecutive) the
  // selects cost more than the reads: 20.34 against 20.42 ms.
  uint32_t oa[NP][TPP], ob[NP][TPP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const uint32_t own = (uint32_t)t_loc<W>(nd[p].w) * 8u;
#pragma unroll
    for (int k = 0; k < TPP; ++k) {
      const uint32_t nb = t_loc<W>(sl[p][k].loc) * 8u;
      const bool lo = nd[p].x < sl[p][k].nbr;
      oa[p][k] = SELH ? nb : (lo ? own : nb);
      ob[p][k] = SELH ? (lo ? 1u : 0u) : (lo ? nb : own);
    }
  }
  // sweeps 2..S from shared memory: sweep q for the local nodes of level <= S-q
#pragma unroll
  for (int q = 2; q <= S; ++q) {
    __syncthreads();
    const double* src = fb[(q - 2) & 1];
    double* dst = fb[(q - 1) & 1];
    const int cq = cnt[S - q];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int j = tid + p * TW;
      if (j < cq) {
        const int v = nd[p].x, L = nd[p].y;
        const double f0 = fv[p], hv = 0.5 * f0;
        const char* sb = reinterpret_cast<const char*>(src);
        double d[TPP];  // dval_h of every padded slot
        if (SELH) {  // the owner's half from its register, the neighbour's from shared memory
#pragma unroll
          for (int k = 0; k < TPP; ++k) {
            const double hw = *reinterpret_cast<const double*>(sb + oa[p][k]);
            const bool lo = (ob[p][k] & 1u) != 0;
            d[k] = (sl[p][k].D + (lo ? hv : hw)) + (lo ? hw : hv);
          }
        } else {
#pragma unroll
        for (int k = 0; k < TPP; ++k)
          d[k] = (sl[p][k].D + *reinterpret_cast<const double*>(sb + oa[p][k])) +
                 *reinterpret_cast<const double*>(sb + ob[p][k]);
        }
        Top3 t3 = top3_of4(d[0], d[1], d[2], d[3]);  // the same as pushing them
#pragma unroll
        for (int k = 4; k < TPP; ++k) t3.push(d[k]);
#pragma unroll 1
        for (int k = TPP; k < L; ++k) {
          const TSlot o = ov[(nd[p].z + k - TPP) & (TOV - 1)];
          t3.push(dval_h(o.D, hv, src[t_loc<W>(o.loc)], v, o.nbr));
        }
        const double g = t3.grad();
        const double f = (FROZEN && fz[p]) ? f0 : f0 + g;
        if (q == S) {
          Fout[v] = f;
          gabs = fmax(gabs, fabs(g));
        } else {
          fv[p] = f;
          dst[nd[p].w] = 0.5 * f;
        }
      }
    }
  }
  if (maxslots) tile_max_to<W>(gabs, maxslots + (blockIdx.x & 31) * 16);
}

// cuda_partition_v4 (graph_cuda_partition.hpp:89-115)
__global__ void s_part_v4(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ deg,
                          int32_t* __restrict__ act, int kk, int32_t* __restrict__ nbr, double* __restrict__ D,
                          int32_t* __restrict__ orig, const double* __restrict__ F) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) act[v] = part_v4_node(v, rp[v], deg[v], kk, nbr, D, orig, F);
}

// block_GLOB's partition with PV4_LANES lanes per node (part_v4_lanes)
__global__ void __launch_bounds__(BS) s_part_v4_q(int n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ deg, int32_t* __restrict__ act, int kk,
                                                  int32_t* __restrict__ nbr, double* __restrict__ D,
                                                  int32_t* __restrict__ orig, const double* __restrict__ F) {
  const int64_t t = (int64_t)blockIdx.x * BS + threadIdx.x;
  const int v = (int)(t / PV4_LANES), q = (int)(t % PV4_LANES);
  if (v >= n) return;  // the four lanes of a node leave together
  const int dg = deg[v];
  if (dg < kk) {
    if (q == 0) act[v] = dg;
    return;
  }
  const int p = part_v4_lanes(v, q, rp[v], dg, kk, nbr, D, orig, F);
  if (q == 0) act[v] = p;
}

// check_opt's partition with PV4_LANES lanes per node (part_v2_lanes)
__global__ void __launch_bounds__(BS) s_part_v2_q(int n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ deg, int32_t* __restrict__ act,
                                                  double thr, int32_t* __restrict__ nbr, double* __restrict__ D,
                                                  int32_t* __restrict__ orig, const double* __restrict__ F) {
  const int64_t t = (int64_t)blockIdx.x * BS + threadIdx.x;
  const int v = (int)(t / PV4_LANES), q = (int)(t % PV4_LANES);
  if (v >= n) return;  // the four lanes of a node leave together
  const int dg = deg[v];
  if (dg < 10) {
    if (q == 0) act[v] = dg;
    return;
  }
  const int p = part_v2_lanes(v, q, rp[v], dg, thr, nbr, D, orig, F);
  if (q == 0) act[v] = p;
}

// cuda_partition_v2 (graph_cuda_partition.hpp:57-69)
__global__ void s_part_v2(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ deg,
                          int32_t* __restrict__ act, double thr, int32_t* __restrict__ nbr,
                          double* __restrict__ D, int32_t* __restrict__ orig, const double* __restrict__ F) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) act[v] = part_v2_node(v, rp[v], deg[v], thr, nbr, D, orig, F);
}

// cuda_partition_v3 with threshold 0.25 over the active prefix + class counts
__global__ void s_classify(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ act,
                           int32_t* __restrict__ clen, int32_t* __restrict__ A, int32_t* __restrict__ T,
                           double eps, int32_t* __restrict__ nbr, double* __restrict__ D,
                           int32_t* __restrict__ orig, const double* __restrict__ F) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int cl, a, t;
  classify_node(v, rp[v], act[v], eps, nbr, D, orig, F, cl, a, t);
  clen[v] = cl;
  A[v] = a;
  T[v] = t;
}

// ---- union-find over tight edges (roots are component minima) ----------
__global__ void s_uf_init(int n, int32_t* par) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) par[v] = v;
}
__global__ void s_uf_union(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ clen, double eps,
                           const int32_t* __restrict__ nbr, const double* __restrict__ D,
                           const double* __restrict__ F, int32_t* par) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int b = rp[v];
  const double Fv = F[v];
  for (int k = 0; k < clen[v]; ++k) {
    const int w = nbr[b + k];
    if (w <= v) continue;
    const double d = dval(D[b + k], Fv, F[w], v, w);
    if (d >= -eps && d <= eps) uf_unite(par, v, w);
  }
}
__global__ void s_uf_root(int n, const int32_t* __restrict__ T, int32_t* par, int32_t* __restrict__ root) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) root[v] = T[v] > 0 ? uf_find(par, v) : -1;
}

// comb regularity (comb::solve_simple, comb_thread.cpp:180-195): irregular if
// some node has 2A + T != 4.
__global__ void s_irr_mark(int n, const int32_t* __restrict__ root, const int32_t* __restrict__ A,
                           const int32_t* __restrict__ T, int32_t* __restrict__ irr) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int r = root[v];
  if (r >= 0 && 4 - 2 * A[v] - T[v] != 0) irr[r] = 1;
}

// x per working slot (primal_node)
__global__ void s_primal(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ deg,
                         const int32_t* __restrict__ clen, double eps, const int32_t* __restrict__ nbr,
                         const double* __restrict__ D, const double* __restrict__ F,
                         const int32_t* __restrict__ root, const int32_t* __restrict__ irr,
                         double* __restrict__ xw, uint8_t* __restrict__ flag) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int r = root[v];
  primal_node(v, rp[v], deg[v], clen[v], eps, nbr, D, F, r >= 0 && irr[r] == 0, xw, flag);
}

__global__ void s_to_canon(int64_t m, const int32_t* __restrict__ orig, const double* __restrict__ xw,
                           const uint8_t* __restrict__ fw, double* __restrict__ xc, uint8_t* __restrict__ fc) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  const int o = orig[s];
  xc[o] = xw[s];
  fc[o] = fw[s];
}

// ---- irregular comb node lists --------------------------------------------
__global__ void s_irr_count(int n, const int32_t* __restrict__ root, const int32_t* __restrict__ irr,
                            int32_t* __restrict__ cnt) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int r = root[v];
  if (r >= 0 && irr[r]) atomicAdd(&cnt[r], 1);
}
__global__ void s_flag_pos(int n, const int32_t* __restrict__ cnt, int32_t* __restrict__ f) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) f[v] = cnt[v] > 0 ? 1 : 0;
}
__global__ void s_root_list(int n, const int32_t* __restrict__ cnt, const int32_t* __restrict__ ridx,
                            int32_t* __restrict__ roots) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n && cnt[v] > 0) roots[ridx[v]] = v;
}
__global__ void s_irr_fill(int n, const int32_t* __restrict__ root, const int32_t* __restrict__ irr,
                           const int32_t* __restrict__ off, int32_t* __restrict__ cur, int32_t* __restrict__ list) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int r = root[v];
  if (r >= 0 && irr[r]) list[off[r] + atomicAdd(&cur[r], 1)] = v;
}
// Ascending sort of each comb's node segment: one block per comb, bitonic
// network with implicit +inf padding (mirror-compare variant, so padding
// stays at the tail and out-of-range pairs are skipped).
__global__ void s_seg_sort(const int32_t* __restrict__ roots, const int32_t* __restrict__ off,
                           const int32_t* __restrict__ cnt, int32_t* __restrict__ list) {
  const int r = roots[blockIdx.x];
  int32_t* a = list + off[r];
  const int m = cnt[r];
  int p2 = 1;
  while (p2 < m) p2 <<= 1;
  for (int size = 2; size <= p2; size <<= 1) {
    const int half = size >> 1;
    for (int idx = threadIdx.x; idx < p2 / 2; idx += blockDim.x) {
      const int blk = idx / half, pos = idx % half;
      const int i = blk * size + pos, j = blk * size + size - 1 - pos;
      if (j < m && a[i] > a[j]) { int t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    __syncthreads();
    for (int k = half >> 1; k >= 1; k >>= 1) {
      for (int idx = threadIdx.x; idx < p2 / 2; idx += blockDim.x) {
        const int blk = idx / k, pos = idx % k;
        const int i = blk * 2 * k + pos, j = i + k;
        if (j < m && a[i] > a[j]) { int t = a[i]; a[i] = a[j]; a[j] = t; }
      }
      __syncthreads();
    }
  }
}

// least squares (axb_comb) for one irregular comb per block
struct RP32 {
  const int32_t* rp;
  __device__ __forceinline__ int64_t operator()(int v) const { return rp[v]; }
};
__global__ void __launch_bounds__(BS) s_axb(const int32_t* __restrict__ roots, const int32_t* __restrict__ off,
                                            const int32_t* __restrict__ cnt, const int32_t* __restrict__ list,
                                            const int32_t* __restrict__ rp, const int32_t* __restrict__ clen,
                                            const int32_t* __restrict__ A, const int32_t* __restrict__ T,
                                            double eps, const int32_t* __restrict__ nbr, const double* __restrict__ D,
                                            const double* __restrict__ F, double* __restrict__ xw,
                                            double* __restrict__ xg, double* __restrict__ rn,
                                            int32_t* __restrict__ status) {
  __shared__ double sh[BS / 64];
  const int r = roots[blockIdx.x];
  const int go = axb_comb(list + off[r], cnt[r], RP32{rp}, clen, A, T, eps, nbr, D, F, xw, xg, rn, sh);
  if (threadIdx.x == 0) status[blockIdx.x] = go;
}

// shift_comb_IRREGULAR_B (graph_cuda_comb.hpp:365-403): one thread per
// IRREGULAR_B comb; "first violating node" = smallest id.
__global__ void s_shift(int nb, const int32_t* __restrict__ broots, const int32_t* __restrict__ off,
                        const int32_t* __restrict__ cnt, const int32_t* __restrict__ list,
                        const int32_t* __restrict__ A, const int32_t* __restrict__ T, double sv, int type,
                        double* __restrict__ F, uint8_t* __restrict__ frozen) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nb) return;
  const int r = broots[c];
  for (int i = 0; i < cnt[r]; ++i) {
    const int v = list[off[r] + i];
    const int s = 4 - 2 * A[v] - T[v];
    if (s != 0) {
      F[v] += sv * (s > 0 ? 1.0 : -1.0);
      frozen[v] = 1;
      if (type == 1) break;
    }
  }
}

// ---- deterministic reductions (fixed grid, fixed per-thread order, fixed tree)
__device__ __forceinline__ double block_sum_det(double v) {
  __shared__ double wsum[BS / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wsum[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += wsum[k];
  return s;
}
// partials: [0] objective, [1] sum F, [2] sum min(0,d).  SUM_LANES lanes per
// node (lane q: slots q, q + 4, ..., read together four at a time); the
// order of the sums is fixed by the launch shape, so the results are
// deterministic.
constexpr int SUM_LANES = 4;
__global__ void __launch_bounds__(BS) s_sums(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ gnbr,
                                             const double* __restrict__ gcost, const double* __restrict__ xc,
                                             const int32_t* __restrict__ wnbr, const double* __restrict__ wD,
                                             const double* __restrict__ F, double* __restrict__ part) {
  double obj = 0.0, sf = 0.0, sneg = 0.0;
  const int q = threadIdx.x % SUM_LANES;
  const int64_t groups = (int64_t)gridDim.x * (BS / SUM_LANES);
  for (int64_t vv = (int64_t)blockIdx.x * (BS / SUM_LANES) + threadIdx.x / SUM_LANES; vv < n; vv += groups) {
    const int v = (int)vv;
    const int b = rp[v], e = rp[v + 1];
    const double Fv = F[v];
    if (q == 0) sf += Fv;
    for (int s0 = b + q; s0 < e; s0 += 4 * SUM_LANES) {
      int gw[4], ww[4];
      double x[4], cs[4], dd[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int s = s0 + u * SUM_LANES < e ? s0 + u * SUM_LANES : e - 1;  // past the list: unused
        gw[u] = gnbr[s];
        x[u] = xc[s];
        cs[u] = gcost[s];
        ww[u] = wnbr[s];
        dd[u] = wD[s];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (s0 + u * SUM_LANES >= e) continue;
        if (gw[u] > v) obj += x[u] * cs[u];
        if (ww[u] > v) sneg += fmin(0.0, dval(dd[u], Fv, F[ww[u]], v, ww[u]));
      }
    }
  }
  obj = block_sum_det(obj);
  __syncthreads();
  sf = block_sum_det(sf);
  __syncthreads();
  sneg = block_sum_det(sneg);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = obj;
    part[RED_BLOCKS + blockIdx.x] = sf;
    part[2 * RED_BLOCKS + blockIdx.x] = sneg;
  }
}
__global__ void s_sums_final(const double* __restrict__ part, double* __restrict__ out) {
  // one block of BS threads, RED_BLOCKS partials per quantity, fixed order
  for (int q = 0; q < 3; ++q) {
    double v = 0.0;
    for (int i = threadIdx.x; i < RED_BLOCKS; i += BS) v += part[q * RED_BLOCKS + i];
    v = block_sum_det(v);
    if (threadIdx.x == 0) out[q] = v;
    __syncthreads();
  }
}

// MAX edge rank (graph_tour.hpp:169-180): over edges in X, min of the two
// endpoint ranks (0-based position in the node's cost-sorted list).
// With mir (the exact stage's mirror slots: v's position in w's list), the
// reverse rank is read instead of searched for.
__global__ void s_maxrank(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ gnbr,
                          const uint8_t* __restrict__ fc, const int32_t* __restrict__ mir,
                          int32_t* __restrict__ out) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int best = 0;
  for (int s = rp[v]; s < rp[v + 1]; ++s) {
    const int w = gnbr[s];
    if (w <= v || !fc[s]) continue;
    int r = s - rp[v];
    if (mir) {
      r = min(r, mir[s] - rp[w]);
    } else {
      for (int q = rp[w]; q < rp[w + 1]; ++q)
        if (gnbr[q] == v) { r = min(r, q - rp[w]); break; }
    }
    best = max(best, r);
  }
  if (best > 0) atomicMax(out, best);
}

}  // namespace

// ===========================================================================
// Host driver
// ===========================================================================
namespace {

// device max slots: [0, 512) gradient max (32 slots x 16), [512, 1024)
// early-exit max of the last sweep of a chunk
constexpr int MAXS = 1024;

struct Runner {
  f2m_ctx* c;
  SolveWs& w;
  f2m_opts o;
  int n;
  int64_t m;
  const int32_t* rp;    // canonical rowptr (dev)
  const int32_t* gnbr;  // canonical nbr (dev)
  const double* gcost;  // canonical cost (dev)
  double dmin = 0, eps = 0;
  int64_t sweeps = 0;
  int lpn = 4;
  bool padded = true;  // padded layout (default) or compacted lists (F2M_SWEEP=compact)
  int pad_cfg = 2;     // padded variant (stride,lanes): see launch_padded; default 4,2
  int pp = 4;          // padded stride of pad_cfg
  // multi-sweep launches over spatial tiles (padded layout only): up to `fuse`
  // sweeps per launch (1: one sweep per launch, s_sweep_pad); the tiles are
  // rebuilt with the lists, and usable once their overflow flag reads clear
  int fuse = 1;
  int ntiles = 0;                    // tiles of the current lists
  int ts_ef = TS, ts_glob = TS;  // nodes per tile of the block_EF / block_GLOB lists
  int tw = TW_EF, tw_ef = TW_EF, tw_glob = TW_EF;  // tile workgroup shape: current, block_EF's, block_GLOB's
  int tile_xcd = kXcds;              // XCDs for the tile -> workgroup mapping (F2M_TILE_XCD; 1: tile = workgroup)
  // tile boundaries balanced by work (s_tile_balance) for the block_EF lists
  // (F2M_TILE_BALANCE=1, default), for both list kinds (2) or for neither (0).
  // n = 100k Gaussian mixture: 20.81 -> 20.60 ms per solve with the block_EF
  // lists balanced.  Balancing the block_GLOB lists too evens their local sets,
  // so more levels fit (5.4 sweeps per launch instead of 5.0) but the heavier
  // launches made the solve slower (21.19 ms).
  int balance = 1;
  int64_t tile_cap = 0;              // most tiles of either list kind
  bool tiles_pending = false;
  int width = 1;  // sweeps per launch on the current lists: min(fuse, fit_min + 1)
  // halo levels the next build of short (0) / long (1) lists runs: one more
  // than the last build of that kind kept (0: fuse - 1).  Levels no launch
  // uses are not built; the sweeps' bits do not depend on the width.
  int depth_next[2] = {0, 0};
  // builds of each kind so far: the level past the last fit is probed on
  // every probe_every-th build only (F2M_TILE_PROBE; 1: every build)
  int builds[2] = {0, 0};
  int probe_every = 8;
  int pending_kind = 0;
  int tpp = 4;    // padded slots per local node on the current lists
  bool selh = true;  // block_GLOB tile sweeps take the owner's half from a register (F2M_TILE_SELH)
  // the convergence checks after a list rebuild take their gradient from the
  // tile build instead of a separate pass (F2M_BUILD_GRAD=0: the pass)
  bool build_grad = true;
  bool grad_ready = false;  // region 0 holds the last build's max |g|
  // what the tile build reads (F2M_BUILD_ROWS): 0 the full padded lists,
  // built with every list rebuild; 1 the rows, plus the padded copy of the
  // first slots for the block_EF lists; 2 the rows only.  The full padded
  // lists are then built only before a padded sweep uses them.
  int build_rows = 1;
  bool mir_ok = false;  // the exact stage left every canonical slot's mirror in ws.xmir
  bool pad_ready = false;  // the padded lists (psl, povo) hold the current lists
  int64_t launches = 0, sampled_launches = 0, sampled_sweeps = 0;
  // instrumentation: active slots of the current lists, totals
  // with profiling on: sweeps run on the current lists not yet counted, and a
  // device accumulator of active slots x sweeps (no host round trip per list)
  int64_t pending = 0;
  unsigned long long* acc = nullptr;
  size_t prof_used = 0;
  unsigned long long* maxs = nullptr;  // device max slots
  bool region_clear[2] = {false, false};  // max regions 0 / 1 hold zeros (cleared by s_read_max)
  int part_lanes = PV4_LANES;  // lanes per node of the list partitions (F2M_PART_LANES: 4, or 1)
  double* part = nullptr;              // reduction partials
  double* sums = nullptr;              // final sums [3]
  int32_t* ibuf = nullptr;             // misc ints
  // irregular-comb lists
  int32_t nc = 0;
  std::vector<int32_t> broots_h;

  double* F() const { return (w.cur ? w.F1 : w.F0).as<double>(); }
  double* Fo() const { return (w.cur ? w.F0 : w.F1).as<double>(); }
  int32_t* I(DevBuf& b) const { return b.as<int32_t>(); }

  f2m_status alloc() {
    F2M_HIP(c, w.nbr.reserve(m * 4));
    F2M_HIP(c, w.D.reserve(m * 8));
    F2M_HIP(c, w.orig.reserve(m * 4));
    F2M_HIP(c, w.xw.reserve(m * 8));
    F2M_HIP(c, w.xc.reserve(m * 8));
    F2M_HIP(c, w.sortk.reserve(m * 2));  // flags per working + canonical slot
    F2M_HIP(c, w.F0.reserve(n * 8));
    F2M_HIP(c, w.F1.reserve(n * 8));
    F2M_HIP(c, w.g.reserve(n * 8));
    F2M_HIP(c, w.act.reserve(n * 4));
    F2M_HIP(c, w.clen.reserve(n * 4));
    F2M_HIP(c, w.deg.reserve(n * 4));
    F2M_HIP(c, w.A.reserve(n * 4));
    F2M_HIP(c, w.T.reserve(n * 4));
    F2M_HIP(c, w.root.reserve(n * 4));
    F2M_HIP(c, w.irr.reserve((n + 1) * 4));
    F2M_HIP(c, w.frozen.reserve(n));
    F2M_HIP(c, w.sortv.reserve((int64_t)(n + 1) * 4 * 6));  // cnt, off, cur, ridx, roots, list
    F2M_HIP(c, w.axb.reserve(m * 8 + (int64_t)n * 8 + (int64_t)(n + 1) * 4));
    F2M_HIP(c, w.ef.reserve(m * 16));
    F2M_HIP(c, w.efoff.reserve((n + 1) * 4));
    F2M_HIP(c, w.psl.reserve(((int64_t)n * PP_MAX + m) * 16));
    F2M_HIP(c, w.povo.reserve((n + 1) * 4));
    F2M_HIP(c, w.red.reserve(MAXS * 8 + 3 * RED_BLOCKS * 8 + 64 * 8));
    maxs = w.red.as<unsigned long long>();
    part = reinterpret_cast<double*>(maxs + MAXS);
    sums = part + 3 * RED_BLOCKS;
    acc = reinterpret_cast<unsigned long long*>(sums + 4);
    ibuf = reinterpret_cast<int32_t*>(sums + 8);
    return F2M_OK;
  }

  f2m_status setup() {
    hipLaunchKernelGGL(s_zero_u64, dim3(4), dim3(256), 0, c->stream, maxs, (int64_t)MAXS, INF_BITS);
    region_clear[0] = region_clear[1] = false;
    hipLaunchKernelGGL(s_zero_u64, dim3(1), dim3(64), 0, c->stream, acc, (int64_t)2, 0ull);
    int nb = grid_for(m, BS);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(s_min_pos, dim3(nb), dim3(BS), 0, c->stream, gcost, m, maxs);
    F2M_LAUNCH_CHECK(c);
    unsigned long long* hp = static_cast<unsigned long long*>(c->pinned);
    F2M_HIP(c, hipMemcpyAsync(hp, maxs, 8, hipMemcpyDeviceToHost, c->stream));
    F2M_HIP(c, hipStreamSynchronize(c->stream));
    if (hp[0] == INF_BITS) return fail(c, F2M_ERR_DEGENERATE, "no positive edge cost");
    std::memcpy(&dmin, hp, 8);
    hipLaunchKernelGGL(s_init_slots, dim3(grid_for(m, BS)), dim3(BS), 0, c->stream, m, gnbr, gcost, maxs,
                       I(w.nbr), w.D.as<double>(), I(w.orig));
    w.cur = 0;
    hipLaunchKernelGGL(s_init_nodes, dim3(grid_for(n, BS)), dim3(BS), 0, c->stream, n, rp, I(w.deg),
                       w.F0.as<double>(), w.frozen.as<uint8_t>());
    F2M_LAUNCH_CHECK(c);
    eps = o.err_grad * o.err_comb_mult;
    lpn = env_int("F2M_LPN", 4);
    part_lanes = env_int("F2M_PART_LANES", PV4_LANES) == 1 ? 1 : PV4_LANES;
    const char* v = std::getenv("F2M_SWEEP");
    padded = !(v && std::strcmp(v, "compact") == 0);
    if (const char* pv = std::getenv("F2M_PAD")) {
      static const char* names[] = {"8,4", "4,4", "4,2", "8,8", "8,4,g", "4,1", "2,2", "6,2", "4,2,g", "2,1", "3,1"};
      for (int i = 0; i < 11; ++i)
        if (std::strcmp(pv, names[i]) == 0) pad_cfg = i;
    }
    static const int strides[] = {8, 4, 4, 8, 8, 4, 2, 6, 4, 2, 3};
    pp = strides[pad_cfg];
    fuse = padded ? std::max(1, std::min(TMAXS, env_int("F2M_FUSE", DEFAULT_FUSE))) : 1;
    if (fuse > 1) {
      // default tile size: the same whole number of tiles on every CU, as
      // large as that allows up to TS (n = 100k on 256 CUs: 196 nodes, 511
      // tiles, two per CU).  A launch lasts as long as its busiest CU: with
      // TS = 256 (391 tiles) half the CUs run two tiles, the others one.
      int ncu = 0;
      F2M_HIP(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
      ncu = std::max(1, ncu);
      // one 768-thread tile per CU by default: its halo is a smaller share of its local set than
      // two 384-thread tiles' (n = 100k Gaussian mixture: 24.45 -> 22.69 ms per solve)
      // (a 1024-thread shape with 2048 local nodes measured slower: 24.9-25.2 against 22.4 ms)
      auto shape = [](int v) { return v == TW_EF ? TW_EF : TW_GLOB; };
      tw_ef = shape(env_int("F2M_TILE_W", TW_GLOB));
      tw_glob = shape(env_int("F2M_TILE_W_GLOB", TW_GLOB));
      selh = env_int("F2M_TILE_SELH", 1) != 0;
      build_grad = env_int("F2M_BUILD_GRAD", 1) != 0;
      build_rows = std::min(2, std::max(0, env_int("F2M_BUILD_ROWS", 1)));
      probe_every = std::max(1, env_int("F2M_TILE_PROBE", 8));
      // tile nodes of a workgroup shape: the same whole number of tiles on every CU, as large as
      // the shape's cap allows (256 nodes for 384 threads, 512 for 768)
      auto balanced = [&](int tw_) {
        const int cap = tw_ == TW_EF ? TS : TS_W;
        const int64_t per_cu = (n + (int64_t)cap * ncu - 1) / ((int64_t)cap * ncu);
        return (int)std::max<int64_t>(64, std::min<int64_t>(cap, (n + per_cu * ncu - 1) / (per_cu * ncu)));
      };
      ts_ef = std::max(32, std::min(tw_ef == TW_EF ? TS : TS_W, env_int("F2M_TILE", balanced(tw_ef))));
      ts_glob = std::max(32, std::min(tw_glob == TW_EF ? TS : TS_W, env_int("F2M_TILE_GLOB", balanced(tw_glob))));
      const int64_t mt = (n + std::min(ts_ef, ts_glob) - 1) / std::min(ts_ef, ts_glob);  // most tiles
      tile_cap = mt;
      F2M_HIP(c, w.thdr.reserve((mt * THDR + 16 + 2 * (mt + 1)) * 4));
      F2M_HIP(c, w.tnode.reserve(mt * TL_MAX * 16));
      F2M_HIP(c, w.tslot.reserve(mt * TL_MAX * TPP_MAX * 16));
      F2M_HIP(c, w.tovf.reserve(mt * TOV_MAX * 16));
      tile_xcd = std::max(1, std::min(64, env_int("F2M_TILE_XCD", kXcds)));
      balance = std::max(0, std::min(2, env_int("F2M_TILE_BALANCE", 1)));
      for (int kind = 0; kind < 2; ++kind) {  // uniform boundaries until a build of the kind balances them
        const int ts = kind ? ts_glob : ts_ef, T = (int)((n + ts - 1) / ts);
        hipLaunchKernelGGL(s_tile_uniform, dim3((T + 256) / 256), dim3(256), 0, c->stream, T, n, ts, tile_starts(kind));
      }
      hipLaunchKernelGGL(s_set_i32, dim3(1), dim3(64), 0, c->stream, tile_fit(), INT_MAX);
      F2M_LAUNCH_CHECK(c);
    }
    return F2M_OK;
  }
  // the tiles' fit counter: INT_MAX until a build takes the min of its tiles'
  // fits into it; reset to INT_MAX where the host reads it
  int32_t* tile_fit() const { return w.thdr.as<int32_t>() + tile_cap * THDR; }
  // first ids of the tiles of list kind k (0: block_EF, 1: block_GLOB) and n
  int32_t* tile_starts(int k) const { return w.thdr.as<int32_t>() + tile_cap * THDR + 16 + k * (tile_cap + 1); }
  void set_width(int32_t fit_min) {
    width = std::max(1, std::min(fuse, fit_min + 1));
    if (fit_min >= 0) depth_next[pending_kind] = fit_min + 1;
    tiles_pending = false;
  }

  template <int LPN>
  void launch_sweep_t(bool write, bool frozen, unsigned long long* mslots) {
    dim3 grid(grid_for((int64_t)n * LPN, BS)), block(BS);
    double* fin = F();
    double* fout = Fo();
    const uint8_t* fz = w.frozen.as<uint8_t>();
    const int32_t* eo = I(w.efoff);
    const ESlot* ef = w.ef.as<ESlot>();
    if (!write)
      hipLaunchKernelGGL((s_sweep<LPN, false, false>), grid, block, 0, c->stream, n, eo, ef, fin, nullptr,
                         w.g.as<double>(), fz, mslots);
    else if (frozen)
      hipLaunchKernelGGL((s_sweep<LPN, true, true>), grid, block, 0, c->stream, n, eo, ef, fin, fout, nullptr, fz,
                         mslots);
    else
      hipLaunchKernelGGL((s_sweep<LPN, false, true>), grid, block, 0, c->stream, n, eo, ef, fin, fout, nullptr, fz,
                         mslots);
  }
  // ev0/ev1: when set, the runtime records them at the kernel's own start and
  // end (hipExtLaunchKernelGGL), so sampled timings exclude event-marker latency
  template <int PPT, int LPN, bool SPEC, int NPG = 1>
  void launch_pad_t(bool write, bool frozen, unsigned long long* mslots, hipEvent_t ev0, hipEvent_t ev1) {
    dim3 grid(grid_for(((int64_t)n + NPG - 1) / NPG * LPN, BS)), block(BS);
    const uint8_t* fz = w.frozen.as<uint8_t>();
    auto go = [&](auto kern) {
      if (ev0)
        hipExtLaunchKernelGGL(kern, grid, block, 0, c->stream, ev0, ev1, 0, n, (const int32_t*)I(w.act),
                              (const int32_t*)I(w.povo), (const ESlot*)w.psl.as<ESlot>(), (const double*)F(), Fo(),
                              fz, mslots);
      else
        hipLaunchKernelGGL(kern, grid, block, 0, c->stream, n, I(w.act), I(w.povo), w.psl.as<ESlot>(), F(), Fo(), fz,
                           mslots);
    };
    if (!write) go(s_sweep_pad<PPT, LPN, SPEC, NPG, false, false>);
    else if (frozen) go(s_sweep_pad<PPT, LPN, SPEC, NPG, true, true>);
    else go(s_sweep_pad<PPT, LPN, SPEC, NPG, false, true>);
  }
  // ev0/ev1: when set, the runtime records them at the kernel's own start and
  // end (hipExtLaunchKernelGGL), so sampled timings exclude event-marker latency
  void launch_padded(bool write, bool frozen, unsigned long long* mslots, hipEvent_t ev0, hipEvent_t ev1) {
    switch (pad_cfg) {
      case 1: launch_pad_t<4, 4, true>(write, frozen, mslots, ev0, ev1); break;
      case 2: launch_pad_t<4, 2, true>(write, frozen, mslots, ev0, ev1); break;
      case 3: launch_pad_t<8, 8, true>(write, frozen, mslots, ev0, ev1); break;
      case 4: launch_pad_t<8, 4, false>(write, frozen, mslots, ev0, ev1); break;
      case 5: launch_pad_t<4, 1, true>(write, frozen, mslots, ev0, ev1); break;
      case 6: launch_pad_t<2, 2, true>(write, frozen, mslots, ev0, ev1); break;
      case 7: launch_pad_t<6, 2, true>(write, frozen, mslots, ev0, ev1); break;
      case 8: launch_pad_t<4, 2, false>(write, frozen, mslots, ev0, ev1); break;
      case 9: launch_pad_t<2, 1, true>(write, frozen, mslots, ev0, ev1); break;
      case 10: launch_pad_t<3, 1, true>(write, frozen, mslots, ev0, ev1); break;
      default: launch_pad_t<8, 4, true>(write, frozen, mslots, ev0, ev1); break;
    }
  }
  f2m_status launch_sweep(bool write, bool frozen, unsigned long long* mslots) {
    const bool timed = write && fuse == 1 && c->prof_every > 0 && (sweeps % c->prof_every) == 0;
    if (timed) ++sampled_sweeps;
    if (timed) {
      while (c->prof_ev.size() < 2 * (prof_used + 1)) {
        hipEvent_t e;
        F2M_HIP(c, hipEventCreate(&e));
        c->prof_ev.push_back(e);
      }
      if (!padded) F2M_HIP(c, hipEventRecord(c->prof_ev[2 * prof_used], c->stream));
    }
    if (padded && !pad_ready) {
      f2m_status sp = pad_lists();
      if (sp) return sp;
    }
    if (padded) launch_padded(write, frozen, mslots, timed ? c->prof_ev[2 * prof_used] : nullptr,
                              timed ? c->prof_ev[2 * prof_used + 1] : nullptr);
    else switch (lpn) {
      case 1: launch_sweep_t<1>(write, frozen, mslots); break;
      case 2: launch_sweep_t<2>(write, frozen, mslots); break;
      case 8: launch_sweep_t<8>(write, frozen, mslots); break;
      case 16: launch_sweep_t<16>(write, frozen, mslots); break;
      default: launch_sweep_t<4>(write, frozen, mslots); break;
    }
    F2M_LAUNCH_CHECK(c);
    if (timed) {
      if (!padded) F2M_HIP(c, hipEventRecord(c->prof_ev[2 * prof_used + 1], c->stream));
      ++prof_used;
      ++sampled_launches;
      sample_slots(1);
    }
    if (write) {
      w.cur ^= 1;
      ++sweeps;
      ++launches;
      if (c->prof_every > 0) ++pending;
    }
    return F2M_OK;
  }
  // acc[1] += active slots of the current lists x sweeps of a timed launch
  void sample_slots(int S) {
    hipLaunchKernelGGL(s_acc_slots, dim3(1), dim3(1), 0, c->stream, I(w.efoff) + n, acc + 1,
                       (unsigned long long)S);
  }
  // adds (active slots of the current lists) x (sweeps run on them) on the device
  f2m_status flush_slots() {
    if (c->prof_every > 0 && pending > 0) {
      hipLaunchKernelGGL(s_acc_slots, dim3(1), dim3(1), 0, c->stream, I(w.efoff) + n, acc,
                         (unsigned long long)pending);
      F2M_LAUNCH_CHECK(c);
    }
    pending = 0;
    return F2M_OK;
  }
  // harvest sampled sweep timings and slot counts into ctx->prof
  f2m_status harvest() {
    f2m_status st = flush_slots();
    if (st) return st;
    unsigned long long slots = 0, sslots = 0;
    if (c->prof_every > 0) {
      unsigned long long* hp = static_cast<unsigned long long*>(c->pinned);
      F2M_HIP(c, hipMemcpyAsync(hp, acc, 16, hipMemcpyDeviceToHost, c->stream));
      F2M_HIP(c, hipStreamSynchronize(c->stream));
      slots = hp[0];
      sslots = hp[1];
    }
    const int64_t active_slots = (int64_t)slots;
    const double alg_bytes = kAlgSlotBytes * (double)slots + kAlgNodeBytes * (double)n * (double)sweeps;
    double ms = 0;
    for (size_t i = 0; i < prof_used; ++i) {
      float t = 0;
      F2M_HIP(c, hipEventElapsedTime(&t, c->prof_ev[2 * i], c->prof_ev[2 * i + 1]));
      ms += t;
    }
    c->prof.sweep_launches = launches;
    c->prof.sweeps = sweeps;
    c->prof.sampled_sweeps = sampled_sweeps;
    c->prof.sampled_alg_bytes = kAlgSlotBytes * (double)sslots + kAlgNodeBytes * (double)n * (double)sampled_sweeps;
    c->prof.sweep_samples = sampled_launches;
    c->prof.sweep_ms_sampled = ms;
    c->prof.sweep_alg_bytes = alg_bytes;
    c->prof.active_slots = active_slots;
    return F2M_OK;
  }
  // The padded lists in full (s_sweep_pad's layout): overflow offsets by an
  // exclusive scan of max(act - pp, 0), then the copy.
  f2m_status pad_lists() {
    F2M_HIP(c, scan_excess_i32(c, I(w.act), pp, I(w.povo), n, nullptr, 0));
    hipLaunchKernelGGL(s_pad_fill, dim3(grid_for((int64_t)n * PF_LANES, BS)), dim3(BS), 0, c->stream, n, pp, rp,
                       I(w.act), (const int32_t*)I(w.povo), I(w.nbr), w.D.as<double>(), w.psl.as<ESlot>(),
                       (int32_t*)nullptr, 0);
    F2M_LAUNCH_CHECK(c);
    pad_ready = true;
    return F2M_OK;
  }
  // update_EF (graph_cuda.cu:229-252): offsets = scan(act), copy the active
  // prefixes.  With profiling on, the sweeps run on the old lists are counted
  // first (on the device).
  // grad: the convergence check that follows takes its gradient from the
  // tile build (s_tile_build with Fg), when the lists are tiled
  f2m_status ef_rebuild(bool long_lists = false, bool grad = false) {
    f2m_status st0 = flush_slots();
    if (st0) return st0;
    // offsets of the compacted lists: the padded layout needs them only for
    // the instrumentation's slot counts (efoff[n])
    if (!padded || c->prof_every > 0) F2M_HIP(c, scan_i32_to_i32(c, I(w.act), I(w.efoff), n, nullptr));
    const dim3 gn(grid_for(n, BS)), bs(BS);
    if (padded) {
      const int kind = long_lists ? 1 : 0;
      // (a probe level that does not fit in every tile costs a level pass, a
      // larger sort and more records on every build: n = 100k Gaussian
      // mixture, block_GLOB lists fit four levels and the fifth is probed)
      const bool probe = builds[kind]++ % probe_every == 0;
      const int depth = depth_next[kind] > 0 ? std::min(fuse - 1, depth_next[kind] - (probe ? 0 : 1)) : fuse - 1;
      const bool tiled = fuse > 1;
      if (tiled) ntiles = (n + (long_lists ? ts_glob : ts_ef) - 1) / (long_lists ? ts_glob : ts_ef);
      // the tile build reads the lists' rows; the padded copy of the first pp
      // slots only for the block_EF lists (one dependent read fewer per
      // level), the full padded lists (overflow offsets by a scan) only when a
      // padded sweep runs on them (pad_lists).  (n = 100k Gaussian mixture:
      // 19.30 -> 18.82 ms against building the full padded lists with every
      // rebuild.)
      const bool copy = tiled && build_rows < 2 && !long_lists;
      pad_ready = false;
      if (!tiled || build_rows == 0) {
        f2m_status sp = pad_lists();
        if (sp) return sp;
      } else if (copy) {
        hipLaunchKernelGGL(s_pad_fill, dim3(grid_for((int64_t)n * PF_LANES, BS)), bs, 0, c->stream, n, pp, rp,
                           I(w.act), (const int32_t*)nullptr, I(w.nbr), w.D.as<double>(), w.psl.as<ESlot>(),
                           (int32_t*)nullptr, 0);
      }
      if (fuse > 1) {
        tpp = long_lists ? 6 : 4;
        const int ts = long_lists ? ts_glob : ts_ef;
        tw = long_lists ? tw_glob : tw_ef;
        pending_kind = kind;
        const bool bg = grad && build_grad;
        if (bg) {
          f2m_status sz = zero_region(0);
          if (sz) return sz;
        }
        auto build = [&](auto kern, int threads) {
          hipLaunchKernelGGL(kern, dim3(ntiles), dim3(threads), 0, c->stream, n, (const int32_t*)tile_starts(kind), pp,
                             tpp, depth,
                             (const int32_t*)I(w.act), rp, (const double*)w.D.as<double>(), (const int32_t*)I(w.nbr),
                             (pad_ready || copy) ? (const ESlot*)w.psl.as<ESlot>() : (const ESlot*)nullptr,
                             w.thdr.as<int32_t>(), w.tnode.as<int4>(), w.tslot.as<TSlot>(), w.tovf.as<TSlot>(),
                             tile_fit(), bg ? (const double*)F() : (const double*)nullptr, bg ? maxs : nullptr);
        };
        grad_ready = bg;
        if (tw == TW_GLOB) build(s_tile_build<TW_GLOB, 1024>, 1024);
        else build(s_tile_build<TW_EF>, TCfg<TW_EF>::TL);
        if (balance > kind && ntiles <= BAL_MAX) {  // boundaries for the next build of this kind
          const int cap = tw == TW_EF ? TS : TS_W;
          hipLaunchKernelGGL(s_tile_balance, dim3(1), dim3(BAL_T), 0, c->stream, ntiles, n, ts, cap, fuse,
                             tpp == 4 ? 256.0 : 2.0, (const int32_t*)w.thdr.as<int32_t>(),
                             (const int32_t*)tile_fit(), tile_starts(kind));
        }
        tiles_pending = true;
        width = 1;
      }
    } else {
      hipLaunchKernelGGL(s_ef_fill, gn, bs, 0, c->stream, n, rp, I(w.act), I(w.efoff), I(w.nbr), w.D.as<double>(),
                         w.ef.as<ESlot>());
    }
    F2M_LAUNCH_CHECK(c);
    return F2M_OK;
  }
  // max over the 32 strided slots of region r (0: gradient, 1: early exit)
  // (one launch that also reads the pending tiles' fit and clears the region)
  // The host waits by reading the kernel's sequence number in pinned memory
  // (a few us sooner than a stream synchronisation); after 50 ms it falls
  // back to the synchronisation, which also reports any error.
  f2m_status read_max(int r, double& out) {
    unsigned long long* rb = static_cast<unsigned long long*>(c->pinned) + 2048;
    const bool tp = tiles_pending;
    const unsigned long long want = ++c->read_seq;
    hipLaunchKernelGGL(s_read_max, dim3(1), dim3(64), 0, c->stream, maxs + 512 * r, tp ? tile_fit() : nullptr, rb,
                       want);
    F2M_LAUNCH_CHECK(c);
    bool seen = false;
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned it = 1;; ++it) {
      if (__atomic_load_n(&rb[2], __ATOMIC_ACQUIRE) == want) {
        seen = true;
        break;
      }
      if ((it & 1023u) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) break;
    }
    if (!seen) F2M_HIP(c, hipStreamSynchronize(c->stream));
    region_clear[r] = true;
    if (tp) set_width((int32_t)(uint32_t)__atomic_load_n(&rb[1], __ATOMIC_ACQUIRE));
    const unsigned long long bits = __atomic_load_n(&rb[0], __ATOMIC_ACQUIRE);
    std::memcpy(&out, &bits, 8);
    return F2M_OK;
  }
  // zeros region r before kernels take their max into it (skipped when the
  // last read left it cleared)
  f2m_status zero_region(int r) {
    if (!region_clear[r]) {
      hipLaunchKernelGGL(s_zero_u64, dim3(2), dim3(256), 0, c->stream, maxs + 512 * r, (int64_t)512, 0ull);
      F2M_LAUNCH_CHECK(c);
    }
    region_clear[r] = false;
    return F2M_OK;
  }
  // gradient over the active lists + max |g| (host round trip)
  f2m_status grad_max(double& M) {
    f2m_status s;
    if (grad_ready) {  // the last tile build took the gradient into region 0
      grad_ready = false;
      return read_max(0, M);
    }
    if ((s = zero_region(0))) return s;
    if ((s = launch_sweep(false, false, maxs))) return s;
    return read_max(0, M);
  }
  // sweeps per launch on the current lists (reads the tiles' fit when the
  // read-back of a max has not already done so)
  f2m_status tile_width(int& out) {
    if (tiles_pending) {
      int32_t* hfit = reinterpret_cast<int32_t*>(static_cast<unsigned long long*>(c->pinned) + 1024);
      F2M_HIP(c, hipMemcpyAsync(hfit, tile_fit(), 4, hipMemcpyDeviceToHost, c->stream));
      F2M_HIP(c, hipStreamSynchronize(c->stream));
      set_width(*hfit);
      hipLaunchKernelGGL(s_set_i32, dim3(1), dim3(64), 0, c->stream, tile_fit(), INT_MAX);
      F2M_LAUNCH_CHECK(c);
    }
    out = width;
    return F2M_OK;
  }
  template <int S>
  void launch_tile_t(bool frozen, unsigned long long* mslots) {
    const int TW = tw;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(ntiles), dim3(TW), 0, c->stream, (const int32_t*)w.thdr.as<int32_t>(),
                         (const int4*)w.tnode.as<int4>(), (const TSlot*)w.tslot.as<TSlot>(),
                         (const TSlot*)w.tovf.as<TSlot>(), (const double*)F(), Fo(),
                         (const uint8_t*)w.frozen.as<uint8_t>(), mslots, tile_xcd);
    };
    if (tpp == 6 && tw == TW_GLOB && selh) {
      if (frozen) go(s_sweep_tile<S, true, 6, TW_GLOB, true>);
      else go(s_sweep_tile<S, false, 6, TW_GLOB, true>);
    } else if (tpp == 6 && tw == TW_GLOB) {
      if (frozen) go(s_sweep_tile<S, true, 6, TW_GLOB>);
      else go(s_sweep_tile<S, false, 6, TW_GLOB>);
    } else if (tw == TW_GLOB) {
      if (frozen) go(s_sweep_tile<S, true, 4, TW_GLOB>);
      else go(s_sweep_tile<S, false, 4, TW_GLOB>);
    } else if (tpp == 6) {
      if (frozen) go(s_sweep_tile<S, true, 6, TW_EF>);
      else go(s_sweep_tile<S, false, 6, TW_EF>);
    } else {
      if (frozen) go(s_sweep_tile<S, true, 4, TW_EF>);
      else go(s_sweep_tile<S, false, 4, TW_EF>);
    }
  }
  // S sweeps in one launch over the tiles
  f2m_status launch_tile(int S, bool frozen, unsigned long long* mslots) {
    switch (S) {
      case 1: launch_tile_t<1>(frozen, mslots); break;
      case 2: launch_tile_t<2>(frozen, mslots); break;
      case 3: launch_tile_t<3>(frozen, mslots); break;
      case 4: launch_tile_t<4>(frozen, mslots); break;
      case 5: launch_tile_t<5>(frozen, mslots); break;
      case 6: launch_tile_t<6>(frozen, mslots); break;
      case 7: launch_tile_t<7>(frozen, mslots); break;
      case 8: launch_tile_t<8>(frozen, mslots); break;
      case 9: launch_tile_t<9>(frozen, mslots); break;
      case 10: launch_tile_t<10>(frozen, mslots); break;
      case 11: launch_tile_t<11>(frozen, mslots); break;
      default: launch_tile_t<12>(frozen, mslots); break;
    }
    F2M_LAUNCH_CHECK(c);
    ++launches;
    w.cur ^= 1;
    sweeps += S;
    if (c->prof_every > 0) pending += S;
    return F2M_OK;
  }
  // `count` sweeps; mlast (when set) receives max |g| of the last one
  f2m_status run_sweeps(int count, bool frozen, unsigned long long* mlast) {
    f2m_status s;
    int wd = 1;
    if (fuse > 1 && count > 1)
      if ((s = tile_width(wd))) return s;
    if (wd > 1) {
      // instrumented solves: one event pair around every run of two or more
      // back-to-back launches of the configured width on the block_EF lists
      // (the dominant kernel), the gaps between them included; a shorter
      // last launch stays outside
      const int full = count / wd;
      const bool win = c->prof_every > 0 && wd == fuse && tpp == 4 && full >= 2;
      if (win) {
        while (c->prof_ev.size() < 2 * (prof_used + 1)) {
          hipEvent_t e;
          F2M_HIP(c, hipEventCreate(&e));
          c->prof_ev.push_back(e);
        }
        F2M_HIP(c, hipEventRecord(c->prof_ev[2 * prof_used], c->stream));
      }
      for (int done = 0, i = 0; done < count; ++i) {
        const int S = std::min(wd, count - done);
        done += S;
        if ((s = launch_tile(S, frozen, done == count ? mlast : nullptr))) return s;
        if (win && i + 1 == full) {
          F2M_HIP(c, hipEventRecord(c->prof_ev[2 * prof_used + 1], c->stream));
          ++prof_used;
          sampled_launches += full;
          sampled_sweeps += (int64_t)full * wd;
          sample_slots(full * wd);
        }
      }
      return F2M_OK;
    }
    for (int i = 0; i < count; ++i)
      if ((s = launch_sweep(true, frozen, i + 1 == count ? mlast : nullptr))) return s;
    return F2M_OK;
  }
  f2m_status sweep_n(int count, bool frozen) { return run_sweeps(count, frozen, nullptr); }

  // block_GLOB (graph_cuda.cu:363-413)
  f2m_status block_glob(int32_t& rounds) {
    hipLaunchKernelGGL(s_act_min, dim3(grid_for(n, BS)), dim3(BS), 0, c->stream, n, I(w.deg), I(w.act), o.glob_k);
    F2M_LAUNCH_CHECK(c);
    int go = 0;
    f2m_status s;
    const bool long_lists = o.glob_k > 4;  // every list holds min(deg, glob_k) or more slots here
    if ((s = ef_rebuild(long_lists))) return s;
    for (; go < o.glob_rounds; ++go) {
      if ((s = sweep_n(o.glob_sweeps, false))) return s;
      if (part_lanes == PV4_LANES)
        hipLaunchKernelGGL(s_part_v4_q, dim3(grid_for((int64_t)n * PV4_LANES, BS)), dim3(BS), 0, c->stream, n, rp,
                           I(w.deg), I(w.act), o.glob_k, I(w.nbr), w.D.as<double>(), I(w.orig), F());
      else
        hipLaunchKernelGGL(s_part_v4, dim3(grid_for(n, BS)), dim3(BS), 0, c->stream, n, rp, I(w.deg), I(w.act),
                           o.glob_k, I(w.nbr), w.D.as<double>(), I(w.orig), F());
      F2M_LAUNCH_CHECK(c);
      if ((s = ef_rebuild(long_lists, true))) return s;
      double M;
      if ((s = grad_max(M))) return s;
      if (M < o.glob_exit) { ++go; break; }
    }
    rounds = go;
    return F2M_OK;
  }
  // check_OPT + update_EF (graph_cuda.cu:254-271, 229-252)
  f2m_status check_opt(double& M) {
    if (part_lanes == PV4_LANES)
      hipLaunchKernelGGL(s_part_v2_q, dim3(grid_for((int64_t)n * PV4_LANES, BS)), dim3(BS), 0, c->stream, n, rp,
                         I(w.deg), I(w.act), 1.0, I(w.nbr), w.D.as<double>(), I(w.orig), F());
    else
      hipLaunchKernelGGL(s_part_v2, dim3(grid_for(n, BS)), dim3(BS), 0, c->stream, n, rp, I(w.deg), I(w.act), 1.0,
                         I(w.nbr), w.D.as<double>(), I(w.orig), F());
    F2M_LAUNCH_CHECK(c);
    f2m_status s = ef_rebuild(false, true);
    if (s) return s;
    return grad_max(M);
  }
  // block_EF (graph_cuda.cu:308-338).  Its opening max |g| is MA: every
  // caller runs it right after check_opt, whose last step computed that same
  // max on the same lists and potentials.
  f2m_status block_ef(int N, bool frozen, double MA) {
    f2m_status s;
    if (!o.early_exit) {
      int cnt = (MA < o.err_grad) ? std::min(N, o.check_every) : N;
      return sweep_n(cnt, frozen);
    }
    int done = 0;
    while (done < N) {
      int chunk = std::min(o.check_every, N - done);
      if (chunk == o.check_every) {  // max |g| of the chunk's last sweep
        if ((s = zero_region(1))) return s;
        if ((s = run_sweeps(chunk, frozen, maxs + 512))) return s;
        double MB;
        if ((s = read_max(1, MB))) return s;
        done += chunk;
        if (MB < o.err_grad) break;
      } else {
        if ((s = run_sweeps(chunk, frozen, nullptr))) return s;
        done += chunk;
      }
    }
    return F2M_OK;
  }
  // run_start_cuda (graph_cuda.cu:163-199)
  f2m_status run_start(double& M, int32_t& rounds) {
    f2m_status s;
    if ((s = check_opt(M))) return s;
    int factor = o.factor0, go = 1;
    for (; go <= o.max_outer; ++go) {
      if ((s = block_ef(factor, false, M))) return s;
      if ((s = check_opt(M))) return s;
      if (factor < o.factor_max) factor += o.factor_step;
      if (M < o.err_grad) break;
    }
    rounds = go;
    return F2M_OK;
  }
  // run_fix_cuda (graph_cuda.cu:204-226)
  f2m_status run_fix(double& M) {
    f2m_status s;
    if ((s = check_opt(M))) return s;
    for (int go = 1; go <= o.fix_rounds; ++go) {
      if ((s = block_ef(o.fix_factor, false, M))) return s;
      if ((s = check_opt(M))) return s;
      if (M < o.err_grad) break;
    }
    return F2M_OK;
  }

  int32_t* cnt() const { return w.sortv.as<int32_t>(); }
  int32_t* off() const { return cnt() + (n + 1); }
  int32_t* cur() const { return off() + (n + 1); }
  int32_t* ridx() const { return cur() + (n + 1); }
  int32_t* roots() const { return ridx() + (n + 1); }
  int32_t* list() const { return roots() + (n + 1); }
  double* xg() const { return w.axb.as<double>(); }
  double* rn() const { return xg() + m; }
  int32_t* status() const { return reinterpret_cast<int32_t*>(rn() + n); }
  uint8_t* fw() const { return w.sortk.as<uint8_t>(); }
  uint8_t* fc() const { return fw() + m; }

  // get_combs + host_combs_check (graph_cuda_comb.hpp:2-28, 296-356):
  // returns the number of IRREGULAR_B combs; fills n_combs / n_irregular_a.
  f2m_status combs_check(int32_t& ncombs, int32_t& irr_a, int32_t& irr_b) {
    const dim3 gn(grid_for(n, BS)), bs(BS);
    double* f = F();
    hipLaunchKernelGGL(s_classify, gn, bs, 0, c->stream, n, rp, I(w.act), I(w.clen), I(w.A), I(w.T), eps,
                       I(w.nbr), w.D.as<double>(), I(w.orig), f);
    int32_t* par = I(w.irr);  // reuse as parent array, then as irregular flags
    hipLaunchKernelGGL(s_uf_init, gn, bs, 0, c->stream, n, par);
    hipLaunchKernelGGL(s_uf_union, gn, bs, 0, c->stream, n, rp, I(w.clen), eps, I(w.nbr), w.D.as<double>(), f,
                       par);
    hipLaunchKernelGGL(s_uf_root, gn, bs, 0, c->stream, n, I(w.T), par, I(w.root));
    hipLaunchKernelGGL(s_zero_i32, gn, bs, 0, c->stream, par, (int64_t)n);
    hipLaunchKernelGGL(s_irr_mark, gn, bs, 0, c->stream, n, I(w.root), I(w.A), I(w.T), par);
    hipLaunchKernelGGL(s_primal, gn, bs, 0, c->stream, n, rp, I(w.deg), I(w.clen), eps, I(w.nbr),
                       w.D.as<double>(), f, I(w.root), par, w.xw.as<double>(), fw());
    F2M_LAUNCH_CHECK(c);
    // number of combs: roots are nodes with root[v] == v; count via flags
    hipLaunchKernelGGL(s_zero_i32, gn, bs, 0, c->stream, cnt(), (int64_t)n);
    hipLaunchKernelGGL(s_irr_count, gn, bs, 0, c->stream, n, I(w.root), par, cnt());
    hipLaunchKernelGGL(s_flag_pos, gn, bs, 0, c->stream, n, cnt(), cur());
    F2M_LAUNCH_CHECK(c);
    int32_t nirr = 0;
    F2M_HIP(c, scan_i32_to_i32(c, cur(), ridx(), n, &nirr));
    nc = nirr;
    irr_a = nirr;
    irr_b = 0;
    broots_h.clear();
    // with exact recovery on (opts.exact >= 1), irregular combs go to the
    // exact stage: their least squares (and the dual repairs it would start)
    // are not run
    if (nirr > 0 && o.exact == 0) {
      int32_t tot = 0;
      F2M_HIP(c, scan_i32_to_i32(c, cnt(), off(), n, &tot));
      hipLaunchKernelGGL(s_root_list, gn, bs, 0, c->stream, n, cnt(), ridx(), roots());
      hipLaunchKernelGGL(s_zero_i32, gn, bs, 0, c->stream, cur(), (int64_t)n);
      hipLaunchKernelGGL(s_irr_fill, gn, bs, 0, c->stream, n, I(w.root), par, off(), cur(), list());
      hipLaunchKernelGGL(s_seg_sort, dim3(nirr), bs, 0, c->stream, roots(), off(), cnt(), list());
      hipLaunchKernelGGL(s_axb, dim3(nirr), bs, 0, c->stream, roots(), off(), cnt(), list(), rp, I(w.clen), I(w.A),
                         I(w.T), eps, I(w.nbr), w.D.as<double>(), f, w.xw.as<double>(), xg(), rn(), status());
      F2M_LAUNCH_CHECK(c);
      std::vector<int32_t> st(nirr), rts(nirr);
      F2M_HIP(c, hipMemcpyAsync(st.data(), status(), nirr * 4, hipMemcpyDeviceToHost, c->stream));
      F2M_HIP(c, hipMemcpyAsync(rts.data(), roots(), nirr * 4, hipMemcpyDeviceToHost, c->stream));
      F2M_HIP(c, hipStreamSynchronize(c->stream));
      for (int i = 0; i < nirr; ++i)
        if (st[i] == 0) broots_h.push_back(rts[i]);
      irr_b = (int32_t)broots_h.size();
    }
    // total combs = distinct roots: count nodes with root[v] == v
    ncombs = 0;
    {
      hipLaunchKernelGGL(s_zero_i32, dim3(1), dim3(64), 0, c->stream, ibuf, (int64_t)4);
      F2M_LAUNCH_CHECK(c);
      count_roots(ncombs);
    }
    return F2M_OK;
  }
  f2m_status count_roots(int32_t& out);

  f2m_status repair(int32_t& irr_b, int32_t& n_repairs, double& M, int32_t& ncombs, int32_t& irr_a) {
    double shiftV = o.shift0;
    f2m_status s;
    DevBuf& bro = c->stage[5];
    auto attempt = [&](double sv, int type) -> f2m_status {
      ++n_repairs;
      hipLaunchKernelGGL(s_zero_u8, dim3(grid_for(n, BS)), dim3(BS), 0, c->stream, w.frozen.as<uint8_t>(),
                         (int64_t)n);
      F2M_HIP(c, bro.reserve(broots_h.size() * 4 + 4));
      F2M_HIP(c, hipMemcpyAsync(bro.p, broots_h.data(), broots_h.size() * 4, hipMemcpyHostToDevice, c->stream));
      hipLaunchKernelGGL(s_shift, dim3(grid_for((int64_t)broots_h.size(), 64)), dim3(64), 0, c->stream,
                         (int)broots_h.size(), bro.as<int32_t>(), off(), cnt(), list(), I(w.A), I(w.T), sv, type,
                         F(), w.frozen.as<uint8_t>());
      F2M_LAUNCH_CHECK(c);
      f2m_status s2;
      if ((s2 = sweep_n(o.frozen_sweeps, true))) return s2;
      hipLaunchKernelGGL(s_zero_u8, dim3(grid_for(n, BS)), dim3(BS), 0, c->stream, w.frozen.as<uint8_t>(),
                         (int64_t)n);
      if ((s2 = run_fix(M))) return s2;
      return combs_check(ncombs, irr_a, irr_b);
    };
    for (;;) {
      if ((s = attempt(shiftV, 0))) return s;
      if (irr_b == 0) break;
      if ((s = attempt(shiftV, 1))) return s;
      if (irr_b == 0) break;
      shiftV = -shiftV;
      if ((s = attempt(shiftV, 0))) return s;
      if (irr_b == 0) break;
      shiftV = -shiftV;
      if ((s = attempt(shiftV, 1))) return s;
      if (irr_b == 0) break;
      shiftV = -shiftV;
      shiftV *= 2;
      if (std::fabs(shiftV) > o.shift_max) break;
    }
    return F2M_OK;
  }

  // x and X-flags per canonical slot
  f2m_status to_canon() {
    hipLaunchKernelGGL(s_to_canon, dim3(grid_for(m, BS)), dim3(BS), 0, c->stream, m, I(w.orig), w.xw.as<double>(),
                       fw(), w.xc.as<double>(), fc());
    F2M_LAUNCH_CHECK(c);
    return F2M_OK;
  }
  // exact recovery where the reference recovery is not exact (opts.exact 1:
  // status 1 or an irregular component, whose x came from the float least
  // squares), or always (2)
  f2m_status exact(f2m_stats& z) {
    const bool run = o.exact >= 2 || (o.exact == 1 && (z.status != 0 || z.n_irregular_a > 0));
    if (!run) return F2M_OK;
    f2m_stats t = z;
    int rc = 0;
    f2m_status s = exact_stage(c, w, n, m, rp, gnbr, gcost, dmin, o, F(), w.xc.as<double>(), fc(), I(w.root), t, rc);
    if (s) return s;
    mir_ok = rc >= 0;  // every slot's mirror found
    z.exact_ran = t.exact_ran;
    if (rc >= 0) {
      z = t;
      z.status = rc == 0 ? 0 : 1;
    } else {
      z.status = 1;
    }
    return F2M_OK;
  }
  f2m_status finish(f2m_stats& z) {
    hipLaunchKernelGGL(s_sums, dim3(RED_BLOCKS), dim3(BS), 0, c->stream, n, rp, gnbr, gcost, w.xc.as<double>(),
                       I(w.nbr), w.D.as<double>(), F(), part);
    hipLaunchKernelGGL(s_sums_final, dim3(1), dim3(BS), 0, c->stream, part, sums);
    hipLaunchKernelGGL(s_zero_i32, dim3(1), dim3(64), 0, c->stream, ibuf, (int64_t)4);
    hipLaunchKernelGGL(s_maxrank, dim3(grid_for(n, BS)), dim3(BS), 0, c->stream, n, rp, gnbr, fc(),
                       mir_ok ? (const int32_t*)w.xmir.as<int32_t>() : (const int32_t*)nullptr, ibuf);
    F2M_LAUNCH_CHECK(c);
    double hs[3];
    int32_t mr = 0;
    F2M_HIP(c, hipMemcpyAsync(hs, sums, sizeof hs, hipMemcpyDeviceToHost, c->stream));
    F2M_HIP(c, hipMemcpyAsync(&mr, ibuf, 4, hipMemcpyDeviceToHost, c->stream));
    F2M_HIP(c, hipStreamSynchronize(c->stream));
    z.objective = hs[0];
    z.dual = dmin * (-hs[1] + hs[2]);
    z.max_rank = mr;
    return F2M_OK;
  }
};

// one atomic per workgroup (one per wave queued ~1,500 atomics on one
// address at n = 100k)
__global__ void __launch_bounds__(1024) s_count_roots(int n, const int32_t* __restrict__ root, int32_t* out) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long m = __ballot(v < n && root[v] == v);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&cnt, __popcll(m));
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(out, cnt);
}

f2m_status Runner::count_roots(int32_t& out) {
  hipLaunchKernelGGL(s_count_roots, dim3(grid_for(n, 1024)), dim3(1024), 0, c->stream, n, I(w.root), ibuf);
  F2M_LAUNCH_CHECK(c);
  F2M_HIP(c, hipMemcpyAsync(&out, ibuf, 4, hipMemcpyDeviceToHost, c->stream));
  F2M_HIP(c, hipStreamSynchronize(c->stream));
  return F2M_OK;
}

}  // namespace

f2m_status solve_graph(f2m_ctx* c, SolveWs& w, const f2m_graph* g, const f2m_opts& o, f2m_stats* st,
                       const double* F_init) {
  Runner R{c, w, o, g->n, g->nslots, g->rowptr, g->nbr, g->cost};
  f2m_status s;
  w.n = g->n;
  w.m = g->nslots;
  if ((s = R.alloc())) return s;
  if ((s = R.setup())) return s;
  c->last_dmin = R.dmin;
  c->last_k = g->k;
  f2m_stats z;
  std::memset(&z, 0, sizeof z);
  z.d_min = R.dmin;
  if (F_init) {  // warm start: the same schedule from the given potentials
    F2M_HIP(c, hipMemcpyAsync(w.F0.p, F_init, (size_t)g->n * 8, hipMemcpyDeviceToDevice, c->stream));
    w.cur = 0;
  }
  if ((s = R.block_glob(z.glob_rounds))) return s;
  double M = 0;
  if (o.stop_phase != 1)
    if ((s = R.run_start(M, z.outer_rounds))) return s;
  z.max_grad = M;
  if (o.stop_phase == 1 || o.stop_phase == 2) {
    F2M_HIP(c, hipStreamSynchronize(c->stream));
    if ((s = R.harvest())) return s;
    z.sweeps = R.sweeps;
    c->last_n = g->n;
    c->last_m = g->nslots;
    if (st) *st = z;
    return F2M_OK;
  }
  int32_t irr_b = 0;
  if ((s = R.combs_check(z.n_combs, z.n_irregular_a, irr_b))) return s;
  z.n_irregular_b = irr_b;
  if (irr_b > 0 && o.repair) {
    if ((s = R.repair(irr_b, z.n_repairs, M, z.n_combs, z.n_irregular_a))) return s;
    z.max_grad = M;
  }
  z.status = irr_b == 0 ? 0 : 1;
  if ((s = R.to_canon())) return s;
  if ((s = R.exact(z))) return s;
  if ((s = R.finish(z))) return s;
  if ((s = R.harvest())) return s;
  z.sweeps = R.sweeps;
  c->last_n = g->n;
  c->last_m = g->nslots;
  if (st) *st = z;
  return F2M_OK;
}

}  // namespace f2m

using namespace f2m;

// Test entry: nsweeps Jacobi sweeps on caller arrays, through the same
// compaction (s_ef_fill) and sweep kernel as f2m_solve.
namespace {
template <int LPN>
void launch_plain(hipStream_t s, int n, const int32_t* efoff, const ESlot* ef, const double* fin, double* fout,
                  double* g, unsigned long long* mx) {
  hipLaunchKernelGGL((s_sweep<LPN, false, true>), dim3(grid_for((int64_t)n * LPN, BS)), dim3(BS), 0, s, n, efoff,
                     ef, fin, fout, g, nullptr, mx);
}
}  // namespace

extern "C" {

f2m_status f2m_solve(f2m_ctx* c, const f2m_graph* g, const f2m_opts* o, f2m_stats* st) {
  if (!c || !g || g->n < 2 || g->nslots < 1) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  f2m_opts oo;
  if (o) oo = *o; else f2m_default_opts(&oo);
  auto t0 = std::chrono::steady_clock::now();
  f2m_status s = solve_graph(c, c->ws, g, oo, st, nullptr);
  c->prof.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return s;
}

f2m_status f2m_solve_warm(f2m_ctx* c, const f2m_graph* g, const f2m_opts* o, const double* F_init, f2m_stats* st) {
  if (!c || !g || g->n < 2 || g->nslots < 1 || !F_init) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  f2m_opts oo;
  if (o) oo = *o; else f2m_default_opts(&oo);
  const double* fi = (const double*)stage_in(c, F_init, (size_t)g->n * 8, 3);
  if (!fi) return F2M_ERR_HIP;
  auto t0 = std::chrono::steady_clock::now();
  f2m_status s = solve_graph(c, c->ws, g, oo, st, fi);
  c->prof.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return s;
}

f2m_status f2m_knn_certify(f2m_ctx* c, const f2m_graph* g, int64_t* violations) {
  if (!c || !g || !violations) return F2M_ERR_ARG;
  if (g->k < 1) return fail(c, F2M_ERR_ARG, "the graph was not built by f2m_knn_build (k unknown)");
  if (c->last_n != g->n || c->last_k != g->k) return fail(c, F2M_ERR_ARG, "solve this graph first");
  F2M_HIP(c, hipSetDevice(c->device));
  SolveWs& w = c->ws;
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(w.red.as<double>());
  hipLaunchKernelGGL(s_zero_u64, dim3(1), dim3(64), 0, c->stream, cnt, (int64_t)1, 0ull);
  hipLaunchKernelGGL(s_knn_certify, dim3(grid_for(g->n, BS)), dim3(BS), 0, c->stream, g->n, g->rowptr, g->cost,
                     g->k, c->last_dmin, (w.cur ? w.F1 : w.F0).as<double>(), cnt);
  F2M_LAUNCH_CHECK(c);
  unsigned long long h = 0;
  F2M_HIP(c, copy_out(c, &h, cnt, 8));
  *violations = (int64_t)h;
  return F2M_OK;
}

f2m_status f2m_ctx_set_profile(f2m_ctx* c, int32_t sample_every) {
  if (!c || sample_every < 0) return F2M_ERR_ARG;
  c->prof_every = sample_every;
  return F2M_OK;
}

f2m_status f2m_get_profile(f2m_ctx* c, f2m_profile* out) {
  if (!c || !out) return F2M_ERR_ARG;
  *out = c->prof;
  return F2M_OK;
}

f2m_status f2m_result_copy(f2m_ctx* c, double* F, double* x, int32_t* comb_root) {
  if (!c || c->last_n <= 0) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  SolveWs& w = c->ws;
  F2M_HIP(c, copy_out(c, F, (w.cur ? w.F1 : w.F0).p, c->last_n * 8));
  F2M_HIP(c, copy_out(c, x, w.xc.p, c->last_m * 8));
  F2M_HIP(c, copy_out(c, comb_root, w.root.p, c->last_n * 4));
  return F2M_OK;
}

f2m_status f2m_sweep(f2m_ctx* c, int32_t n, const int32_t* rowptr, const int32_t* nbr, const double* D,
                     const int32_t* len, const double* F_in, double* F_out, double* g_out, int32_t nsweeps,
                     double* max_abs_g) {
  if (!c || n < 1 || !rowptr || !nbr || !D || !len || !F_in || !F_out || nsweeps < 1) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  int32_t m = 0;
  F2M_HIP(c, hipMemcpy(&m, rowptr + n, 4, hipMemcpyDefault));
  const int32_t* drp = (const int32_t*)stage_in(c, rowptr, (n + 1) * 4ull, 0);
  const int32_t* dnb = (const int32_t*)stage_in(c, nbr, m * 4ull, 1);
  const double* dD = (const double*)stage_in(c, D, m * 8ull, 2);
  const int32_t* dl = (const int32_t*)stage_in(c, len, n * 4ull, 3);
  const double* dF = (const double*)stage_in(c, F_in, n * 8ull, 4);
  if (!drp || !dnb || !dD || !dl || !dF) return F2M_ERR_HIP;
  SolveWs& w = c->ws;
  F2M_HIP(c, w.F0.reserve(n * 8ull));
  F2M_HIP(c, w.F1.reserve(n * 8ull));
  F2M_HIP(c, w.g.reserve(n * 8ull));
  F2M_HIP(c, w.red.reserve(4096));
  F2M_HIP(c, hipMemcpyAsync(w.F0.p, dF, n * 8ull, hipMemcpyDeviceToDevice, c->stream));
  unsigned long long* mx = w.red.as<unsigned long long>();
  F2M_HIP(c, w.efoff.reserve((n + 1) * 4ull));
  F2M_HIP(c, w.ef.reserve(m * 16ull + 16));
  F2M_HIP(c, scan_i32_to_i32(c, dl, w.efoff.as<int32_t>(), n, nullptr));
  hipLaunchKernelGGL(s_ef_fill, dim3(grid_for(n, BS)), dim3(BS), 0, c->stream, n, drp, dl, w.efoff.as<int32_t>(), dnb,
                     dD, w.ef.as<ESlot>());
  F2M_LAUNCH_CHECK(c);
  const int32_t* eo = w.efoff.as<int32_t>();
  const ESlot* ef = w.ef.as<ESlot>();
  int lpn = env_int("F2M_LPN", 4);
  int cur = 0;
  for (int i = 0; i < nsweeps; ++i) {
    double* fin = cur ? w.F1.as<double>() : w.F0.as<double>();
    double* fout = cur ? w.F0.as<double>() : w.F1.as<double>();
    unsigned long long* ms = (i == nsweeps - 1) ? mx : nullptr;
    if (ms) hipLaunchKernelGGL(s_zero_u64, dim3(1), dim3(512), 0, c->stream, mx, (int64_t)512, 0ull);
    switch (lpn) {
      case 1: launch_plain<1>(c->stream, n, eo, ef, fin, fout, w.g.as<double>(), ms); break;
      case 2: launch_plain<2>(c->stream, n, eo, ef, fin, fout, w.g.as<double>(), ms); break;
      case 8: launch_plain<8>(c->stream, n, eo, ef, fin, fout, w.g.as<double>(), ms); break;
      case 16: launch_plain<16>(c->stream, n, eo, ef, fin, fout, w.g.as<double>(), ms); break;
      case 64: launch_plain<64>(c->stream, n, eo, ef, fin, fout, w.g.as<double>(), ms); break;
      default: launch_plain<4>(c->stream, n, eo, ef, fin, fout, w.g.as<double>(), ms); break;
    }
    F2M_LAUNCH_CHECK(c);
    cur ^= 1;
  }
  F2M_HIP(c, copy_out(c, F_out, cur ? w.F1.p : w.F0.p, n * 8ull));
  if (g_out) F2M_HIP(c, copy_out(c, g_out, w.g.p, n * 8ull));
  if (max_abs_g) {
    unsigned long long h[512];
    F2M_HIP(c, copy_out(c, h, mx, sizeof h));
    unsigned long long b = 0;
    for (int k = 0; k < 32; ++k) b = h[16 * k] > b ? h[16 * k] : b;
    std::memcpy(max_abs_g, &b, 8);
  }
  return F2M_OK;
}

}  // extern "C"
