// f2m_batch.hip -- batched F2M: thousands of independent instances in one
// persistent launch (new; the reference solves one instance per process).
//
// Each workgroup pulls instances from a queue ordered by decreasing size
// (longest-processing-time first) and runs the complete plain schedule of
// graph::RUN for that instance -- warm start, active-set descent, tight
// components, least squares, dual repair, objective -- with workgroup
// barriers where the single-graph path has kernel boundaries.  Every step
// uses the same device routines and the same arithmetic as f2m_solve, so an
// instance's potentials, components and x equal those of a plain solve of
// the same graph bit for bit.
//
// Instances are grouped into size buckets; each bucket is one persistent
// launch with its own (lanes per node) schedule.  The whole sequence (queue
// reset kernels + one launch per bucket) is captured once into a hipGraph
// and replayed on later calls with the same batch and options.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

#include "f2m_exact_device.h"

namespace f2m {
namespace {


struct BArgs {
  int n_inst;
  const int64_t* node_off;
  const int64_t* rowptr;
  const int32_t* gnbr;
  const double* gcost;
  int32_t* nbr;
  double* D;
  int32_t* orig;
  double* F0;
  double* F1;
  int32_t *act, *deg, *clen, *A, *T, *root, *irr, *aux, *list;
  uint8_t* frozen;
  double *xw, *xc, *xg, *rn;
  uint8_t *fw, *fc;
  int32_t* efoff;  // per instance n+1 entries at node_off[i] + i
  ESlot* ef;       // compacted active lists, per instance at its slot base
                   // (16 B per slot: D doubles, then nbr ints, see Inst)
  int32_t fcap;    // doubles of dynamic shared memory for the potentials
  int32_t sweep_regs;    // 1, 2: runs of sweeps hold the lists in registers where they can (wg_sweeps_reg; 2: SH)
  int32_t part_lanes;    // 1: partitions with PV4_LANES lanes per node (part_v4_lanes / part_v2_lanes)
  const int32_t* order;  // instance ids of this bucket, decreasing size
  int32_t n_order;
  int32_t* queue;
  f2m_opts o;
  f2m_stats* stats;
  int32_t* cur_out;
  long long* clk;  // per-instance elapsed wall_clock64 ticks (autotuning), or null
  // exact recovery: per slot mirror + flow; per node prices, excesses,
  // lists, certificate terms (arrays of n_total at these bases)
  int64_t n_total;
  const double* F_init;  // warm start (n_total, normalised), or null
  int32_t* xmir;
  uint8_t* xf;
  double* xp;      // pL | pR | P terms | D terms
  int32_t* xe;     // eL | eR
  int32_t* xact;   // L list | R list
};

// Instance view: local node ids, pointers offset to the instance.
struct Inst {
  int n;
  int64_t nb, sb, m;
  const int64_t* rp;  // global slot offsets of local nodes (rowptr + nb)
  int32_t* nbr;
  double* D;
  int32_t* orig;
  const int32_t* gnbr;
  const double* gcost;
  double* F[2];
  int32_t *act, *deg, *clen, *A, *T, *root, *irr, *aux, *list;
  uint8_t* frozen;
  double *xw, *xc, *xg, *rn;
  uint8_t *fw, *fc;
  int32_t* efoff;
  double* efD;      // active slot k of node v at efoff[v] + k: its D ...
  int32_t* efN;     // ... and its neighbour (split arrays, 12 B per slot)
  XState X;            // exact recovery state (instance-local ids and slots)
  double *xpv, *xdv;   // certificate terms
  int32_t *xactL, *xactR;
  __device__ __forceinline__ int64_t b(int v) const { return rp[v] - sb; }
};

struct LocalRP {
  const int64_t* rp;
  int64_t sb;
  __device__ __forceinline__ int64_t operator()(int v) const { return rp[v] - sb; }
};

// Where a pass reads the potentials: FS_HBM from I.F[cur] (global memory),
// LDS copies otherwise -- FS_ONE: one shared-memory buffer holding F[cur]
// (after a writing pass it is refreshed from the new global values), FS_TWO:
// two buffers indexed like I.F (b_dyn + cur*n holds F[cur]).
enum { FS_HBM = 0, FS_ONE = 1, FS_TWO = 2 };
constexpr int PASS_U = 4;  // active slots per node loaded together in the shared-memory pass

// What a pass reads and writes, by value: the callers copy it into locals,
// so the pointers stay in scalar registers (the instance view itself lives
// in private memory, and a flat pointer to it would be re-read after every
// global store).
struct PView {
  int n;
  const int32_t* efoff;
  const double* efD;
  const int32_t* efN;
  double* F0;
  double* F1;
  const uint8_t* frozen;
  int lofs;  // lists in shared memory at b_dyn + lofs (efD, 16-bit efN; offsets efoff), or -1: global
  int ltot;  // their slots
};
__device__ __forceinline__ PView pview(const Inst& I) {
  return PView{I.n, I.efoff, I.efD, I.efN, I.F[0], I.F[1], I.frozen, -1, 0};
}

// The workgroup's dynamic shared memory (b_persistent's launch size): the
// potentials and active lists of the passes, later the exact stage's sums.
// Addressed from this symbol, so that the compiler emits LDS instructions.
extern __shared__ double b_dyn[];

// One pass over all nodes: gradient over the compacted active lists from
// F[cur]; if write, F[cur^1] = F[cur] + g (frozen nodes keep F).  Returns
// max |g| to every thread (barriers inside); 0 when !want_max (shared-memory
// modes: one barrier instead of the reduction).
//
// FS_HBM: LPN lanes per node, three dependent global loads per node (list
// bounds, slots, neighbour potentials).  Shared-memory modes: one thread per
// node; the list bounds of the thread's next node are loaded one step ahead,
// its first PASS_U slots together, the neighbour potentials from shared
// memory, so a step waits on one global load.  The arithmetic (dval, Top3,
// F + g) is the same in every mode and Top3 keeps the three smallest values
// whatever the order they come in, so all modes give the same bits.
// pre: list bounds of the thread's first two nodes (tid, tid + BS) as
// (b, e, b, e), loaded once per run of sweeps on the same lists
__device__ __forceinline__ int4 first_bounds(const PView& I, int bs) {
  int4 p = make_int4(0, 0, 0, 0);
  const int u = threadIdx.x;
  if (u < I.n) {
    p.x = I.efoff[u];
    p.y = I.efoff[u + 1];
  }
  if (u + bs < I.n) {
    p.z = I.efoff[u + bs];
    p.w = I.efoff[u + bs + 1];
  }
  return p;
}

template <int BS, int LPN, bool LL>
__device__ __forceinline__ double wg_pass(const PView I, int cur, bool write, bool use_frozen, double* sh, int fmode,
                                          int4 pre, bool want_max = true) {
  const double* Fin = cur ? I.F1 : I.F0;
  double* Fout = cur ? I.F0 : I.F1;
  double gmax = 0.0;
  if (fmode == FS_HBM) {
    const int total = I.n * LPN;
    for (int base = 0; base < total; base += BS) {
      const int gt = base + threadIdx.x;
      const int v = gt / LPN, lane = gt % LPN;
      Top3 t;
      double Fv = 0.0;
      const bool ok = v < I.n;
      if (ok) {
        const int b = I.efoff[v], e = I.efoff[v + 1];
        Fv = Fin[v];
        for (int k = b + lane; k < e; k += LPN) {
          const int w = I.efN[k];
          t.push(dval(I.efD[k], Fv, Fin[w], v, w));
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
      if (ok && lane == 0) {
        const double g = t.grad();
        if (write) Fout[v] = (use_frozen && I.frozen[v]) ? Fv : Fv + g;
        gmax = fmax(gmax, fabs(g));
      }
    }
    return block_max_all(gmax, sh);
  }
  const double* Fl = b_dyn + (fmode == FS_TWO ? cur * I.n : 0);
  double* Fl2 = b_dyn + (cur ^ 1) * I.n;  // written in FS_TWO only
  const int32_t* eO = I.efoff;
  const double* eD = I.efD;
  const int32_t* eN = I.efN;
  // neighbours as 16 bits: in global memory at efN (the shared-memory modes
  // hold instances of at most 65536 nodes), or in shared memory after D
  const uint16_t* eN16 = reinterpret_cast<const uint16_t*>(I.efN);
  if (LL) {  // lists in shared memory; offsets in global memory (prefetched)
    eD = b_dyn + I.lofs;
    eN16 = reinterpret_cast<const uint16_t*>(eD + I.ltot);
  }
  auto nbr_at = [&](int k) -> int { return (int)eN16[k]; };
  // node u's list bounds and first PASS_U slots (past its list: D = +inf and
  // the node itself, which give +inf, above every sentinel of Top3, so no
  // mask is needed); nodes past n have empty lists
  auto bounds = [&](int u, int& bb, int& ee) {
    bb = ee = 0;
    if (u < I.n) {
      bb = eO[u];
      ee = eO[u + 1];
    }
  };
  auto slots = [&](int u, int bb, int ee, double* Dk, int* wk) {
#pragma unroll
    for (int k = 0; k < PASS_U; ++k) {
      Dk[k] = __builtin_inf();
      wk[k] = u;
      if (k < ee - bb) {
        Dk[k] = eD[bb + k];
        wk[k] = nbr_at(bb + k);
      }
    }
  };
  // software pipeline over this thread's nodes v, v+BS, ...: bounds two
  // nodes ahead, slots one node ahead
  int b = pre.x, e = pre.y, b1 = pre.z, e1 = pre.w;
  double Dk[PASS_U];
  int wk[PASS_U];
  slots(threadIdx.x, b, e, Dk, wk);
#pragma unroll 2
  for (int base = 0; base < I.n; base += BS) {
    const int v = base + threadIdx.x;
    int b2, e2;
    bounds(v + 2 * BS, b2, e2);
    double Dn[PASS_U];
    int wn[PASS_U];
    slots(v + BS, b1, e1, Dn, wn);
    if (v < I.n) {
      const double Fv = Fl[v], hv = 0.5 * Fv;
      static_assert(PASS_U == 4, "the first slots go through top3_of4");
      double d[PASS_U];
#pragma unroll
      for (int k = 0; k < PASS_U; ++k) d[k] = dval_h(Dk[k], hv, 0.5 * Fl[wk[k]], v, wk[k]);  // +inf past the list
      Top3 t = top3_of4(d[0], d[1], d[2], d[3]);
      for (int k = b + PASS_U; k < e; ++k) {
        const int w = nbr_at(k);
        t.push(dval_h(eD[k], hv, 0.5 * Fl[w], v, w));
      }
      const double g = t.grad();
      if (write) {
        const double f = (use_frozen && I.frozen[v]) ? Fv : Fv + g;
        Fout[v] = f;
        if (fmode == FS_TWO) Fl2[v] = f;
      }
      gmax = fmax(gmax, fabs(g));
    }
    b = b1;
    e = e1;
    b1 = b2;
    e1 = e2;
#pragma unroll
    for (int k = 0; k < PASS_U; ++k) {
      Dk[k] = Dn[k];
      wk[k] = wn[k];
    }
  }
  if (write && fmode == FS_ONE) {
    __syncthreads();  // every read of the old potentials is done
    for (int v = threadIdx.x; v < I.n; v += BS) b_dyn[v] = Fout[v];  // this thread's own stores
  }
  if (!want_max) {
    __syncthreads();  // publishes the new shared values
    return 0.0;
  }
  return block_max_all(gmax, sh);  // its barriers publish the new shared values
}

// A run of cnt sweeps on fixed lists (sweep_n) for an instance of at most
// NP*BS nodes whose potentials (both buffers) and active lists sit in shared
// memory.  Each node's first RK active slots are read once for the whole
// run and kept in registers: D and the byte offsets of the two halves that
// dval_h adds to it, the lower id's first (padded slots: D = +inf and the
// node itself twice).  So a sweep reads only the halves, and the slots past
// RK, from shared memory, and never writes global memory.  RK = 4 (PASS_U):
// at two nodes per thread of a 1024-thread workgroup (128 registers) 6 or 8
// slots in registers made the C5 solve slower (19.9 / 25.8 against 18.0 ms).
//  - The thread's NP nodes are updated together: their 2*NP*RK reads are
//    issued before the first is used (a missing node reads slot 0 with
//    D = +inf and is not written).
//  - SORT: nodes owned in the order of their list lengths (a counting sort at
//    the start of the run; thread u takes ranks u, u + BS, ...), so that the
//    waves whose lists all fit the registers skip the loop over further
//    slots and the others loop to similar lengths.  It pays for lists longer
//    than RK on average (the block_GLOB lists); for shorter ones the
//    scattered reads of the owners' own halves cost more than it saves.
//    Which thread updates a node does not change its operands.
//  - During the run the two shared buffers hold the halves 0.5*F (sweep i
//    reads buffer i & 1); the owner keeps its nodes' full potentials in
//    registers and writes them back at the end, to the shared copy of F[c]
//    and to F[c] in global memory.  The other shared buffer is left
//    undefined: no pass reads it before writing it.
//  - 0.5*F is exact, and the operands and arithmetic of every node update
//    are wg_pass's, so the bits are the same.
// Returns the max |g| of the last sweep when want_max (else 0); barriers
// inside.  Shared scratch of the sort: the buffer F[c0 ^ 1] (n doubles,
// undefined on entry), which holds its KSORT + n ints when n >= KSORT.
//  - SH: a slot's owner half comes from the owner's register and only the
//    neighbour's from shared memory, the two put in id order by selects:
//    one shared-memory read per slot instead of two.  At 2000 nodes a sweep
//    reads 128 KB of halves otherwise, ~1000 cycles of a CU's shared-memory
//    bandwidth.
constexpr int KSORT = 64;  // list lengths sorted by; longer lists share the last bin
template <int BS, int NP, bool SORT, int RK, bool SH = false>
__device__ __attribute__((noinline)) double wg_sweeps_reg(const PView P, int c0, int cnt, bool use_frozen,
                                                         double* sh, bool want_max) {
  const int n = P.n;
  const double* eD = b_dyn + P.lofs;
  const uint16_t* eN16 = reinterpret_cast<const uint16_t*>(eD + P.ltot);
  const double* Fc = b_dyn + c0 * n;
  int* hist = reinterpret_cast<int*>(b_dyn + (c0 ^ 1) * n);
  int* rank = hist + KSORT;  // n entries: node of each rank
  if (SORT) {  // counting sort of the nodes by list length, in the free buffer
    for (int k = threadIdx.x; k < KSORT; k += BS) hist[k] = 0;
    __syncthreads();
    int len[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int v = threadIdx.x + p * BS;
      len[p] = 0;
      if (v < n) {
        len[p] = min(P.efoff[v + 1] - P.efoff[v], KSORT - 1);
        atomicAdd(&hist[len[p]], 1);
      }
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of the KSORT counts by one wave
      static_assert(KSORT == 64, "one count per lane");
      const int cnt_l = hist[threadIdx.x];
      int x = cnt_l;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if ((int)threadIdx.x >= off) x += y;
      }
      hist[threadIdx.x] = x - cnt_l;
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int v = threadIdx.x + p * BS;
      if (v < n) rank[atomicAdd(&hist[len[p]], 1)] = v;
    }
    __syncthreads();
  }
  int vv[NP];
  double Dk[NP][RK];
  uint32_t oa[NP][RK], ob[NP][RK];
  uint32_t lom[NP];  // SH: bit k = the owner has the lower id of slot k
  int bb[NP], L[NP];
  double fv[NP];
  bool fz[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int r = threadIdx.x + p * BS;
    const int v = (SORT && r < n) ? rank[r] : r;
    vv[p] = v;
    bb[p] = 0;
    L[p] = 0;
    fv[p] = 0.0;
    fz[p] = false;
    lom[p] = 0u;
    if (v < n) {
      bb[p] = P.efoff[v];
      L[p] = P.efoff[v + 1] - bb[p];
      fv[p] = Fc[v];
      fz[p] = use_frozen && P.frozen[v] != 0;
    }
#pragma unroll
    for (int k = 0; k < RK; ++k) {
      int w = v < n ? v : 0;  // a missing node reads slot 0, D = +inf
      Dk[p][k] = __builtin_inf();
      if (k < L[p]) {
        Dk[p][k] = eD[bb[p] + k];
        w = (int)eN16[bb[p] + k];
      }
      const int u = v < n ? v : 0;
      const bool lo = u < w;
      if (SH) {
        oa[p][k] = (uint32_t)w * 8u;
        ob[p][k] = 0u;
        if (lo) lom[p] |= 1u << k;
      } else {
        oa[p][k] = (uint32_t)(lo ? u : w) * 8u;
        ob[p][k] = (uint32_t)(lo ? w : u) * 8u;
      }

    }
  }
  __syncthreads();  // every read of F[c0] and of the ranks is done
#pragma unroll
  for (int p = 0; p < NP; ++p)
    if (vv[p] < n) b_dyn[vv[p]] = 0.5 * fv[p];
  __syncthreads();
  double gmax = 0.0;
  // one sweep reading the halves of buffer Q, writing buffer Q ^ 1
  auto sweep = [&](auto qc, auto lastc) {
    constexpr int Q = decltype(qc)::value;
    constexpr bool LAST = decltype(lastc)::value;
    const char* src = reinterpret_cast<const char*>(b_dyn + Q * n);
    const double* srcd = b_dyn + Q * n;
    double* dst = b_dyn + (Q ^ 1) * n;
    double d[NP][RK];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const double hv = 0.5 * fv[p];
#pragma unroll
      for (int k = 0; k < RK; ++k) {
        if (SH) {
          const double hw = *reinterpret_cast<const double*>(src + oa[p][k]);
          const bool lo = ((lom[p] >> k) & 1u) != 0;
          d[p][k] = (Dk[p][k] + (lo ? hv : hw)) + (lo ? hw : hv);
        } else {
          d[p][k] = (Dk[p][k] + *reinterpret_cast<const double*>(src + oa[p][k])) +
                    *reinterpret_cast<const double*>(src + ob[p][k]);
        }
      }
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int v = vv[p];
      Top3 t = top3_of4(d[p][0], d[p][1], d[p][2], d[p][3]);  // the same as pushing them
#pragma unroll
      for (int k = 4; k < RK; ++k) t.push(d[p][k]);
      if (L[p] > RK) {
        const double hv = 0.5 * fv[p];
        for (int k = RK; k < L[p]; ++k) {
          const int w = (int)eN16[bb[p] + k];
          t.push(dval_h(eD[bb[p] + k], hv, srcd[w], v, w));
        }
      }
      if (v < n) {
        const double g = t.grad();
        const double f = fz[p] ? fv[p] : fv[p] + g;
        fv[p] = f;
        dst[v] = 0.5 * f;
        if (LAST) gmax = fmax(gmax, fabs(g));
      }
    }
    __syncthreads();
  };
  using Q0 = std::integral_constant<int, 0>;
  using Q1 = std::integral_constant<int, 1>;
  using NL = std::integral_constant<bool, false>;
  using LA = std::integral_constant<bool, true>;
  int i = 0;
  for (; i + 2 < cnt; i += 2) {
    sweep(Q0{}, NL{});
    sweep(Q1{}, NL{});
  }
  if (cnt - i == 2) {
    sweep(Q0{}, NL{});
    if (want_max) sweep(Q1{}, LA{});
    else sweep(Q1{}, NL{});
  } else if (cnt - i == 1) {
    if (want_max) sweep(Q0{}, LA{});
    else sweep(Q0{}, NL{});
  }
  // full potentials back, to F[c] (the owner's entries; nobody reads the
  // shared potentials before the barrier below)
  const int c = c0 ^ (cnt & 1);
  double* Fg = c ? P.F1 : P.F0;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int v = vv[p];
    if (v < n) {
      b_dyn[c * n + v] = fv[p];
      Fg[v] = fv[p];
    }
  }
  if (want_max) return block_max_all(gmax, sh);  // its barriers publish the potentials
  __syncthreads();
  return 0.0;
}

// update_EF (graph_cuda.cu:229-252) inside the workgroup: efoff = exclusive
// scan of act (global memory), then each node's active prefix copied into the
// split lists -- in shared memory at b_dyn + lofs (efD doubles, then efN as
// 16 bits; the offsets stay in global memory) when they fit in room bytes
// (lofs >= 0, n <= 65536), else in global memory (neighbours as 16 bits too
// when lofs >= 0: the potentials then fit in shared memory, so n <= 20480).  Returns in (lofs, ltot) the shared offset (-1: global) and the
// slot count.
template <int BS>
__device__ void wg_ef_rebuild(const Inst& I, int* shi, int lofs, int room, int& lofs_out, int& ltot) {
  // the shared-memory passes (lofs >= 0) read 16-bit neighbours, the others 32-bit
  uint16_t* gN16 = reinterpret_cast<uint16_t*>(I.efN);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int carry = 0;
  for (int base = 0; base < I.n; base += BS) {
    const int v = base + threadIdx.x;
    const int a = v < I.n ? I.act[v] : 0;
    int inc = a;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(inc, off, 64);
      if (lane >= off) inc += t;
    }
    __syncthreads();
    if (lane == 63) shi[wid] = inc;
    __syncthreads();
    int pre = 0, tot = 0;
    for (int k = 0; k < BS / 64; ++k) {
      if (k < wid) pre += shi[k];
      tot += shi[k];
    }
    if (v < I.n) I.efoff[v] = carry + pre + inc - a;
    carry += tot;
  }
  if (threadIdx.x == 0) I.efoff[I.n] = carry;
  __syncthreads();
  // carry (the total) is the same in every thread
  const bool ll = lofs >= 0 && I.n <= 65536 && (int64_t)carry * 10 <= room;
  if (ll) {
    double* dD = b_dyn + lofs;
    uint16_t* dN = reinterpret_cast<uint16_t*>(dD + carry);
    for (int v = threadIdx.x; v < I.n; v += BS) {
      const int64_t b = I.b(v);
      const int o = I.efoff[v], L = I.act[v];
      for (int k = 0; k < L; ++k) {
        dD[o + k] = I.D[b + k];
        dN[o + k] = (uint16_t)I.nbr[b + k];
      }
    }
  } else {
    for (int v = threadIdx.x; v < I.n; v += BS) {
      const int64_t b = I.b(v);
      const int o = I.efoff[v], L = I.act[v];
      for (int k = 0; k < L; ++k) {
        I.efD[o + k] = I.D[b + k];
        if (lofs >= 0) gN16[o + k] = (uint16_t)I.nbr[b + k];
        else I.efN[o + k] = I.nbr[b + k];
      }
    }
  }
  __syncthreads();
  lofs_out = ll ? lofs : -1;
  ltot = carry;
}

// Ascending ids v in [0,n) with pred(v) into out; returns the count.
template <int BS, class P>
__device__ int wg_compact(int n, P pred, int32_t* out, int* shi) {
  int total = 0;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int base = 0; base < n; base += BS) {
    const int v = base + threadIdx.x;
    const int f = (v < n && pred(v)) ? 1 : 0;
    const unsigned long long bal = __ballot(f);
    const int wpre = __popcll(bal & ((1ull << lane) - 1ull));
    __syncthreads();
    if (lane == 0) shi[wid] = __popcll(bal);
    __syncthreads();
    int off = 0, tot = 0;
    for (int k = 0; k < BS / 64; ++k) {
      if (k < wid) off += shi[k];
      tot += shi[k];
    }
    if (f) out[total + off + wpre] = v;
    total += tot;
  }
  __syncthreads();
  return total;
}

// The lane partitions as calls: inlined, their registers join the
// persistent solve's and it spills at 1024 threads
__device__ __attribute__((noinline)) int b_part_v4_lanes(int v, int q, int64_t b, int dg, int kk,
                                                         int32_t* __restrict__ nbr, double* __restrict__ D,
                                                         int32_t* __restrict__ orig, const double* __restrict__ F) {
  return part_v4_lanes(v, q, b, dg, kk, nbr, D, orig, F);
}
__device__ __attribute__((noinline)) int b_part_v2_lanes(int v, int q, int64_t b, int dg, double thr,
                                                         int32_t* __restrict__ nbr, double* __restrict__ D,
                                                         int32_t* __restrict__ orig, const double* __restrict__ F) {
  return part_v2_lanes(v, q, b, dg, thr, nbr, D, orig, F);
}

template <int BS, int LPN>
struct Sched {
  const Inst& I;
  const f2m_opts& o;
  double* sh;
  int* shi;
  int fmode;              // FS_HBM, FS_ONE or FS_TWO
  PView pv;               // the passes' view of I (lists in shared or global memory)
  int dynb;               // bytes of shared memory at b_dyn
  int regs = 0;           // runs of sweeps may hold the lists in registers (wg_sweeps_reg)
  int plane = 0;          // partitions with PV4_LANES lanes per node
  int cur = 0;
  int64_t sweeps = 0;
  double eps;
  bool fok = false;       // the shared copy holds F[cur]

  // (re)load the shared copy after the potentials were set outside a pass
  __device__ void fs_load() {
    if (fmode == FS_HBM || fok) return;
    double* dst = b_dyn + (fmode == FS_TWO ? cur * I.n : 0);
    for (int v = threadIdx.x; v < I.n; v += BS) dst[v] = I.F[cur][v];
    __syncthreads();
    fok = true;
  }
  // act[v] = part(v, F) for every node, F = the current potentials: the
  // shared copy when there is one (two call sites, so that each reads its
  // memory with its own instructions)
  template <class P>
  __device__ void partition(P part) {
    fs_load();
    if (fmode == FS_HBM) {
      const double* F = I.F[cur];
      for (int v = threadIdx.x; v < I.n; v += BS) I.act[v] = part(v, F);
    } else {
      const double* F = b_dyn + (fmode == FS_TWO ? cur * I.n : 0);
      for (int v = threadIdx.x; v < I.n; v += BS) I.act[v] = part(v, F);
    }
    __syncthreads();
  }
  // the same with PV4_LANES lanes per node: part(v, q, F) is called by the
  // node's four lanes together and its result is valid on lane q = 0
  template <class P>
  __device__ void partition_lanes(P part) {
    fs_load();
    const int q = threadIdx.x % PV4_LANES;
    if (fmode == FS_HBM) {
      const double* F = I.F[cur];
      for (int v = threadIdx.x / PV4_LANES; v < I.n; v += BS / PV4_LANES) {
        const int p = part(v, q, F);
        if (q == 0) I.act[v] = p;
      }
    } else {
      const double* F = b_dyn + (fmode == FS_TWO ? cur * I.n : 0);
      for (int v = threadIdx.x / PV4_LANES; v < I.n; v += BS / PV4_LANES) {
        const int p = part(v, q, F);
        if (q == 0) I.act[v] = p;
      }
    }
    __syncthreads();
  }
  // active lists rebuilt after a partition; into the shared memory after the
  // potentials when they fit
  __device__ void rebuild() {
    int lofs = -1, room = 0;
    if (fmode != FS_HBM) {
      lofs = (fmode == FS_TWO ? 2 : 1) * I.n;  // doubles of potentials before the lists
      room = dynb - lofs * 8;
    }
    wg_ef_rebuild<BS>(I, shi, lofs, room, pv.lofs, pv.ltot);
  }
  __device__ double grad_act() {
    fs_load();
    const int4 pre = first_bounds(pv, BS);
    if (pv.lofs >= 0) return wg_pass<BS, LPN, true>(pv, cur, false, false, sh, fmode, pre);
    return wg_pass<BS, LPN, false>(pv, cur, false, false, sh, fmode, pre);
  }
  // cnt Jacobi sweeps; returns max |g| of the last one when want_max (else
  // 0).  The loop runs on local copies of the state.
  __device__ double sweep_n(int cnt, bool frozen, bool want_max = false) {
    fs_load();
    const PView P = pv;
    double* const shm = sh;
    const int fm = fmode;
    int c = cur;
    double m = 0.0;
    if (regs && cnt > 0 && fm == FS_TWO && P.lofs >= 0 && P.n <= 2 * BS && P.n >= KSORT) {
      // sorted ownership where the lists are longer than the registers hold on average
      if (regs == 2) {  // the owners' halves from registers
        if (P.ltot > PASS_U * P.n)
          m = P.n <= BS ? wg_sweeps_reg<BS, 1, true, PASS_U, true>(P, c, cnt, frozen, shm, want_max)
                        : wg_sweeps_reg<BS, 2, true, PASS_U, true>(P, c, cnt, frozen, shm, want_max);
        else
          m = P.n <= BS ? wg_sweeps_reg<BS, 1, false, PASS_U, true>(P, c, cnt, frozen, shm, want_max)
                        : wg_sweeps_reg<BS, 2, false, PASS_U, true>(P, c, cnt, frozen, shm, want_max);
      } else if (P.ltot > PASS_U * P.n) {
        m = P.n <= BS ? wg_sweeps_reg<BS, 1, true, PASS_U>(P, c, cnt, frozen, shm, want_max)
                      : wg_sweeps_reg<BS, 2, true, PASS_U>(P, c, cnt, frozen, shm, want_max);
      } else {
        m = P.n <= BS ? wg_sweeps_reg<BS, 1, false, PASS_U>(P, c, cnt, frozen, shm, want_max)
                      : wg_sweeps_reg<BS, 2, false, PASS_U>(P, c, cnt, frozen, shm, want_max);
      }
      cur = c ^ (cnt & 1);
      sweeps += cnt;
      return m;
    }
    const int4 pre = first_bounds(P, BS);  // the lists do not change during these sweeps
    if (P.lofs >= 0) {
      for (int i = 0; i < cnt; ++i) {
        m = wg_pass<BS, LPN, true>(P, c, true, frozen, shm, fm, pre, want_max && i == cnt - 1);
        c ^= 1;
      }
    } else {
      for (int i = 0; i < cnt; ++i) {
        m = wg_pass<BS, LPN, false>(P, c, true, frozen, shm, fm, pre, want_max && i == cnt - 1);
        c ^= 1;
      }
    }
    cur = c;
    sweeps += cnt;
    return m;
  }
  __device__ int block_glob() {
    for (int v = threadIdx.x; v < I.n; v += BS) I.act[v] = I.deg[v] > o.glob_k ? o.glob_k : I.deg[v];
    __syncthreads();
    rebuild();
    int go = 0;
    for (; go < o.glob_rounds; ++go) {
      sweep_n(o.glob_sweeps, false);
      if (plane)
        partition_lanes([&](int v, int q, const double* F) {
          const int dg = I.deg[v];
          return dg < o.glob_k ? dg : b_part_v4_lanes(v, q, I.b(v), dg, o.glob_k, I.nbr, I.D, I.orig, F);
        });
      else
        partition([&](int v, const double* F) {
          return part_v4_node(v, I.b(v), I.deg[v], o.glob_k, I.nbr, I.D, I.orig, F);
        });
      rebuild();
      const double M = grad_act();
      if (M < o.glob_exit) { ++go; break; }
    }
    return go;
  }
  __device__ double check_opt() {
    if (plane)
      partition_lanes([&](int v, int q, const double* F) {
        const int dg = I.deg[v];
        return dg < 10 ? dg : b_part_v2_lanes(v, q, I.b(v), dg, 1.0, I.nbr, I.D, I.orig, F);
      });
    else
      partition([&](int v, const double* F) {
        return part_v2_node(v, I.b(v), I.deg[v], 1.0, I.nbr, I.D, I.orig, F);
      });
    rebuild();
    return grad_act();
  }
  // block_EF; MA = its opening max |g|, which check_opt (run right before
  // by every caller) computed on the same lists and potentials
  __device__ void block_ef(int N, bool frozen, double MA) {
    if (!o.early_exit) {
      const int cnt = (MA < o.err_grad) ? min(N, o.check_every) : N;
      sweep_n(cnt, frozen);
      return;
    }
    int done = 0;
    while (done < N) {
      const int chunk = min(o.check_every, N - done);
      const double MB = sweep_n(chunk, frozen, true);
      done += chunk;
      if (chunk == o.check_every && MB < o.err_grad) break;
    }
  }
  __device__ int run_start(double& M) {
    M = check_opt();
    int factor = o.factor0, go = 1;
    for (; go <= o.max_outer; ++go) {
      block_ef(factor, false, M);
      M = check_opt();
      if (factor < o.factor_max) factor += o.factor_step;
      if (M < o.err_grad) break;
    }
    return go;
  }
  __device__ double run_fix() {
    double M = check_opt();
    for (int go = 1; go <= o.fix_rounds; ++go) {
      block_ef(o.fix_factor, false, M);
      M = check_opt();
      if (M < o.err_grad) break;
    }
    return M;
  }
  // classes, tight components, primal, least squares; returns #IRREGULAR_B
  __device__ int combs_check(int& ncombs, int& irr_a) {
    const double* F = I.F[cur];
    for (int v = threadIdx.x; v < I.n; v += BS) {
      int cl, a, t;
      classify_node(v, I.b(v), I.act[v], eps, I.nbr, I.D, I.orig, F, cl, a, t);
      I.clen[v] = cl;
      I.A[v] = a;
      I.T[v] = t;
      I.aux[v] = v;  // union-find parents
    }
    __syncthreads();
    for (int v = threadIdx.x; v < I.n; v += BS) {
      const int64_t b = I.b(v);
      const double Fv = F[v];
      for (int k = 0; k < I.clen[v]; ++k) {
        const int w = I.nbr[b + k];
        if (w <= v) continue;
        const double d = dval(I.D[b + k], Fv, F[w], v, w);
        if (d >= -eps && d <= eps) uf_unite(I.aux, v, w);
      }
    }
    __syncthreads();
    for (int v = threadIdx.x; v < I.n; v += BS) I.root[v] = I.T[v] > 0 ? uf_find(I.aux, v) : -1;
    __syncthreads();
    for (int v = threadIdx.x; v < I.n; v += BS) I.irr[v] = 0;
    __syncthreads();
    double nc = 0.0;
    for (int v = threadIdx.x; v < I.n; v += BS) {
      const int r = I.root[v];
      if (r >= 0 && 4 - 2 * I.A[v] - I.T[v] != 0) I.irr[r] = 1;
      if (r == v) nc += 1.0;
    }
    ncombs = (int)block_sum_all(nc, sh);
    for (int v = threadIdx.x; v < I.n; v += BS) {
      const int r = I.root[v];
      primal_node(v, I.b(v), I.deg[v], I.clen[v], eps, I.nbr, I.D, F, r >= 0 && I.irr[r] == 0, I.xw, I.fw);
    }
    __syncthreads();
    const Inst& II = I;
    const int nirr = wg_compact<BS>(I.n, [&](int v) { return II.root[v] == v && II.irr[v] != 0; }, I.aux, shi);
    irr_a = nirr;
    int irr_b = 0;
    LocalRP rpf{I.rp, I.sb};
    // with exact recovery on, irregular combs go to the exact stage: their
    // least squares (and the dual repairs it would start) are not run
    for (int i = 0; o.exact == 0 && i < nirr; ++i) {
      const int r = I.aux[i];
      const int m = wg_compact<BS>(I.n, [&](int v) { return II.root[v] == r; }, I.list, shi);
      const int go = axb_comb(I.list, m, rpf, I.clen, I.A, I.T, eps, I.nbr, I.D, F, I.xw, I.xg, I.rn, sh);
      if (go == 0) {
        if (threadIdx.x == 0) I.irr[r] = 2;
        ++irr_b;
      }
      __syncthreads();
    }
    return irr_b;
  }
  // shift_comb_IRREGULAR_B (graph_cuda_comb.hpp:365-403) on every comb with
  // irr == 2; type 1 shifts only the smallest violating node of each comb.
  // Not inlined: its divergent stores end in a join block, and inlined the
  // compiler placed a spill of the caller's shift value at the top of that
  // block, ahead of the exec restore (the build's ISA scan rejects that).
  __device__ __attribute__((noinline)) void shift(double sv, int type) {
    double* F = I.F[cur];
    fok = false;  // F changes below
    for (int v = threadIdx.x; v < I.n; v += BS) {
      I.frozen[v] = 0;
      I.list[v] = INT_MAX;
    }
    __syncthreads();
    if (type == 1) {
      for (int v = threadIdx.x; v < I.n; v += BS) {
        const int r = I.root[v];
        if (r >= 0 && I.irr[r] == 2 && 4 - 2 * I.A[v] - I.T[v] != 0) atomicMin(&I.list[r], v);
      }
      __syncthreads();
    }
    for (int v = threadIdx.x; v < I.n; v += BS) {
      const int r = I.root[v];
      const int s = 4 - 2 * I.A[v] - I.T[v];
      if (r >= 0 && I.irr[r] == 2 && s != 0 && (type == 0 || I.list[r] == v)) {
        F[v] += sv * (s > 0 ? 1.0 : -1.0);
        I.frozen[v] = 1;
      }
    }
    __syncthreads();
  }
};

// Fixed-order sums of the certificate terms, equal to the single-graph
// stage's: DET_T contiguous chunks, each summed in order from 0, then the
// chunk sums in order from 0.  xdet: 2*BS doubles of shared memory.
template <int BS>
__device__ void wg_det_sum(const double* pv, const double* dv, int n, double* xdet, double& P, double& D) {
  const int64_t chunk = ((int64_t)n + DET_T - 1) / DET_T;
  double accP = 0.0, accD = 0.0;
  for (int k0 = 0; k0 < DET_T; k0 += BS) {
    const int t = k0 + threadIdx.x;
    double a = 0.0, d = 0.0;
    if (t < DET_T) {
      const int64_t b = t * chunk;
      const int64_t e = b + chunk < n ? b + chunk : n;
      for (int64_t i = b; i < e; ++i) {
        a += pv[i];
        d += dv[i];
      }
    }
    xdet[threadIdx.x] = a;
    xdet[BS + threadIdx.x] = d;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int j = 0; j < BS && k0 + j < DET_T; ++j) {
        accP += xdet[j];
        accD += xdet[BS + j];
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    xdet[0] = accP;
    xdet[1] = accD;
  }
  __syncthreads();
  P = xdet[0];
  D = xdet[1];
  __syncthreads();
}

// Exact primal recovery of one instance by its workgroup: the schedule of the
// single-graph stage (f2m_exact.hip) with workgroup barriers for kernel
// boundaries and shared-memory list counters, so rounds cost two barriers.
// F: the instance's current potentials (normalised), overwritten with the
// certifying ones; x / X-flags into I.xc / I.fc; root: x = 1/2 components.
// Returns 0 certified, 1 not certified, -1 asymmetric, -2 degree below 2.
template <int BS>
__device__ int wg_exact(const Inst& I, const f2m_opts& o, double dmin, double* F, f2m_stats& z, double* sh,
                        double* xdet, int* xcnt) {
  const XState& X = I.X;
  const LocalRP rpf{I.rp, I.sb};
  const int lane = threadIdx.x % XG, gid = threadIdx.x / XG;
  constexpr int NG = BS / XG;
  if (threadIdx.x == 0) xcnt[4] = xcnt[5] = xcnt[6] = 0;
  __syncthreads();
  for (int v = threadIdx.x; v < I.n; v += BS) {
    if (I.b(v + 1) - I.b(v) < 2) atomicOr(&xcnt[5], 1);
    bool nonint = false;
    if (!x_mirror_node(v, rpf, X.nbr, X.c, const_cast<int32_t*>(X.mir), nonint)) atomicOr(&xcnt[4], 1);
    if (nonint) atomicOr(&xcnt[6], 1);
  }
  __syncthreads();
  z.exact_ran = 1;
  if (xcnt[4]) return -1;
  if (xcnt[5]) return -2;
  const bool integral = xcnt[6] == 0;
  for (int v = threadIdx.x; v < I.n; v += BS) {
    const double p = 0.5 * (dmin * F[v]);
    X.pL[v] = p;
    X.pR[v] = -p;
  }
  __syncthreads();
  for (int v = threadIdx.x; v < I.n; v += BS) x_flow_node<false>(X, v, (int)I.b(v), (int)I.b(v + 1), 0.0);
  __syncthreads();
  double eps = o.exact_eps0 * dmin;
  int64_t r = 0, rounds = 0;
  int phases = 0;
  bool stuck = false, certified = false;
  double P = 0.0, D = 0.0;
  for (;;) {
    if (phases > 0) {
      for (int v = threadIdx.x; v < I.n; v += BS) x_flow_node<true, true>(X, v, (int)I.b(v), (int)I.b(v + 1), eps);
      __syncthreads();
    }
    // excesses; lists of round r: counters xcnt[par] (L), xcnt[2 + par] (R)
    if (threadIdx.x < 4) xcnt[threadIdx.x] = 0;
    __syncthreads();
    {
      const int par = (int)(r & 1);
      for (int v = threadIdx.x; v < I.n; v += BS) {
        int el, er;
        if (phases == 0) {
          x_excess_node(X, v, (int)I.b(v), (int)I.b(v + 1), el, er);
        } else {  // maintained through the flips of the saturation pass
          el = X.eL[v];
          er = X.eR[v];
        }
        if (el > 0) I.xactL[atomicAdd(&xcnt[par], 1)] = v;
        if (er > 0) I.xactR[atomicAdd(&xcnt[2 + par], 1)] = v;
      }
    }
    __syncthreads();
    for (;;) {
      if (rounds >= o.exact_max_rounds) { stuck = true; break; }
      const int par = (int)(r & 1);
      const int nl = xcnt[par];
      if (threadIdx.x == 0) xcnt[par ^ 1] = 0;  // L list of round r+1, filled by the R step
      for (int i = gid; i < nl; i += NG) {
        const int v = I.xactL[i];
        x_discharge_l_any<4>(X, v, (int)I.b(v), (int)I.b(v + 1), lane, eps,
                      [&](int w) { I.xactR[atomicAdd(&xcnt[2 + par], 1)] = w; });
      }
      __syncthreads();
      const int nr = xcnt[2 + par];
      if (threadIdx.x == 0) xcnt[2 + (par ^ 1)] = 0;  // R list of round r+1, filled by its L step
      for (int i = gid; i < nr; i += NG) {
        const int w = I.xactR[i];
        x_discharge_r_any<4>(X, w, (int)I.b(w), (int)I.b(w + 1), lane, eps,
                      [&](int v) { I.xactL[atomicAdd(&xcnt[par ^ 1], 1)] = v; });
      }
      __syncthreads();
      ++r;
      if (nl + nr == 0) break;
      ++rounds;
    }
    ++phases;
    if (stuck) break;
    for (int v = threadIdx.x; v < I.n; v += BS) x_cert_node(X, v, (int)I.b(v), (int)I.b(v + 1), I.xpv[v], I.xdv[v]);
    __syncthreads();
    wg_det_sum<BS>(I.xpv, I.xdv, I.n, xdet, P, D);
    const double tol = fmax(integral ? 0.4 : 0.0, o.exact_tol * fabs(P));
    if (P - D <= tol) { certified = true; break; }
    eps = eps / (double)o.exact_alpha;
    if (phases >= o.exact_max_phases) break;
  }
  if (stuck) {
    for (int v = threadIdx.x; v < I.n; v += BS) x_cert_node(X, v, (int)I.b(v), (int)I.b(v + 1), I.xpv[v], I.xdv[v]);
    __syncthreads();
    wg_det_sum<BS>(I.xpv, I.xdv, I.n, xdet, P, D);
  }
  // outputs, then the x = 1/2 components (parents in I.aux)
  for (int v = threadIdx.x; v < I.n; v += BS) {
    for (int64_t s = I.b(v); s < I.b(v + 1); ++s) {
      const double xv = 0.5 * (double)(X.f[s] + X.f[X.mir[s]]);
      I.xc[s] = xv;
      I.fc[s] = xv > 0.0 ? 1 : 0;
    }
    F[v] = (X.pL[v] - X.pR[v]) / dmin;
    I.aux[v] = v;
  }
  __syncthreads();
  for (int v = threadIdx.x; v < I.n; v += BS)
    for (int64_t s = I.b(v); s < I.b(v + 1); ++s) {
      const int w = X.nbr[s];
      if (w > v && X.f[s] + X.f[X.mir[s]] == 1) uf_unite(I.aux, v, w);
    }
  __syncthreads();
  double nroots = 0.0;
  for (int v = threadIdx.x; v < I.n; v += BS) {
    bool has = false;
    for (int64_t s = I.b(v); s < I.b(v + 1) && !has; ++s) has = X.f[s] + X.f[X.mir[s]] == 1;
    const int rt = has ? uf_find(I.aux, v) : -1;
    I.root[v] = rt;
    if (rt == v) nroots += 1.0;
  }
  z.n_combs = (int)block_sum_all(nroots, sh);
  z.exact_phases = phases;
  z.exact_rounds = rounds;
  z.exact_gap = P - D;
  z.exact_certified = certified ? 1 : 0;
  return certified ? 0 : 1;
}

template <int BS, int LPN>
__device__ __attribute__((noinline)) void solve_instance(const BArgs& a, int inst, double* sh, int* shi, double* xdet, int* xcnt,
                               double* sdmin) {
  Inst I;
  I.nb = a.node_off[inst];
  I.n = (int)(a.node_off[inst + 1] - I.nb);
  I.rp = a.rowptr + I.nb;
  I.sb = I.rp[0];
  I.m = I.rp[I.n] - I.sb;
  I.nbr = a.nbr + I.sb;
  I.D = a.D + I.sb;
  I.orig = a.orig + I.sb;
  I.gnbr = a.gnbr + I.sb;
  I.gcost = a.gcost + I.sb;
  I.F[0] = a.F0 + I.nb;
  I.F[1] = a.F1 + I.nb;
  I.act = a.act + I.nb;
  I.deg = a.deg + I.nb;
  I.clen = a.clen + I.nb;
  I.A = a.A + I.nb;
  I.T = a.T + I.nb;
  I.root = a.root + I.nb;
  I.irr = a.irr + I.nb;
  I.aux = a.aux + I.nb;
  I.list = a.list + I.nb;
  I.frozen = a.frozen + I.nb;
  I.xw = a.xw + I.sb;
  I.xc = a.xc + I.sb;
  I.xg = a.xg + I.sb;
  I.rn = a.rn + I.nb;
  I.fw = a.fw + I.sb;
  I.fc = a.fc + I.sb;
  I.efoff = a.efoff + I.nb + inst;
  I.efD = reinterpret_cast<double*>(a.ef + I.sb);  // the instance's 16 B per slot hold
  I.efN = reinterpret_cast<int32_t*>(I.efD + I.m);  // m doubles, then m ints
  I.X.nbr = I.gnbr;
  I.X.c = I.gcost;
  I.X.mir = a.xmir + I.sb;
  I.X.f = a.xf + I.sb;
  I.X.pL = a.xp + I.nb;
  I.X.pR = a.xp + a.n_total + I.nb;
  I.xpv = a.xp + 2 * a.n_total + I.nb;
  I.xdv = a.xp + 3 * a.n_total + I.nb;
  I.X.eL = a.xe + I.nb;
  I.X.eR = a.xe + a.n_total + I.nb;
  I.xactL = a.xact + I.nb;
  I.xactR = a.xact + a.n_total + I.nb;
  const f2m_opts& o = a.o;
  f2m_stats z;
  memset(&z, 0, sizeof z);

  // d_min: smallest positive cost (graph_tour.hpp:94)
  double mn = 1.0 / 0.0;
  for (int64_t s = threadIdx.x; s < I.m; s += BS) {
    const double c = I.gcost[s];
    if (c > 0 && c < mn) mn = c;
  }
  const double dmin = -block_max_all(-mn, sh);
  if (!(dmin > 0) || isinf(dmin)) {
    if (threadIdx.x == 0) {
      z.status = -1;
      a.stats[inst] = z;
    }
    return;
  }
  // d_min is kept in shared memory for its later uses: a register holding it
  // across the whole solve gets spilled in the 1024-thread variants, and the
  // compiler has placed such a spill store where no lane is active (before a
  // join block's exec restore), losing it (the build scans the ISA for this)
  if (threadIdx.x == 0) *sdmin = dmin;
  for (int64_t s = threadIdx.x; s < I.m; s += BS) {
    I.nbr[s] = I.gnbr[s];
    I.D[s] = I.gcost[s] / dmin;
    I.orig[s] = (int32_t)s;
  }
  for (int v = threadIdx.x; v < I.n; v += BS) {
    I.deg[v] = (int)(I.rp[v + 1] - I.rp[v]);
    I.F[0][v] = a.F_init ? a.F_init[I.nb + v] : 0.0;
    I.frozen[v] = 0;
  }
  __syncthreads();
  z.d_min = dmin;
  // the potentials in shared memory when they fit (both buffers, else one);
  // the same space holds xdet in the exact stage, after the last pass
  const int fmode = 2 * I.n <= a.fcap ? FS_TWO : (I.n <= a.fcap ? FS_ONE : FS_HBM);
  Sched<BS, LPN> S{I, o, sh, shi, fmode, pview(I), a.fcap * 8};
  S.regs = a.sweep_regs;
  S.plane = a.part_lanes;
  S.eps = o.err_grad * o.err_comb_mult;
  z.glob_rounds = S.block_glob();  // from F_init on a warm start
  double M = 0.0;
  if (o.stop_phase != 1) z.outer_rounds = S.run_start(M);
  z.max_grad = M;
  if (o.stop_phase == 1 || o.stop_phase == 2) {
    z.sweeps = S.sweeps;
    if (threadIdx.x == 0) {
      a.stats[inst] = z;
      a.cur_out[inst] = S.cur;
    }
    __syncthreads();
    return;
  }
  int irr_b = S.combs_check(z.n_combs, z.n_irregular_a);
  z.n_irregular_b = irr_b;
  if (irr_b > 0 && o.repair) {
    double shiftV = o.shift0;
    auto attempt = [&](double sv, int type) {
      ++z.n_repairs;
      S.shift(sv, type);
      S.sweep_n(o.frozen_sweeps, true);
      for (int v = threadIdx.x; v < I.n; v += BS) I.frozen[v] = 0;
      __syncthreads();
      M = S.run_fix();
      return S.combs_check(z.n_combs, z.n_irregular_a);
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
      if (fabs(shiftV) > o.shift_max) break;
    }
    z.max_grad = M;
  }
  z.status = irr_b == 0 ? 0 : 1;
  // x to canonical slots
  for (int64_t s = threadIdx.x; s < I.m; s += BS) {
    const int oo = I.orig[s];
    I.xc[oo] = I.xw[s];
    I.fc[oo] = I.fw[s];
  }
  __syncthreads();
  // exact recovery where the reference recovery is not exact (opts.exact 1)
  // or always (2), as in the single-graph solve
  if (o.exact >= 2 || (o.exact == 1 && (z.status != 0 || z.n_irregular_a > 0))) {
    f2m_stats t = z;
    const int rc = wg_exact<BS>(I, o, *sdmin, I.F[S.cur], t, sh, xdet, xcnt);
    if (rc >= 0) {
      z = t;
      z.status = rc == 0 ? 0 : 1;
    } else {
      z.exact_ran = 1;
      z.status = 1;
    }
    __syncthreads();
  }
  // objective, dual, MAX edge rank
  const double* F = I.F[S.cur];
  double obj = 0.0, sf = 0.0, sneg = 0.0, mr = 0.0;
  for (int v = threadIdx.x; v < I.n; v += BS) {
    const int64_t b = I.b(v), e = I.b(v + 1);
    const double Fv = F[v];
    sf += Fv;
    for (int64_t s = b; s < e; ++s) {
      const int w = I.gnbr[s];
      if (w > v) {
        obj += I.xc[s] * I.gcost[s];
        if (I.fc[s]) {
          int r = (int)(s - b);
          for (int64_t q = I.b(w); q < I.b(w + 1); ++q)
            if (I.gnbr[q] == v) { r = min(r, (int)(q - I.b(w))); break; }
          mr = fmax(mr, (double)r);
        }
      }
      const int ww = I.nbr[s];
      if (ww > v) sneg += fmin(0.0, dval(I.D[s], Fv, F[ww], v, ww));
    }
  }
  obj = block_sum_all(obj, sh);
  sf = block_sum_all(sf, sh);
  sneg = block_sum_all(sneg, sh);
  mr = block_max_all(mr, sh);
  z.objective = obj;
  z.dual = *sdmin * (-sf + sneg);
  z.max_rank = (int)mr;
  z.sweeps = S.sweeps;
  if (threadIdx.x == 0) {
    a.stats[inst] = z;
    a.cur_out[inst] = S.cur;
  }
  __syncthreads();
}

// Dynamic shared memory: a.fcap doubles (at least 2*BS) holding the
// potentials of the instance's passes and, where they fit, its active lists;
// after the last pass, the exact stage's xdet.
// 256-thread workgroups: at least 3 waves per SIMD (<= 168 VGPRs), three
// workgroups per CU; the larger ones run one workgroup per CU whatever their
// register count below 256.
template <int BS, int LPN>
__global__ void __launch_bounds__(BS, BS == 256 ? 3 : 1) b_persistent(BArgs a) {
  double* const xdet = b_dyn;
  __shared__ double sh[BS / 64];
  __shared__ int shi[BS / 64 + 1];
  __shared__ int xcnt[8];
  __shared__ double sdmin;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) shi[BS / 64] = atomicAdd(a.queue, 1);
    __syncthreads();
    const int idx = shi[BS / 64];
    if (idx >= a.n_order) break;
    const long long t0 = wall_clock64();
    solve_instance<BS, LPN>(a, a.order[idx], sh, shi, xdet, xcnt, &sdmin);
    if (a.clk && threadIdx.x == 0) a.clk[a.order[idx]] = wall_clock64() - t0;
  }
}

// the potentials of instance blockIdx.x from its current buffer
__global__ void b_gather_F(const int64_t* __restrict__ node_off, const int32_t* __restrict__ cur,
                           const double* __restrict__ F0, const double* __restrict__ F1, double* __restrict__ out) {
  const int i = blockIdx.x;
  const int64_t a = node_off[i], e = node_off[i + 1];
  const double* src = cur[i] ? F1 : F0;
  for (int64_t v = a + threadIdx.x; v < e; v += blockDim.x) out[v] = src[v];
}

__global__ void b_zero_i32(int32_t* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---- batched kNN helpers ---------------------------------------------------
__global__ void b_localize(int64_t m, int n_inst, const int64_t* __restrict__ node_off,
                           const int64_t* __restrict__ rowptr, int32_t* __restrict__ nbr) {
  // nbr holds global node ids; rewrite as instance-local ids.  Slot s belongs
  // to the node whose global id is nbr's own instance: binary search node_off.
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  const int64_t w = nbr[s];
  int lo = 0, hi = n_inst;  // largest i with node_off[i] <= w
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (node_off[mid] <= w) lo = mid; else hi = mid;
  }
  nbr[s] = (int32_t)(w - node_off[lo]);
}

}  // namespace

// defined in f2m_knn.hip
f2m_status knn_directed_batched(f2m_ctx* c, int n_inst, const int64_t* node_off_h, const int64_t* node_off_d,
                                const double* x, const double* y, int K, int mode, int32_t* out_id,
                                double* out_c);
f2m_status symmetrise_i64(f2m_ctx* c, int64_t n, int K, const int32_t* knn, const double* kc, DevBuf& rpb,
                          DevBuf& nbb, DevBuf& cb, int64_t* nslots_out);

namespace {

struct BatchPlan {
  std::vector<int32_t> order;  // all instances, bucketed, decreasing size
  std::vector<int> bucket_start, bucket_lpn, bucket_bs, bucket_lds, bucket_grid;
  static constexpr int MAXBUCKETS = 16;  // queue counters (b_zero_i32 clears 16)
  std::vector<size_t> bucket_dyn;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  std::vector<unsigned char> key;
  // side streams for overlapping buckets (F2M_BATCH_OVERLAP), default priority
  static constexpr int MAXSIDE = 3;
  hipStream_t side[MAXSIDE] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[MAXSIDE] = {nullptr, nullptr, nullptr};
};

// The context keeps the two most recently used plans, most recent first, so a
// training loop that alternates warm-started and cold solves of one batch (two
// keys) replays both graphs instead of capturing one anew on every call.  The
// queue order lives in one device buffer (b_order): order_of names the plan
// whose order it holds.
struct BatchPlans {
  BatchPlan* slot[2] = {nullptr, nullptr};
  const BatchPlan* order_of = nullptr;
};

static void destroy_plan(BatchPlan* P) {
  if (!P) return;
  if (P->exec) (void)hipGraphExecDestroy(P->exec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  if (P->ev_fork) (void)hipEventDestroy(P->ev_fork);
  for (int i = 0; i < BatchPlan::MAXSIDE; ++i) {
    if (P->ev_join[i]) (void)hipEventDestroy(P->ev_join[i]);
    if (P->side[i]) (void)hipStreamDestroy(P->side[i]);
  }
  delete P;
}

// Schedule of an instance of n nodes: (lanes per node, workgroup threads).
// F2M_BATCH_LPN / F2M_BATCH_BS override; else the context's autotuned table
// (f2m_batch_autotune); else lanes 2 below 1000 nodes and 4 above, 256
// threads below 1000 nodes, 512 below 2000, 1024 above (F2M_BATCH_N512,
// F2M_BATCH_N1024) -- and when the whole batch fits the GPU at a larger
// workgroup (n_inst * threads <= CUs * 1024), the larger workgroup, since idle
// CUs would otherwise wait on the slowest instance.  With the buckets
// overlapped, larger workgroups for mid-size instances finish them sooner and
// pack better: the 4096-instance batch took 1.02 s at 4000 / 10000 nodes,
// 0.85 s at 2000 / 5000, 0.81-0.82 s from 2000 / 3000 down to 500 / 2000.
void batch_cfg_for(const f2m_ctx* c, int64_t n, int n_inst, int ncu, int& lpn, int& bs, int& lds) {
  lpn = n < 1000 ? 2 : 4;
  const int64_t n512 = env_int("F2M_BATCH_N512", 1000), n1024 = env_int("F2M_BATCH_N1024", 2000);
  bs = n < n512 ? 256 : (n < n1024 ? 512 : 1024);
  while (bs < 1024 && (int64_t)n_inst * bs * 2 <= (int64_t)ncu * 1024 && n >= 2 * bs) bs *= 2;
  lds = 3;
  for (const auto& t : c->b_tune)
    if (n <= t.max_n) {
      lpn = t.lpn;
      bs = t.bs;
      if (t.lds >= 0) lds = t.lds;
      break;
    }
  if (const char* e = std::getenv("F2M_BATCH_LPN")) lpn = std::atoi(e);
  if (const char* e = std::getenv("F2M_BATCH_BS")) bs = std::atoi(e);
  if (const char* e = std::getenv("F2M_BATCH_LDS")) lds = std::atoi(e);
  if (lpn != 1 && lpn != 2 && lpn != 8) lpn = 4;
  if (bs != 512 && bs != 1024) bs = 256;
  if (lds < 0 || lds > 3) lds = 3;
}

template <int BS, int LPN>
const void* bucket_kernel() {
  return reinterpret_cast<const void*>(&b_persistent<BS, LPN>);
}

const void* bucket_kernel_any(int bs, int lpn) {
  auto pick = [&](auto k1, auto k2, auto k4, auto k8) {
    return lpn == 1 ? k1 : (lpn == 2 ? k2 : (lpn == 8 ? k8 : k4));
  };
  if (bs == 1024)
    return pick(bucket_kernel<1024, 1>(), bucket_kernel<1024, 2>(), bucket_kernel<1024, 4>(),
                bucket_kernel<1024, 8>());
  if (bs == 512)
    return pick(bucket_kernel<512, 1>(), bucket_kernel<512, 2>(), bucket_kernel<512, 4>(), bucket_kernel<512, 8>());
  return pick(bucket_kernel<256, 1>(), bucket_kernel<256, 2>(), bucket_kernel<256, 4>(), bucket_kernel<256, 8>());
}

// Launch shape of a bucket whose largest instance has nmax nodes: dynamic
// shared memory for the potentials -- two buffers of nmax doubles if they
// fit beside the workgroups per CU that the registers allow, else one, else
// what is left (smaller instances of the bucket may still fit) -- at least
// the exact stage's 2*bs doubles; grid = co-resident workgroups, at most one
// per instance.  By default the workgroup gets all the shared memory its
// occupancy allows, so that the active lists can follow the potentials.
// lds (F2M_BATCH_LDS, or the autotuned table): 0 keeps the potentials in
// global memory, 1 allows one shared buffer only, 2 the potentials only, 3
// (default) the potentials and then the lists.
hipError_t bucket_shape(int device, int bs, int lpn, int lds, int64_t nmax, int n_order, int ncu, size_t& dyn,
                        int& grid) {
  const void* k = bucket_kernel_any(bs, lpn);
  hipFuncAttributes fa;
  hipError_t e = hipFuncGetAttributes(&fa, k);
  if (e != hipSuccess) return e;
  int lds_cu = 0;
  e = hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device);
  if (e != hipSuccess) return e;
  const size_t base = (size_t)bs * 2 * 8;
  int per_cu = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, bs, base);
  if (e != hipSuccess) return e;
  per_cu = std::max(1, per_cu);
  size_t room = (size_t)lds_cu / per_cu;
  room = room > fa.sharedSizeBytes + 64 ? (room - fa.sharedSizeBytes - 64) & ~(size_t)255 : 0;
  dyn = base;
  if (lds > 0) {
    const size_t two = (size_t)nmax * 16, one = (size_t)nmax * 8;
    if (lds >= 3) dyn = std::max(base, room);  // potentials, then the lists where they fit
    else dyn = std::max(base, (two <= room && lds >= 2) ? two : (one <= room ? one : room));
  }
  // The limit is a property of the kernel, shared by every bucket that
  // launches it (buckets of one kernel may differ in placement, so in dyn):
  // set it to what the kernel can have at all, never to one bucket's dyn,
  // which would leave a larger bucket of the same kernel above the limit.
  const size_t lim = (size_t)lds_cu > fa.sharedSizeBytes ? (size_t)lds_cu - fa.sharedSizeBytes : 0;
  if (dyn > lim) return hipErrorInvalidValue;
  e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim);
  if (e != hipSuccess) (void)hipGetLastError();  // the default limit may already allow it
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, bs, dyn);
  if (e != hipSuccess) return e;
  grid = std::min(n_order, std::max(1, per_cu) * ncu);
  return hipSuccess;
}

template <int BS, int LPN>
hipError_t launch_bucket(const BArgs& ak, int grid, size_t dyn, hipStream_t st) {
  hipLaunchKernelGGL((b_persistent<BS, LPN>), dim3(grid), dim3(BS), dyn, st, ak);
  return hipGetLastError();
}

template <int BS>
hipError_t launch_bucket_bs(const BArgs& ak, int lpn, int grid, size_t dyn, hipStream_t st) {
  switch (lpn) {
    case 1: return launch_bucket<BS, 1>(ak, grid, dyn, st);
    case 2: return launch_bucket<BS, 2>(ak, grid, dyn, st);
    case 8: return launch_bucket<BS, 8>(ak, grid, dyn, st);
    default: return launch_bucket<BS, 4>(ak, grid, dyn, st);
  }
}

hipError_t launch_bucket_any(const BArgs& ak, int bs, int lpn, int grid, size_t dyn, hipStream_t st) {
  if (bs == 1024) return launch_bucket_bs<1024>(ak, lpn, grid, dyn, st);
  if (bs == 512) return launch_bucket_bs<512>(ak, lpn, grid, dyn, st);
  return launch_bucket_bs<256>(ak, lpn, grid, dyn, st);
}

}  // namespace

}  // namespace f2m

namespace f2m {
void release_batch_plan(f2m_ctx* c) {
  BatchPlans* Q = static_cast<BatchPlans*>(c->b_plan);
  if (!Q) return;
  destroy_plan(Q->slot[0]);
  destroy_plan(Q->slot[1]);
  delete Q;
  c->b_plan = nullptr;
}
}  // namespace f2m

using namespace f2m;

extern "C" {

f2m_status f2m_knn_build_batched(f2m_ctx* c, int32_t n_inst, const int64_t* node_off, const double* x,
                                 const double* y, int32_t k, int32_t cost_mode, f2m_batch* out) {
  if (!c || n_inst < 1 || !node_off || !x || !y || !out || k < 1 || cost_mode < 0 || cost_mode > 2)
    return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  if (k > 80) return fail(c, F2M_ERR_LIMIT, "k > 80");
  const int64_t n_total = node_off[n_inst];
  for (int i = 0; i < n_inst; ++i) {
    const int64_t ni = node_off[i + 1] - node_off[i];
    if (ni < 2 || ni - 1 < k) return fail(c, F2M_ERR_LIMIT, "every instance needs more than k points");
    if (ni > (1 << 30)) return fail(c, F2M_ERR_LIMIT, "instance too large");
  }
  const double* dx = (const double*)stage_in(c, x, n_total * 8ull, 0);
  const double* dy = (const double*)stage_in(c, y, n_total * 8ull, 1);
  if (!dx || !dy) return F2M_ERR_HIP;
  F2M_HIP(c, c->b_knn.reserve(n_total * k * 4ull));
  F2M_HIP(c, c->b_kc.reserve(n_total * k * 8ull));
  int32_t* knn = c->b_knn.as<int32_t>();
  double* kc = c->b_kc.as<double>();
  F2M_HIP(c, c->b_node_off.reserve((n_inst + 1) * 8ull));
  F2M_HIP(c, hipMemcpyAsync(c->b_node_off.p, node_off, (n_inst + 1) * 8ull, hipMemcpyHostToDevice, c->stream));
  f2m_status s = knn_directed_batched(c, n_inst, node_off, c->b_node_off.as<int64_t>(), dx, dy, k, cost_mode, knn, kc);
  if (s) return s;
  int64_t m = 0;
  s = symmetrise_i64(c, n_total, k, knn, kc, c->b_rowptr, c->b_nbr, c->b_cost, &m);
  if (s) return s;
  hipLaunchKernelGGL(b_localize, dim3(grid_for(m, 256)), dim3(256), 0, c->stream, m, n_inst,
                     c->b_node_off.as<int64_t>(), c->b_rowptr.as<int64_t>(), c->b_nbr.as<int32_t>());
  F2M_LAUNCH_CHECK(c);
  F2M_HIP(c, hipStreamSynchronize(c->stream));
  c->b_n_inst = n_inst;
  c->b_n_total = n_total;
  c->b_nslots = m;
  out->n_inst = n_inst;
  out->k = k;
  out->cost_mode = cost_mode;
  out->reserved = 0;
  out->n_total = n_total;
  out->nslots = m;
  out->node_off = c->b_node_off.as<int64_t>();
  out->rowptr = c->b_rowptr.as<int64_t>();
  out->nbr = c->b_nbr.as<int32_t>();
  out->cost = c->b_cost.as<double>();
  return F2M_OK;
}

static f2m_status solve_batched_impl(f2m_ctx* c, const f2m_batch* b, const f2m_opts* o, const double* F_init,
                                     f2m_stats* st);

f2m_status f2m_solve_batched(f2m_ctx* c, const f2m_batch* b, const f2m_opts* o, f2m_stats* st) {
  return solve_batched_impl(c, b, o, nullptr, st);
}

f2m_status f2m_solve_batched_warm(f2m_ctx* c, const f2m_batch* b, const f2m_opts* o, const double* F_init,
                                  f2m_stats* st) {
  if (!c || !b || b->n_inst < 1 || !F_init) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  // copied into a context buffer: a stable address keeps the captured graph
  F2M_HIP(c, c->b_finit.reserve((size_t)b->n_total * 8));
  F2M_HIP(c, hipMemcpyAsync(c->b_finit.p, F_init, (size_t)b->n_total * 8, hipMemcpyDefault, c->stream));
  return solve_batched_impl(c, b, o, c->b_finit.as<double>(), st);
}

static f2m_status solve_batched_impl(f2m_ctx* c, const f2m_batch* b, const f2m_opts* o, const double* F_init,
                                     f2m_stats* st) {
  if (!c || !b || b->n_inst < 1) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  f2m_opts oo;
  if (o) oo = *o; else f2m_default_opts(&oo);
  const int64_t N = b->n_total, M = b->nslots;
  const int n_inst = b->n_inst;
  SolveWs& w = c->bws;
  F2M_HIP(c, w.nbr.reserve(M * 4));
  F2M_HIP(c, w.D.reserve(M * 8));
  F2M_HIP(c, w.orig.reserve(M * 4));
  F2M_HIP(c, w.xw.reserve(M * 8));
  F2M_HIP(c, w.xc.reserve(M * 8));
  F2M_HIP(c, w.axb.reserve(M * 8 + N * 8));
  F2M_HIP(c, w.sortk.reserve(M * 2));
  F2M_HIP(c, w.F0.reserve(N * 8));
  F2M_HIP(c, w.F1.reserve(N * 8));
  F2M_HIP(c, w.act.reserve(N * 4));
  F2M_HIP(c, w.clen.reserve(N * 4));
  F2M_HIP(c, w.deg.reserve(N * 4));
  F2M_HIP(c, w.A.reserve(N * 4));
  F2M_HIP(c, w.T.reserve(N * 4));
  F2M_HIP(c, w.root.reserve(N * 4));
  F2M_HIP(c, w.irr.reserve(N * 4));
  F2M_HIP(c, w.sortv.reserve(N * 8));  // aux + list
  F2M_HIP(c, w.frozen.reserve(N));
  F2M_HIP(c, c->b_ef.reserve(M * 16));
  F2M_HIP(c, c->b_efoff.reserve((N + n_inst + 1) * 4));
  F2M_HIP(c, c->b_stats.reserve(n_inst * sizeof(f2m_stats) + 64 * 4));
  F2M_HIP(c, c->b_order.reserve((n_inst + 64) * 4ull));
  F2M_HIP(c, c->b_cur.reserve((n_inst + 64) * 4ull));
  if (oo.exact > 0) {
    F2M_HIP(c, w.xmir.reserve(M * 4));
    F2M_HIP(c, w.xf.reserve(M));
    F2M_HIP(c, w.xp.reserve(N * 8 * 4));
    F2M_HIP(c, w.xe.reserve(N * 4 * 2));
    F2M_HIP(c, w.xact.reserve(N * 4 * 2));
  }
  if (c->b_clk_on) F2M_HIP(c, c->b_clk.reserve((n_inst + 64) * 8ull));

  int ncu = 0;
  F2M_HIP(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
  // host copy of node offsets -> size buckets -> queue order
  std::vector<int64_t> off(n_inst + 1);
  F2M_HIP(c, hipMemcpyAsync(off.data(), b->node_off, (n_inst + 1) * 8ull, hipMemcpyDefault, c->stream));
  F2M_HIP(c, hipStreamSynchronize(c->stream));
  // key: batch identity + options; a new key rebuilds the captured graph
  std::vector<unsigned char> key(sizeof(f2m_batch) + sizeof(f2m_opts));
  std::memcpy(key.data(), b, sizeof(f2m_batch));
  std::memcpy(key.data() + sizeof(f2m_batch), &oo, sizeof(f2m_opts));
  const char* lpn_env = std::getenv("F2M_BATCH_LPN");
  if (lpn_env) key.insert(key.end(), lpn_env, lpn_env + std::strlen(lpn_env));
  for (const auto& t : c->b_tune) {
    const unsigned char* q = reinterpret_cast<const unsigned char*>(&t);
    key.insert(key.end(), q, q + sizeof t);
  }
  if (const char* e = std::getenv("F2M_BATCH_BS")) key.insert(key.end(), e, e + std::strlen(e));
  for (const char* nm : {"F2M_BATCH_N512", "F2M_BATCH_N1024"})
    if (const char* e = std::getenv(nm)) key.insert(key.end(), e, e + std::strlen(e));
  if (const char* e = std::getenv("F2M_BATCH_LDS")) key.insert(key.end(), e, e + std::strlen(e));
  if (const char* e = std::getenv("F2M_BATCH_OVERLAP")) key.insert(key.end(), e, e + std::strlen(e));
  if (const char* e = std::getenv("F2M_BATCH_REGS")) key.insert(key.end(), e, e + std::strlen(e));
  key.push_back(c->b_clk_on ? 1 : 0);
  {  // the warm-start source and every workspace are captured kernel arguments
    auto addp = [&](const void* p) {
      const unsigned char* q = reinterpret_cast<const unsigned char*>(&p);
      key.insert(key.end(), q, q + sizeof p);
    };
    addp(F_init);
    for (const DevBuf* d : {&w.nbr, &w.D, &w.orig, &w.F0, &w.F1, &w.act, &w.deg, &w.clen, &w.A, &w.T, &w.root,
                            &w.irr, &w.sortv, &w.frozen, &w.xw, &w.xc, &w.axb, &w.sortk, &w.xmir, &w.xf, &w.xp,
                            &w.xe, &w.xact, &c->b_ef, &c->b_efoff, &c->b_stats, &c->b_order, &c->b_cur, &c->b_clk})
      addp(d->p);
  }
  if (!c->b_plan) c->b_plan = new BatchPlans;
  BatchPlans* Q = static_cast<BatchPlans*>(c->b_plan);
  if (Q->slot[1] && Q->slot[1]->key == key && !(Q->slot[0] && Q->slot[0]->key == key))
    std::swap(Q->slot[0], Q->slot[1]);
  if (!Q->slot[0] || Q->slot[0]->key != key) {
    if (Q->slot[1] == Q->order_of) Q->order_of = nullptr;
    destroy_plan(Q->slot[1]);
    Q->slot[1] = Q->slot[0];
    Q->slot[0] = new BatchPlan;
  }
  BatchPlan* P = Q->slot[0];
  const bool rebuild = P->key != key || !P->exec;
  if (rebuild) {
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    P->exec = nullptr;
    P->graph = nullptr;
    P->key = key;
    if (Q->order_of == P) Q->order_of = nullptr;
    std::vector<int32_t> ids(n_inst);
    std::iota(ids.begin(), ids.end(), 0);
    std::stable_sort(ids.begin(), ids.end(), [&](int a, int b2) {
      return off[a + 1] - off[a] > off[b2 + 1] - off[b2];
    });
    // buckets by lanes-per-node schedule (autotuned choice per size class)
    P->order = ids;
    P->bucket_start.clear();
    P->bucket_lpn.clear();
    P->bucket_bs.clear();
    P->bucket_lds.clear();
    int prev = -1;
    for (int i = 0; i < n_inst; ++i) {
      int l, bs, lds;
      batch_cfg_for(c, off[ids[i] + 1] - off[ids[i]], n_inst, ncu, l, bs, lds);
      const int kind = (l * 4096 + bs) * 4 + lds;
      if (kind != prev) {
        P->bucket_start.push_back(i);
        P->bucket_lpn.push_back(l);
        P->bucket_bs.push_back(bs);
        P->bucket_lds.push_back(lds);
        prev = kind;
      }
    }
    if (P->bucket_lpn.size() > (size_t)BatchPlan::MAXBUCKETS)
      return fail(c, F2M_ERR_LIMIT, "the schedule table splits the batch into more than 16 buckets");
    P->bucket_start.push_back(n_inst);
    P->bucket_dyn.clear();
    P->bucket_grid.clear();
    for (size_t k = 0; k + 1 < P->bucket_start.size(); ++k) {
      const int first = P->order[P->bucket_start[k]];  // the bucket's largest instance
      size_t dyn = 0;
      int grid = 0;
      F2M_HIP(c, bucket_shape(c->device, P->bucket_bs[k], P->bucket_lpn[k], P->bucket_lds[k],
                              off[first + 1] - off[first], P->bucket_start[k + 1] - P->bucket_start[k], ncu, dyn,
                              grid));
      P->bucket_dyn.push_back(dyn);
      P->bucket_grid.push_back(grid);
    }
  }
  if (Q->order_of != P) {
    F2M_HIP(c, hipMemcpyAsync(c->b_order.p, P->order.data(), n_inst * 4ull, hipMemcpyHostToDevice, c->stream));
    F2M_HIP(c, hipStreamSynchronize(c->stream));
    Q->order_of = P;
  }
  f2m_stats* dstats = c->b_stats.as<f2m_stats>();
  int32_t* queues = reinterpret_cast<int32_t*>(dstats + n_inst);
  int32_t* curs = c->b_cur.as<int32_t>();

  BArgs a;
  a.n_inst = n_inst;
  a.node_off = b->node_off;
  a.rowptr = b->rowptr;
  a.gnbr = b->nbr;
  a.gcost = b->cost;
  a.nbr = w.nbr.as<int32_t>();
  a.D = w.D.as<double>();
  a.orig = w.orig.as<int32_t>();
  a.F0 = w.F0.as<double>();
  a.F1 = w.F1.as<double>();
  a.act = w.act.as<int32_t>();
  a.deg = w.deg.as<int32_t>();
  a.clen = w.clen.as<int32_t>();
  a.A = w.A.as<int32_t>();
  a.T = w.T.as<int32_t>();
  a.root = w.root.as<int32_t>();
  a.irr = w.irr.as<int32_t>();
  a.aux = w.sortv.as<int32_t>();
  a.list = a.aux + N;
  a.frozen = w.frozen.as<uint8_t>();
  a.xw = w.xw.as<double>();
  a.xc = w.xc.as<double>();
  a.xg = w.axb.as<double>();
  a.rn = a.xg + M;
  a.fw = w.sortk.as<uint8_t>();
  a.fc = a.fw + M;
  a.efoff = c->b_efoff.as<int32_t>();
  a.ef = c->b_ef.as<ESlot>();
  a.fcap = 0;  // set per bucket
  // 2: with the owners' halves from registers (C5's forward 19.8-20.1 -> 19.6 ms)
  a.sweep_regs = std::max(0, std::min(2, env_int("F2M_BATCH_REGS", 2)));
  a.part_lanes = env_int("F2M_BATCH_PART_LANES", 1) != 0;
  a.o = oo;
  a.stats = dstats;
  a.cur_out = curs;
  a.n_total = N;
  a.F_init = F_init;
  a.xmir = w.xmir.as<int32_t>();
  a.xf = w.xf.as<uint8_t>();
  a.xp = w.xp.as<double>();
  a.xe = w.xe.as<int32_t>();
  a.xact = w.xact.as<int32_t>();
  a.clk = nullptr;
  if (c->b_clk_on) a.clk = c->b_clk.as<long long>();

  // The buckets after the first run on up to 3 more streams (bucket k on side
  // stream (k-1) mod s; F2M_BATCH_OVERLAP=s, 0 = one stream), forked after the
  // queue reset: the smaller instances take CUs as the larger ones' workgroups
  // finish instead of waiting for the slowest of them (4096 instances of n = 100..20000: 1.31 -> 1.02 s)
  const int nside = std::min(std::max(0, env_int("F2M_BATCH_OVERLAP", BatchPlan::MAXSIDE)),
                             std::min<int>(BatchPlan::MAXSIDE, (int)P->bucket_lpn.size() - 1));
  const bool overlap = nside > 0;
  if (overlap && !P->ev_fork) F2M_HIP(c, hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming));
  for (int i = 0; i < nside; ++i)
    if (!P->side[i]) {
      F2M_HIP(c, hipStreamCreateWithFlags(&P->side[i], hipStreamNonBlocking));
      F2M_HIP(c, hipEventCreateWithFlags(&P->ev_join[i], hipEventDisableTiming));
    }
  auto enqueue = [&]() -> f2m_status {
    const int nb = (int)P->bucket_lpn.size();
    hipLaunchKernelGGL(b_zero_i32, dim3(1), dim3(64), 0, c->stream, queues, 16);
    F2M_LAUNCH_CHECK(c);
    if (overlap) {
      F2M_HIP(c, hipEventRecord(P->ev_fork, c->stream));
      for (int i = 0; i < nside; ++i) F2M_HIP(c, hipStreamWaitEvent(P->side[i], P->ev_fork, 0));
    }
    for (int k = 0; k < nb; ++k) {
      BArgs ak = a;
      ak.order = c->b_order.as<int32_t>() + P->bucket_start[k];
      ak.n_order = P->bucket_start[k + 1] - P->bucket_start[k];
      ak.queue = queues + k;
      ak.fcap = (int32_t)(P->bucket_dyn[k] / 8);
      F2M_HIP(c, launch_bucket_any(ak, P->bucket_bs[k], P->bucket_lpn[k], P->bucket_grid[k], P->bucket_dyn[k],
                                   overlap && k > 0 ? P->side[(k - 1) % nside] : c->stream));
    }
    for (int i = 0; i < nside; ++i) {
      F2M_HIP(c, hipEventRecord(P->ev_join[i], P->side[i]));
      F2M_HIP(c, hipStreamWaitEvent(c->stream, P->ev_join[i], 0));
    }
    return F2M_OK;
  };
  const bool use_graph = std::getenv("F2M_NO_GRAPH") == nullptr;
  if (use_graph) {
    if (rebuild) {
      F2M_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      f2m_status s = enqueue();
      hipGraph_t gph = nullptr;
      hipError_t e = hipStreamEndCapture(c->stream, &gph);
      if (s) return s;
      F2M_HIP(c, e);
      P->graph = gph;
      F2M_HIP(c, hipGraphInstantiate(&P->exec, gph, nullptr, nullptr, 0));
    }
    F2M_HIP(c, hipGraphLaunch(P->exec, c->stream));
  } else {
    f2m_status s = enqueue();
    if (s) return s;
  }
  F2M_HIP(c, hipStreamSynchronize(c->stream));
  if (st) F2M_HIP(c, hipMemcpy(st, dstats, n_inst * sizeof(f2m_stats), hipMemcpyDeviceToHost));
  return F2M_OK;
}

f2m_status f2m_batch_result_copy(f2m_ctx* c, double* F, double* x, double* objective) {
  if (!c || c->b_n_inst <= 0) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  SolveWs& w = c->bws;
  const int n_inst = c->b_n_inst;
  if (F) {
    // per instance the current buffer may be F0 or F1: one gather launch, into
    // F itself when it is device memory, else into a device buffer then copied
    hipPointerAttribute_t pa;
    const bool dev = hipPointerGetAttributes(&pa, F) == hipSuccess && pa.type == hipMemoryTypeDevice &&
                     pa.device == c->device;
    (void)hipGetLastError();  // a host pointer may leave an error behind
    double* dst = F;
    if (!dev) {
      F2M_HIP(c, c->b_fout.reserve((size_t)c->b_n_total * 8));
      dst = c->b_fout.as<double>();
    }
    hipLaunchKernelGGL(b_gather_F, dim3(n_inst), dim3(256), 0, c->stream, c->b_node_off.as<int64_t>(),
                       (const int32_t*)c->b_cur.as<int32_t>(), (const double*)w.F0.as<double>(),
                       (const double*)w.F1.as<double>(), dst);
    F2M_LAUNCH_CHECK(c);
    if (!dev) F2M_HIP(c, copy_out(c, F, dst, (size_t)c->b_n_total * 8));
    else F2M_HIP(c, hipStreamSynchronize(c->stream));
  }
  F2M_HIP(c, copy_out(c, x, w.xc.p, c->b_nslots * 8ull));
  if (objective) {
    std::vector<f2m_stats> st(n_inst);
    F2M_HIP(c, hipMemcpy(st.data(), c->b_stats.p, n_inst * sizeof(f2m_stats), hipMemcpyDeviceToHost));
    for (int i = 0; i < n_inst; ++i) objective[i] = st[i].objective;
  }
  return F2M_OK;
}

f2m_status f2m_batch_comb_root_copy(f2m_ctx* c, int32_t* comb_root) {
  if (!c || c->b_n_inst <= 0 || !comb_root) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  F2M_HIP(c, copy_out(c, comb_root, c->bws.root.p, c->b_n_total * 4ull));
  return F2M_OK;
}

// Autotuner: picks lanes per node for each instance-size class by timing
// every candidate on the first `sample` instances of the batch (each
// workgroup records its instance's wall_clock64 ticks) and keeping, per
// class, the candidate with the smallest summed time.  All candidates give
// bit-identical results; only speed differs.  The table is kept on the
// context and, when cache_path is given, written there and reused on later
// calls (one line per class: "<max instance size> <lanes per node>
// <threads> <placement>", placement -1 = the default; a table written by
// hand or by bench.py --tune-batch may set it, as F2M_BATCH_LDS).
f2m_status f2m_batch_autotune(f2m_ctx* c, const f2m_batch* b, const f2m_opts* o, int32_t sample,
                              const char* cache_path) {
  if (!c || !b || b->n_inst < 1) return F2M_ERR_ARG;
  static const int64_t bounds[] = {500, 1000, 2000, 4000, 8000, 16000, INT64_MAX};
  constexpr int NC = sizeof(bounds) / sizeof(bounds[0]);
  static const int cand_lpn[] = {1, 2, 4, 8};
  static const int cand_bs[] = {256, 512, 1024};
  using TE = f2m_ctx::TuneEntry;
  if (cache_path) {
    FILE* f = std::fopen(cache_path, "r");
    if (f) {
      std::vector<TE> t;
      long long mx;
      int l, bs, lds;
      while (std::fscanf(f, "%lld %d %d %d", &mx, &l, &bs, &lds) == 4)
        t.push_back(TE{(int64_t)mx, (int32_t)l, (int32_t)bs, (int32_t)lds, 0});
      std::fclose(f);
      if (!t.empty()) {
        c->b_tune = t;
        return F2M_OK;
      }
    }
  }
  const int ns = std::max(1, std::min<int>(sample, b->n_inst));
  f2m_batch sub = *b;
  sub.n_inst = ns;  // node_off / rowptr hold absolute offsets: a prefix is a batch
  std::vector<int64_t> off(ns + 1);
  F2M_HIP(c, hipMemcpy(off.data(), b->node_off, (ns + 1) * 8ull, hipMemcpyDefault));
  double best_t[NC];
  TE best[NC];
  for (int k = 0; k < NC; ++k) { best_t[k] = -1; best[k] = TE{bounds[k], 0, 0, -1, 0}; }
  const auto saved = c->b_tune;
  c->b_clk_on = true;
  std::vector<long long> clk(ns);
  f2m_status st = F2M_OK;
  for (int bs : cand_bs)
    for (int l : cand_lpn) {
      c->b_tune = {TE{INT64_MAX, l, bs, -1, 0}};
      st = f2m_solve_batched(c, &sub, o, nullptr);
      if (st) break;
      if (hipMemcpy(clk.data(), c->b_clk.p, ns * 8ull, hipMemcpyDeviceToHost) != hipSuccess) { st = F2M_ERR_HIP; break; }
      // a workgroup of bs threads holds bs/256 times the resources of the
      // smallest: weigh its instance time accordingly
      double tc[NC] = {0};
      int cnt[NC] = {0};
      for (int i = 0; i < ns; ++i) {
        const int64_t n = off[i + 1] - off[i];
        int k = 0;
        while (n > bounds[k]) ++k;
        tc[k] += (double)clk[i] * (bs / 256);
        ++cnt[k];
      }
      for (int k = 0; k < NC; ++k)
        if (cnt[k] && (best_t[k] < 0 || tc[k] < best_t[k])) { best_t[k] = tc[k]; best[k] = TE{bounds[k], l, bs, -1, 0}; }
    }
  c->b_clk_on = false;
  if (st) { c->b_tune = saved; return st; }
  // classes without samples take the nearest sampled class below, else above
  std::vector<TE> t;
  TE last{0, 0, 0, -1, 0};
  for (int k = 0; k < NC; ++k) if (best[k].lpn) { last = best[k]; break; }
  for (int k = 0; k < NC; ++k) {
    if (best[k].lpn) last = best[k];
    t.push_back(TE{bounds[k], last.lpn ? last.lpn : 4, last.bs ? last.bs : 256, -1, 0});
  }
  c->b_tune = t;
  if (cache_path) {
    FILE* f = std::fopen(cache_path, "w");
    if (f) {
      for (auto& e : t) std::fprintf(f, "%lld %d %d %d\n", (long long)e.max_n, e.lpn, e.bs, e.lds);
      std::fclose(f);
    }
  }
  return F2M_OK;
}

f2m_status f2m_batch_copy(f2m_ctx* c, const f2m_batch* b, int64_t* node_off, int64_t* rowptr, int32_t* nbr,
                          double* cost) {
  if (!c || !b) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  F2M_HIP(c, copy_out(c, node_off, b->node_off, (b->n_inst + 1) * 8ull));
  F2M_HIP(c, copy_out(c, rowptr, b->rowptr, (b->n_total + 1) * 8ull));
  F2M_HIP(c, copy_out(c, nbr, b->nbr, b->nslots * 4ull));
  F2M_HIP(c, copy_out(c, cost, b->cost, b->nslots * 8ull));
  return F2M_OK;
}

}  // extern "C"
