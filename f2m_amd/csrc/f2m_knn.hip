// f2m_knn.hip -- init: grid-bucketed kNN edge construction on the GPU.
//
// Replaces the reference's O(n^2)-per-pass radius search and candidate sort
// (cuda_D2_stat / cuda_D2_fill / thrust::sort, FXY.hpp:58-144, 313-534) and
// the host greedy selection (graph_tour.hpp:28-98).  The selected edge set is
// the same: the union over v of the K nearest neighbours of v, where "nearest"
// orders candidates by the key (cost, id) -- the reference's global
// (d, min id, max id) candidate order restricted to one node's edges.
//
// Pipeline (all on the context stream):
//   bbox -> cell id + count -> scan -> scatter into cell order
//   -> per-point ring search, best-K list in LDS  (k_knn)
//   -> reverse-edge count -> scan -> fill -> per-node sort by (cost, nbr).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "f2m_internal.h"

namespace f2m {

// Edge cost (FXY.hpp:376-383).  -ffp-contract=off keeps dx*dx + dy*dy as two
// roundings, identical to the host; gfx950 f64 sqrt and division lower to
// correctly rounded sequences.
__device__ __forceinline__ double edge_cost(double xu, double yu, double xw, double yw, int mode) {
  double dx = xu - xw, dy = yu - yw;
  double s = dx * dx + dy * dy;
  double c = sqrt(s);
  if (mode == F2M_COST_NINT) return floor(c + 0.5);
  if (mode == F2M_COST_RAW_FP32) return (double)(float)c;
  return c;
}
__device__ __forceinline__ double round_like_cost(double dist, int mode) {
  if (mode == F2M_COST_NINT) return floor(dist + 0.5);
  if (mode == F2M_COST_RAW_FP32) return (double)(float)dist;
  return dist;
}

__device__ __forceinline__ unsigned long long ordered_bits(double v) {
  unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double from_ordered_bits(unsigned long long b) {
  b = (b >> 63) ? (b & 0x7fffffffffffffffull) : ~b;
  double v;
  memcpy(&v, &b, 8);
  return v;
}

// bbox as ordered bits: [min x, min y, max x, max y]; min via max of ~bits.
__global__ void k_bbox(const double* __restrict__ x, const double* __restrict__ y, int64_t n,
                       unsigned long long* __restrict__ out) {
  unsigned long long mnx = 0, mny = 0, mxx = 0, mxy = 0;  // all as "max" keys
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    unsigned long long bx = ordered_bits(x[i]), by = ordered_bits(y[i]);
    mnx = max(mnx, ~bx); mny = max(mny, ~by);
    mxx = max(mxx, bx); mxy = max(mxy, by);
  }
  for (int off = 32; off > 0; off >>= 1) {
    mnx = max(mnx, (unsigned long long)__shfl_xor((long long)mnx, off, 64));
    mny = max(mny, (unsigned long long)__shfl_xor((long long)mny, off, 64));
    mxx = max(mxx, (unsigned long long)__shfl_xor((long long)mxx, off, 64));
    mxy = max(mxy, (unsigned long long)__shfl_xor((long long)mxy, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&out[0], mnx); atomicMax(&out[1], mny);
    atomicMax(&out[2], mxx); atomicMax(&out[3], mxy);
  }
}

__global__ void k_zero_u64(unsigned long long* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void k_zero_i32(int32_t* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

struct GridGeom {
  double mnx, mny, cs;
  int64_t gx, gy;
};

__global__ void k_cell(const double* __restrict__ x, const double* __restrict__ y, int64_t n, GridGeom g,
                       int32_t* __restrict__ cell, int32_t* __restrict__ cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t cx = (int64_t)((x[i] - g.mnx) / g.cs);
  int64_t cy = (int64_t)((y[i] - g.mny) / g.cs);
  cx = cx < 0 ? 0 : (cx >= g.gx ? g.gx - 1 : cx);
  cy = cy < 0 ? 0 : (cy >= g.gy ? g.gy - 1 : cy);
  int32_t c = (int32_t)(cy * g.gx + cx);
  cell[i] = c;
  atomicAdd(&cnt[c], 1);
}

__global__ void k_scatter(const double* __restrict__ x, const double* __restrict__ y, int64_t n,
                          const int32_t* __restrict__ cell, const int32_t* __restrict__ cstart,
                          int32_t* __restrict__ cur, int32_t* __restrict__ order, double* __restrict__ sx,
                          double* __restrict__ sy) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t c = cell[i];
  int32_t p = cstart[c] + atomicAdd(&cur[c], 1);
  order[p] = (int32_t)i;
  sx[p] = x[i];
  sy[p] = y[i];
}

// Ring search of one point over the cells of its grid: the K nearest other
// points by the key (cost, id), ascending, as (bw, bc) columns of this
// thread (stride T) in shared memory.  cstart: cell starts of this grid
// (cells row-major, gx per row); order/sx/sy: points in cell order.  The
// search stops once every unvisited point lies beyond the K-th cost, so the
// result does not depend on the cell size.
template <int T>
__device__ __forceinline__ void knn_search(int K, int mode, const GridGeom& g, const int32_t* __restrict__ cstart,
                                           const int32_t* __restrict__ order, const double* __restrict__ sx,
                                           const double* __restrict__ sy, int32_t v, double xv, double yv,
                                           int64_t cell_v, double* bc, int32_t* bw) {
  const int t = threadIdx.x;
  const int64_t cx = cell_v % g.gx, cy = cell_v / g.gx;
  int cnt = 0;
  const int64_t Rmax = g.gx > g.gy ? g.gx : g.gy;
  for (int64_t R = 0;; ++R) {
    for (int64_t yy = cy - R; yy <= cy + R; ++yy) {
      if (yy < 0 || yy >= g.gy) continue;
      const bool edge_row = (yy == cy - R || yy == cy + R);
      const int64_t step = (edge_row || R == 0) ? 1 : 2 * R;
      for (int64_t xx = cx - R; xx <= cx + R; xx += step) {
        if (xx < 0 || xx >= g.gx) continue;
        const int64_t c = yy * g.gx + xx;
        const int32_t e = cstart[c + 1];
        for (int32_t q = cstart[c]; q < e; ++q) {
          const int32_t u = order[q];
          if (u == v) continue;
          const double cu = edge_cost(xv, yv, sx[q], sy[q], mode);
          if (cnt == K) {
            const double lc = bc[(K - 1) * T + t];
            const int32_t lw = bw[(K - 1) * T + t];
            if (!(cu < lc || (cu == lc && u < lw))) continue;
          }
          int i = cnt < K ? cnt++ : K - 1;
          while (i > 0) {
            const double pc = bc[(i - 1) * T + t];
            const int32_t pw = bw[(i - 1) * T + t];
            if (!(cu < pc || (cu == pc && u < pw))) break;
            bc[i * T + t] = pc;
            bw[i * T + t] = pw;
            --i;
          }
          bc[i * T + t] = cu;
          bw[i * T + t] = u;
        }
      }
    }
    if (R > Rmax) break;
    if (cnt == K) {
      // every unvisited point is farther than R*cs from v
      const double lb = round_like_cost((double)R * g.cs * (1.0 - 1e-9), mode);
      if (bc[(K - 1) * T + t] < lb) break;
    }
  }
}

// One thread per point, visited in cell order so a wave scans nearby cells.
// best-K list (cost, id) sorted ascending in LDS, column-major per thread.
template <int T>
__global__ void __launch_bounds__(T) k_knn(int64_t n, int K, int mode, GridGeom g,
                                           const int32_t* __restrict__ cstart,
                                           const int32_t* __restrict__ order,
                                           const double* __restrict__ sx, const double* __restrict__ sy,
                                           const int32_t* __restrict__ cell,
                                           int32_t* __restrict__ out_id, double* __restrict__ out_c,
                                           int32_t id_offset) {
  extern __shared__ unsigned char smem[];
  double* bc = reinterpret_cast<double*>(smem);
  int32_t* bw = reinterpret_cast<int32_t*>(bc + (size_t)K * T);
  int64_t p = (int64_t)blockIdx.x * T + threadIdx.x;
  if (p >= n) return;
  const int t = threadIdx.x;
  const int32_t v = order[p];
  knn_search<T>(K, mode, g, cstart, order, sx, sy, v, sx[p], sy[p], cell[v], bc, bw);
  for (int j = 0; j < K; ++j) {
    out_id[(int64_t)v * K + j] = bw[j * T + t] + id_offset;
    out_c[(int64_t)v * K + j] = bc[j * T + t];
  }
}

// ---- batched form: every instance of a batch in one pipeline ----------------
// Points carry global ids (instance i = [node_off[i], node_off[i+1])); each
// instance has its own grid, whose cells sit at cell_off[i] (at most 4n+16
// cells per instance, so the offsets are known on the host).

__device__ __forceinline__ int inst_of(int64_t p, const int64_t* __restrict__ node_off, int n_inst) {
  int lo = 0, hi = n_inst;  // largest i with node_off[i] <= p
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (node_off[mid] <= p) lo = mid; else hi = mid;
  }
  return lo;
}

// one workgroup per instance: bbox as ordered bits [min x, min y, max x, max y]
__global__ void __launch_bounds__(256) k_bbox_b(const double* __restrict__ x, const double* __restrict__ y,
                                                const int64_t* __restrict__ node_off,
                                                unsigned long long* __restrict__ out) {
  __shared__ unsigned long long sm[4][4];
  const int i = blockIdx.x;
  unsigned long long m[4] = {0, 0, 0, 0};  // all as "max" keys
  for (int64_t p = node_off[i] + threadIdx.x; p < node_off[i + 1]; p += 256) {
    const unsigned long long bx = ordered_bits(x[p]), by = ordered_bits(y[p]);
    m[0] = max(m[0], ~bx);
    m[1] = max(m[1], ~by);
    m[2] = max(m[2], bx);
    m[3] = max(m[3], by);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    for (int off = 32; off > 0; off >>= 1)
      m[k] = max(m[k], (unsigned long long)__shfl_xor((long long)m[k], off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 4; ++k) sm[wid][k] = m[k];
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long r = sm[0][threadIdx.x];
    for (int w = 1; w < 4; ++w) r = max(r, sm[w][threadIdx.x]);
    out[4 * (int64_t)i + threadIdx.x] = r;
  }
}

// Grid of n points in a w x h box: square cells of about one point each,
// enlarged until at most 4n+16 cells (also the single-instance rule).
__host__ __device__ inline GridGeom grid_geom(double mnx, double mny, double mxx, double mxy, int64_t n) {
  GridGeom g;
  g.mnx = mnx;
  g.mny = mny;
  const double w = mxx - mnx, h = mxy - mny;
  double cs = sqrt(fmax(w * h, 1e-300) / (double)n);
  cs = fmax(cs, fmax(w, h) / (double)n);
  if (!(cs > 0 && cs <= DBL_MAX)) cs = 1.0;  // also catches inf and NaN
  int64_t gx = (int64_t)(w / cs) + 1, gy = (int64_t)(h / cs) + 1;
  while (gx * gy > 4 * n + 16) {
    cs *= 1.5;
    gx = (int64_t)(w / cs) + 1;
    gy = (int64_t)(h / cs) + 1;
  }
  g.cs = cs;
  g.gx = gx;
  g.gy = gy;
  return g;
}

__global__ void k_geom_b(int n_inst, const int64_t* __restrict__ node_off,
                         const unsigned long long* __restrict__ bb, GridGeom* __restrict__ geom) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inst) return;
  const unsigned long long* b = bb + 4 * (int64_t)i;
  geom[i] = grid_geom(from_ordered_bits(~b[0]), from_ordered_bits(~b[1]), from_ordered_bits(b[2]),
                      from_ordered_bits(b[3]), node_off[i + 1] - node_off[i]);
}

__global__ void k_cell_b(int64_t n, int n_inst, const int64_t* __restrict__ node_off,
                         const int64_t* __restrict__ cell_off, const GridGeom* __restrict__ geom,
                         const double* __restrict__ x, const double* __restrict__ y, int32_t* __restrict__ cell,
                         int32_t* __restrict__ cnt) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i = inst_of(p, node_off, n_inst);
  const GridGeom g = geom[i];
  int64_t cx = (int64_t)((x[p] - g.mnx) / g.cs);
  int64_t cy = (int64_t)((y[p] - g.mny) / g.cs);
  cx = cx < 0 ? 0 : (cx >= g.gx ? g.gx - 1 : cx);
  cy = cy < 0 ? 0 : (cy >= g.gy ? g.gy - 1 : cy);
  const int32_t c = (int32_t)(cell_off[i] + cy * g.gx + cx);
  cell[p] = c;
  atomicAdd(&cnt[c], 1);
}

template <int T>
__global__ void __launch_bounds__(T) k_knn_b(int64_t n, int n_inst, int K, int mode,
                                             const int64_t* __restrict__ node_off,
                                             const int64_t* __restrict__ cell_off,
                                             const GridGeom* __restrict__ geom, const int32_t* __restrict__ cstart,
                                             const int32_t* __restrict__ order, const double* __restrict__ sx,
                                             const double* __restrict__ sy, const int32_t* __restrict__ cell,
                                             int32_t* __restrict__ out_id, double* __restrict__ out_c) {
  extern __shared__ unsigned char smem[];
  double* bc = reinterpret_cast<double*>(smem);
  int32_t* bw = reinterpret_cast<int32_t*>(bc + (size_t)K * T);
  const int64_t p = (int64_t)blockIdx.x * T + threadIdx.x;
  if (p >= n) return;
  const int t = threadIdx.x;
  const int32_t v = order[p];
  const int i = inst_of(v, node_off, n_inst);
  const int64_t co = cell_off[i];
  knn_search<T>(K, mode, geom[i], cstart + co, order, sx, sy, v, sx[p], sy[p], cell[v] - co, bc, bw);
  for (int j = 0; j < K; ++j) {
    out_id[(int64_t)v * K + j] = bw[j * T + t];
    out_c[(int64_t)v * K + j] = bc[j * T + t];
  }
}

// v among w's K directed neighbours: every entry read, no early exit, so
// that the row's loads are issued together
__device__ __forceinline__ bool in_list(const int32_t* __restrict__ knn, int64_t w, int K, int32_t v) {
  const int32_t* l = knn + w * K;
  bool f = false;
#pragma unroll 4
  for (int j = 0; j < K; ++j) f |= l[j] == v;
  return f;
}

// Directed edge i = (v -> w): flag[i] = 1 when v is not among w's neighbours
// (the symmetric graph then needs the reverse slot w -> v), counted in rev[w].
__global__ void k_rev_flag(int64_t n, int K, const int32_t* __restrict__ knn, int32_t* __restrict__ rev,
                           uint8_t* __restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * K) return;
  int32_t v = (int32_t)(i / K);
  int32_t w = knn[i];
  const bool need = !in_list(knn, w, K, v);
  flag[i] = need ? 1 : 0;
  if (need) atomicAdd(&rev[w], 1);
}

__global__ void k_deg(int64_t n, int K, const int32_t* __restrict__ rev, int32_t* __restrict__ deg) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) deg[v] = K + rev[v];
}

// Unsorted slots: v's K directed neighbours at rp[v] + j, then the reverse
// slots of the edges flagged, in atomic order; town = the slot's node.
template <class RP>
__global__ void k_fill(int64_t n, int K, const int32_t* __restrict__ knn, const double* __restrict__ kc,
                       const uint8_t* __restrict__ flag, const RP* __restrict__ rp, int32_t* __restrict__ cur,
                       int32_t* __restrict__ town, int32_t* __restrict__ nbr, double* __restrict__ cost) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * K) return;
  int32_t v = (int32_t)(i / K);
  int j = (int)(i % K);
  int32_t w = knn[i];
  double c = kc[i];
  int64_t s = (int64_t)rp[v] + j;
  town[s] = v;
  nbr[s] = w;
  cost[s] = c;
  if (flag[i]) {
    int64_t q = (int64_t)rp[w] + K + atomicAdd(&cur[w], 1);
    town[q] = w;
    nbr[q] = v;
    cost[q] = c;
  }
}

// Each node's slots in (cost, nbr) order: slot s goes to its rank among its
// node's slots (the neighbours of a node are distinct, so the keys are, and
// the ranks are a permutation).  One thread per slot; the node's slots are
// read from the unsorted copy.
template <class RP>
__global__ void k_rank_sort(int64_t m, const RP* __restrict__ rp, const int32_t* __restrict__ town,
                            const int32_t* __restrict__ tnbr, const double* __restrict__ tcost,
                            int32_t* __restrict__ nbr, double* __restrict__ cost) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= m) return;
  const int32_t v = town[s];
  const int64_t b = rp[v], e = rp[v + 1];
  const double c = tcost[s];
  const int32_t w = tnbr[s];
  int64_t r = 0;
  for (int64_t t = b; t < e; ++t) {
    const double ct = tcost[t];
    r += (ct < c || (ct == c && tnbr[t] < w)) ? 1 : 0;
  }
  nbr[b + r] = w;
  cost[b + r] = c;
}

// Directed kNN of one instance into (out_id, out_c) [n*K]; ids + id_offset.
f2m_status knn_directed_into(f2m_ctx* c, int64_t n, const double* x, const double* y, int K, int mode,
                             int32_t* out_id, double* out_c, int32_t id_offset) {
  F2M_HIP(c, c->k_red.reserve(64));
  unsigned long long* bb = c->k_red.as<unsigned long long>();
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, c->stream, bb, (int64_t)4);
  int nb = grid_for(n, 256);
  if (nb > 1024) nb = 1024;
  hipLaunchKernelGGL(k_bbox, dim3(nb), dim3(256), 0, c->stream, x, y, n, bb);
  F2M_LAUNCH_CHECK(c);
  unsigned long long hb[4];
  F2M_HIP(c, hipMemcpyAsync(hb, bb, sizeof hb, hipMemcpyDeviceToHost, c->stream));
  F2M_HIP(c, hipStreamSynchronize(c->stream));
  const GridGeom g = grid_geom(from_ordered_bits(~hb[0]), from_ordered_bits(~hb[1]), from_ordered_bits(hb[2]),
                               from_ordered_bits(hb[3]), n);
  int64_t ncell = g.gx * g.gy;
  F2M_HIP(c, c->k_cell.reserve(n * 4));
  F2M_HIP(c, c->k_cstart.reserve((ncell + 1) * 4));
  F2M_HIP(c, c->k_cur.reserve((ncell + 1) * 4 > n * 4 ? (ncell + 1) * 4 : n * 4));
  F2M_HIP(c, c->k_order.reserve(n * 4));
  F2M_HIP(c, c->k_sx.reserve(n * 8));
  F2M_HIP(c, c->k_sy.reserve(n * 8));
  int32_t* cnt = c->k_cur.as<int32_t>();
  hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(ncell, 256)), dim3(256), 0, c->stream, cnt, ncell);
  hipLaunchKernelGGL(k_cell, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, x, y, n, g,
                     c->k_cell.as<int32_t>(), cnt);
  F2M_LAUNCH_CHECK(c);
  F2M_HIP(c, scan_i32_to_i32(c, cnt, c->k_cstart.as<int32_t>(), ncell, nullptr));
  hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(ncell, 256)), dim3(256), 0, c->stream, cnt, ncell);
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, x, y, n,
                     c->k_cell.as<int32_t>(), c->k_cstart.as<int32_t>(), cnt, c->k_order.as<int32_t>(),
                     c->k_sx.as<double>(), c->k_sy.as<double>());
  F2M_LAUNCH_CHECK(c);
  size_t lds;
  if (K <= 32) {
    lds = (size_t)K * 128 * 12;
    hipLaunchKernelGGL(k_knn<128>, dim3(grid_for(n, 128)), dim3(128), lds, c->stream, n, K, mode, g,
                       c->k_cstart.as<int32_t>(), c->k_order.as<int32_t>(), c->k_sx.as<double>(),
                       c->k_sy.as<double>(), c->k_cell.as<int32_t>(), out_id, out_c, id_offset);
  } else {
    lds = (size_t)K * 64 * 12;
    hipLaunchKernelGGL(k_knn<64>, dim3(grid_for(n, 64)), dim3(64), lds, c->stream, n, K, mode, g,
                       c->k_cstart.as<int32_t>(), c->k_order.as<int32_t>(), c->k_sx.as<double>(),
                       c->k_sy.as<double>(), c->k_cell.as<int32_t>(), out_id, out_c, id_offset);
  }
  F2M_LAUNCH_CHECK(c);
  return F2M_OK;
}

// Directed kNN of every instance of a batch into (out_id, out_c) [N*K], ids
// global: one pipeline (bbox, grid geometry, cells, scan, scatter, ring
// search) over all instances, no host round trip.  node_off_h / node_off_d:
// the instance offsets on the host and on the device.
f2m_status knn_directed_batched(f2m_ctx* c, int n_inst, const int64_t* node_off_h, const int64_t* node_off_d,
                                const double* x, const double* y, int K, int mode, int32_t* out_id,
                                double* out_c) {
  const int64_t n = node_off_h[n_inst];
  std::vector<int64_t> coff(n_inst + 1, 0);
  for (int i = 0; i < n_inst; ++i) coff[i + 1] = coff[i] + 4 * (node_off_h[i + 1] - node_off_h[i]) + 16;
  const int64_t ncell = coff[n_inst];
  if (ncell >= (int64_t)INT32_MAX) return fail(c, F2M_ERR_LIMIT, "batch too large for 32-bit cell offsets");
  F2M_HIP(c, c->k_red.reserve(n_inst * 32ull + 64));
  F2M_HIP(c, c->k_geom.reserve(n_inst * (sizeof(GridGeom) + 8ull) + 64));
  F2M_HIP(c, c->k_cell.reserve(n * 4));
  F2M_HIP(c, c->k_cstart.reserve((ncell + 1) * 4));
  F2M_HIP(c, c->k_cur.reserve((ncell + 1) * 4));
  F2M_HIP(c, c->k_order.reserve(n * 4));
  F2M_HIP(c, c->k_sx.reserve(n * 8));
  F2M_HIP(c, c->k_sy.reserve(n * 8));
  unsigned long long* bb = c->k_red.as<unsigned long long>();
  GridGeom* geom = c->k_geom.as<GridGeom>();
  int64_t* cell_off = reinterpret_cast<int64_t*>(geom + n_inst);
  F2M_HIP(c, hipMemcpyAsync(cell_off, coff.data(), (n_inst + 1) * 8ull, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_bbox_b, dim3(n_inst), dim3(256), 0, c->stream, x, y, node_off_d, bb);
  hipLaunchKernelGGL(k_geom_b, dim3(grid_for(n_inst, 256)), dim3(256), 0, c->stream, n_inst, node_off_d, bb, geom);
  int32_t* cnt = c->k_cur.as<int32_t>();
  hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(ncell, 256)), dim3(256), 0, c->stream, cnt, ncell);
  hipLaunchKernelGGL(k_cell_b, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, n_inst, node_off_d, cell_off,
                     geom, x, y, c->k_cell.as<int32_t>(), cnt);
  F2M_LAUNCH_CHECK(c);
  F2M_HIP(c, scan_i32_to_i32(c, cnt, c->k_cstart.as<int32_t>(), ncell, nullptr));
  hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(ncell, 256)), dim3(256), 0, c->stream, cnt, ncell);
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, x, y, n,
                     c->k_cell.as<int32_t>(), c->k_cstart.as<int32_t>(), cnt, c->k_order.as<int32_t>(),
                     c->k_sx.as<double>(), c->k_sy.as<double>());
  F2M_LAUNCH_CHECK(c);
  if (K <= 32) {
    hipLaunchKernelGGL(k_knn_b<128>, dim3(grid_for(n, 128)), dim3(128), (size_t)K * 128 * 12, c->stream, n, n_inst,
                       K, mode, node_off_d, cell_off, geom, c->k_cstart.as<int32_t>(), c->k_order.as<int32_t>(),
                       c->k_sx.as<double>(), c->k_sy.as<double>(), c->k_cell.as<int32_t>(), out_id, out_c);
  } else {
    hipLaunchKernelGGL(k_knn_b<64>, dim3(grid_for(n, 64)), dim3(64), (size_t)K * 64 * 12, c->stream, n, n_inst,
                       K, mode, node_off_d, cell_off, geom, c->k_cstart.as<int32_t>(), c->k_order.as<int32_t>(),
                       c->k_sx.as<double>(), c->k_sy.as<double>(), c->k_cell.as<int32_t>(), out_id, out_c);
  }
  F2M_LAUNCH_CHECK(c);
  return F2M_OK;
}

static f2m_status knn_directed_dev(f2m_ctx* c, int64_t n, const double* x, const double* y, int K, int mode) {
  F2M_HIP(c, c->k_knn.reserve(n * K * 4));
  F2M_HIP(c, c->k_kcost.reserve(n * K * 8));
  return knn_directed_into(c, n, x, y, K, mode, c->k_knn.as<int32_t>(), c->k_kcost.as<double>(), 0);
}

// Symmetrise a directed kNN (global ids) into a CSR at (rowptr, nbr, cost);
// RP = int32/int64.
template <class RP>
static f2m_status symmetrise(f2m_ctx* c, int64_t n, int K, const int32_t* knn, const double* kc, DevBuf& rpb,
                             DevBuf& nbb, DevBuf& cb, int64_t* nslots_out) {
  F2M_HIP(c, c->k_rev.reserve(n * 4));
  F2M_HIP(c, c->k_deg.reserve(n * 4));
  F2M_HIP(c, c->k_flag.reserve(n * K));
  int32_t* rev = c->k_rev.as<int32_t>();
  uint8_t* flag = c->k_flag.as<uint8_t>();
  hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, rev, n);
  hipLaunchKernelGGL(k_rev_flag, dim3(grid_for(n * K, 256)), dim3(256), 0, c->stream, n, K, knn, rev, flag);
  hipLaunchKernelGGL(k_deg, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, K, rev, c->k_deg.as<int32_t>());
  F2M_LAUNCH_CHECK(c);
  F2M_HIP(c, rpb.reserve((n + 1) * sizeof(RP)));
  RP total = 0;
  if (sizeof(RP) == 4)
    F2M_HIP(c, scan_i32_to_i32(c, c->k_deg.as<int32_t>(), (int32_t*)rpb.p, n, (int32_t*)&total));
  else
    F2M_HIP(c, scan_i32_to_i64(c, c->k_deg.as<int32_t>(), (int64_t*)rpb.p, n, (int64_t*)&total));
  int64_t m = (int64_t)total;
  F2M_HIP(c, nbb.reserve(m * 4));
  F2M_HIP(c, cb.reserve(m * 8));
  F2M_HIP(c, c->k_town.reserve(m * 4));
  F2M_HIP(c, c->k_tnbr.reserve(m * 4));
  F2M_HIP(c, c->k_tcost.reserve(m * 8));
  hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, rev, n);
  hipLaunchKernelGGL(k_fill<RP>, dim3(grid_for(n * K, 256)), dim3(256), 0, c->stream, n, K, knn, kc,
                     (const uint8_t*)flag, (const RP*)rpb.p, rev, c->k_town.as<int32_t>(), c->k_tnbr.as<int32_t>(),
                     c->k_tcost.as<double>());
  hipLaunchKernelGGL(k_rank_sort<RP>, dim3(grid_for(m, 256)), dim3(256), 0, c->stream, m, (const RP*)rpb.p,
                     (const int32_t*)c->k_town.as<int32_t>(), (const int32_t*)c->k_tnbr.as<int32_t>(),
                     (const double*)c->k_tcost.as<double>(), nbb.as<int32_t>(), cb.as<double>());
  F2M_LAUNCH_CHECK(c);
  *nslots_out = m;
  return F2M_OK;
}

f2m_status symmetrise_i64(f2m_ctx* c, int64_t n, int K, const int32_t* knn, const double* kc, DevBuf& rpb,
                          DevBuf& nbb, DevBuf& cb, int64_t* nslots_out) {
  return symmetrise<int64_t>(c, n, K, knn, kc, rpb, nbb, cb, nslots_out);
}

}  // namespace f2m

using namespace f2m;

extern "C" {

f2m_status f2m_knn_directed(f2m_ctx* c, int32_t n, const double* x, const double* y, int32_t k,
                            int32_t cost_mode, int32_t* out) {
  if (!c || !x || !y || !out || n < 2 || k < 1 || cost_mode < 0 || cost_mode > 2) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  int K = k > n - 1 ? n - 1 : k;
  if (K > 80) return fail(c, F2M_ERR_LIMIT, "k > 80");
  const double* dx = (const double*)stage_in(c, x, n * 8ull, 0);
  const double* dy = (const double*)stage_in(c, y, n * 8ull, 1);
  if (!dx || !dy) return F2M_ERR_HIP;
  f2m_status s = knn_directed_dev(c, n, dx, dy, K, cost_mode);
  if (s) return s;
  F2M_HIP(c, copy_out(c, out, c->k_knn.p, (size_t)n * K * 4));
  return F2M_OK;
}

f2m_status f2m_knn_build(f2m_ctx* c, int32_t n, const double* x, const double* y, int32_t k,
                         int32_t cost_mode, f2m_graph* out) {
  if (!c || !x || !y || !out || n < 2 || k < 1 || cost_mode < 0 || cost_mode > 2) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  int K = k > n - 1 ? n - 1 : k;
  if (K > 80) return fail(c, F2M_ERR_LIMIT, "k > 80");
  if ((int64_t)n * K * 2 > 2000000000ll) return fail(c, F2M_ERR_LIMIT, "graph exceeds 2^31 slots");
  const double* dx = (const double*)stage_in(c, x, n * 8ull, 0);
  const double* dy = (const double*)stage_in(c, y, n * 8ull, 1);
  if (!dx || !dy) return F2M_ERR_HIP;
  f2m_status s = knn_directed_dev(c, n, dx, dy, K, cost_mode);
  if (s) return s;
  int64_t m = 0;
  s = symmetrise<int32_t>(c, n, K, c->k_knn.as<int32_t>(), c->k_kcost.as<double>(), c->g_rowptr, c->g_nbr,
                         c->g_cost, &m);
  if (s) return s;
  F2M_HIP(c, hipStreamSynchronize(c->stream));
  out->n = n;
  out->k = k;
  out->cost_mode = cost_mode;
  out->reserved = 0;
  out->nslots = m;
  out->d_min = 0;
  out->rowptr = c->g_rowptr.as<int32_t>();
  out->nbr = c->g_nbr.as<int32_t>();
  out->cost = c->g_cost.as<double>();
  return F2M_OK;
}

namespace {
// Caller CSR checks: rowptr starts at 0, ends at nslots and never decreases;
// neighbours are node ids other than the row's own.  Rows with bad bounds are
// not scanned.
__global__ void k_csr_check(int n, int64_t nslots, const int32_t* __restrict__ rp, const int32_t* __restrict__ nbr,
                            int32_t* __restrict__ bad) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int64_t b = rp[v], e = rp[v + 1];
  bool ok = (v > 0 || b == 0) && (v < n - 1 || e == nslots) && b >= 0 && e >= b && e <= nslots;
  if (ok)
    for (int64_t s = b; s < e; ++s) {
      const int w = nbr[s];
      if (w < 0 || w >= n || w == v) { ok = false; break; }
    }
  if (!ok) atomicOr(bad, 1);
}
__global__ void k_zero1(int32_t* p) { *p = 0; }
}  // namespace

f2m_status f2m_graph_from_csr(f2m_ctx* c, int32_t n, const int32_t* rowptr, const int32_t* nbr,
                              const double* cost, int64_t nslots, f2m_graph* out) {
  if (!c || !rowptr || !nbr || !cost || !out || n < 1 || nslots < 0) return F2M_ERR_ARG;
  if (nslots > 2000000000ll) return fail(c, F2M_ERR_LIMIT, "graph exceeds 2^31 slots");
  F2M_HIP(c, hipSetDevice(c->device));
  F2M_HIP(c, c->g_rowptr.reserve((n + 1) * 4ull));
  F2M_HIP(c, c->g_nbr.reserve(nslots * 4ull));
  F2M_HIP(c, c->g_cost.reserve(nslots * 8ull));
  F2M_HIP(c, hipMemcpyAsync(c->g_rowptr.p, rowptr, (n + 1) * 4ull, hipMemcpyDefault, c->stream));
  F2M_HIP(c, hipMemcpyAsync(c->g_nbr.p, nbr, nslots * 4ull, hipMemcpyDefault, c->stream));
  F2M_HIP(c, hipMemcpyAsync(c->g_cost.p, cost, nslots * 8ull, hipMemcpyDefault, c->stream));
  F2M_HIP(c, c->k_red.reserve(256));
  int32_t* bad = c->k_red.as<int32_t>();
  hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, c->stream, bad);
  hipLaunchKernelGGL(k_csr_check, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, n, nslots,
                     c->g_rowptr.as<int32_t>(), c->g_nbr.as<int32_t>(), bad);
  F2M_LAUNCH_CHECK(c);
  int32_t* hp = static_cast<int32_t*>(c->pinned);
  F2M_HIP(c, hipMemcpyAsync(hp, bad, 4, hipMemcpyDeviceToHost, c->stream));
  F2M_HIP(c, hipStreamSynchronize(c->stream));
  if (hp[0]) return fail(c, F2M_ERR_ARG, "malformed CSR: rowptr must run 0..nslots without decreasing, "
                                         "neighbours must be other node ids in [0, n)");
  out->n = n;
  out->k = 0;
  out->cost_mode = -1;
  out->reserved = 0;
  out->nslots = nslots;
  out->d_min = 0;
  out->rowptr = c->g_rowptr.as<int32_t>();
  out->nbr = c->g_nbr.as<int32_t>();
  out->cost = c->g_cost.as<double>();
  return F2M_OK;
}

f2m_status f2m_graph_copy(f2m_ctx* c, const f2m_graph* g, int32_t* rowptr, int32_t* nbr, double* cost) {
  if (!c || !g) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  F2M_HIP(c, copy_out(c, rowptr, g->rowptr, (g->n + 1) * 4ull));
  F2M_HIP(c, copy_out(c, nbr, g->nbr, g->nslots * 4ull));
  F2M_HIP(c, copy_out(c, cost, g->cost, g->nslots * 8ull));
  return F2M_OK;
}

}  // extern "C"
