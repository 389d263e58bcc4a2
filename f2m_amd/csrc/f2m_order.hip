// f2m_order.hip -- spatial numbering of point sets (new; the reference numbers
// nodes in unordered_map order, graph_tour.hpp:97, i.e. arbitrarily).
//
// Morton (Z-order) permutation: each point's coordinates are quantised to
// 16 bits per axis over its instance's bounding box,
//     q = min(65535, floor(((x - min x) / (max x - min x)) * 65535))
// (q = 0 on a zero extent), the two quantised values are bit-interleaved
// (x in the even bits), and the points are sorted stably by (instance,
// key), ties keeping input order.  Numbering the nodes of a graph in this
// order puts graph neighbours close in memory, which is what the sweeps'
// neighbour-potential gathers need.
//
// The sort is a least-significant-digit radix sort written here: 8-bit
// digits, one pass per digit of the key's used bits (4 for one instance, up
// to 6 for 4096), each pass a per-tile digit histogram, one exclusive scan of
// the digit-major counts (the single-pass look-back scan of f2m_common.hip)
// and a stable scatter.  Stability comes from ranking each tile's elements in
// their input order: items one after another, waves in order within an item,
// lanes in order within a wave (lanes of equal digit found with eight
// ballots).
#include <cmath>
#include <utility>
#include <vector>

#include "f2m_internal.h"

namespace f2m {
namespace {

__device__ __forceinline__ uint32_t spread16(uint32_t v) {
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// one block per instance: [min x, min y, max x, max y]
__global__ void __launch_bounds__(256) o_bbox(const int64_t* __restrict__ off, const double* __restrict__ x,
                                              const double* __restrict__ y, double* __restrict__ box) {
  __shared__ double sh[4][4];
  const int64_t b = off[blockIdx.x], e = off[blockIdx.x + 1];
  double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
  for (int64_t i = b + threadIdx.x; i < e; i += 256) {
    mnx = fmin(mnx, x[i]);
    mny = fmin(mny, y[i]);
    mxx = fmax(mxx, x[i]);
    mxy = fmax(mxy, y[i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    mnx = fmin(mnx, __shfl_xor(mnx, o, 64));
    mny = fmin(mny, __shfl_xor(mny, o, 64));
    mxx = fmax(mxx, __shfl_xor(mxx, o, 64));
    mxy = fmax(mxy, __shfl_xor(mxy, o, 64));
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[wv][0] = mnx;
    sh[wv][1] = mny;
    sh[wv][2] = mxx;
    sh[wv][3] = mxy;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      sh[0][0] = fmin(sh[0][0], sh[k][0]);
      sh[0][1] = fmin(sh[0][1], sh[k][1]);
      sh[0][2] = fmax(sh[0][2], sh[k][2]);
      sh[0][3] = fmax(sh[0][3], sh[k][3]);
    }
    for (int k = 0; k < 4; ++k) box[4 * blockIdx.x + k] = sh[0][k];
  }
}

__device__ __forceinline__ uint32_t quant16(double v, double mn, double mx) {
  const double ext = mx - mn;
  if (!(ext > 0.0)) return 0u;
  const double q = floor(((v - mn) / ext) * 65535.0);
  return !(q > 0.0) ? 0u : (q >= 65535.0 ? 65535u : (uint32_t)q);  // NaN -> 0
}

// key = instance << 32 | morton, value = input index; one block per instance
__global__ void __launch_bounds__(256) o_keys(const int64_t* __restrict__ off, const double* __restrict__ x,
                                              const double* __restrict__ y, const double* __restrict__ box,
                                              unsigned long long* __restrict__ key, int64_t* __restrict__ val) {
  const int inst = blockIdx.x;
  const int64_t b = off[inst], e = off[inst + 1];
  const double mnx = box[4 * inst], mny = box[4 * inst + 1], mxx = box[4 * inst + 2], mxy = box[4 * inst + 3];
  for (int64_t i = b + threadIdx.x; i < e; i += 256) {
    const uint32_t m = spread16(quant16(x[i], mnx, mxx)) | (spread16(quant16(y[i], mny, mxy)) << 1);
    key[i] = ((unsigned long long)inst << 32) | m;
    val[i] = i;
  }
}

constexpr int RS_T = 256, RS_I = 8, RS_TILE = RS_T * RS_I, RS_BITS = 8, RS_BINS = 1 << RS_BITS;
static_assert(RS_T == RS_BINS, "one thread per digit in the per-tile tables");

// cnt[d * tiles + tile] = elements of the tile whose digit (at shift) is d
__global__ void __launch_bounds__(RS_T) o_hist(const unsigned long long* __restrict__ key, int64_t n, int shift,
                                               int32_t* __restrict__ cnt, int tiles) {
  __shared__ int32_t h[RS_BINS];
  const int tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
  for (int it = 0; it < RS_I; ++it) {
    const int64_t i = base + (int64_t)it * RS_T + tid;
    if (i < n) atomicAdd(&h[(unsigned)(key[i] >> shift) & (RS_BINS - 1)], 1);
  }
  __syncthreads();
  cnt[(int64_t)tid * tiles + blockIdx.x] = h[tid];
}

// Stable scatter of one pass: element i of the tile (item it = i / RS_T,
// thread i % RS_T) goes to off[d * tiles + tile] + (elements of digit d
// before it in the tile).  off = exclusive scan of o_hist's counts.
__global__ void __launch_bounds__(RS_T) o_scatter(const unsigned long long* __restrict__ kin,
                                                  const int64_t* __restrict__ vin,
                                                  unsigned long long* __restrict__ kout,
                                                  int64_t* __restrict__ vout, int64_t n, int shift,
                                                  const int32_t* __restrict__ off, int tiles) {
  constexpr int NW = RS_T / 64;
  __shared__ int32_t run[RS_BINS];     // next position of each digit in this tile
  __shared__ int32_t wc[NW][RS_BINS];  // per-wave digit counts of the current item
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  run[tid] = off[(int64_t)tid * tiles + blockIdx.x];
#pragma unroll
  for (int k = 0; k < NW; ++k) wc[k][tid] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int it = 0; it < RS_I; ++it) {
    const int64_t i = base + (int64_t)it * RS_T + tid;
    const bool ok = i < n;
    const unsigned long long k = ok ? kin[i] : 0ull;
    const int d = (int)((unsigned)(k >> shift) & (RS_BINS - 1));
    // lanes of this wave holding a valid element with digit d
    unsigned long long m = __ballot(ok);
#pragma unroll
    for (int b = 0; b < RS_BITS; ++b) {
      const bool bit = (d >> b) & 1;
      const unsigned long long bb = __ballot(ok && bit);
      m &= bit ? bb : ~bb;
    }
    const int r = __popcll(m & lt);
    if (ok && r == 0) wc[w][d] = __popcll(m);  // the lowest such lane
    __syncthreads();
    if (ok) {
      int pos = run[d] + r;
      for (int q = 0; q < w; ++q) pos += wc[q][d];
      if (pos >= 0 && (int64_t)pos < n) {
        kout[pos] = k;
        vout[pos] = vin[i];
      }
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      add += wc[q][tid];
      wc[q][tid] = 0;
    }
    run[tid] += add;
    __syncthreads();
  }
}

__global__ void o_to_i32(int64_t n, const int64_t* __restrict__ a, int32_t* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (int32_t)a[i];
}

// perm (device, int64, n_total) for instances [off[i], off[i+1]) of x, y (device)
f2m_status order_into(f2m_ctx* c, int n_inst, const int64_t* doff, int64_t n_total, const double* x,
                      const double* y, int64_t* perm) {
  DevBuf* B = c->o_buf;
  F2M_HIP(c, B[0].reserve(n_inst * 32ull));
  F2M_HIP(c, B[1].reserve(n_total * 8ull));
  F2M_HIP(c, B[2].reserve(n_total * 8ull));
  F2M_HIP(c, B[3].reserve(n_total * 8ull));
  hipLaunchKernelGGL(o_bbox, dim3(n_inst), dim3(256), 0, c->stream, doff, x, y, B[0].as<double>());
  hipLaunchKernelGGL(o_keys, dim3(n_inst), dim3(256), 0, c->stream, doff, x, y, B[0].as<double>(),
                     B[1].as<unsigned long long>(), B[2].as<int64_t>());
  F2M_LAUNCH_CHECK(c);
  int hb = 0;
  while (hb < 31 && (1ll << hb) < n_inst) ++hb;
  const int end_bit = 32 + hb;
  const int64_t tiles = (n_total + RS_TILE - 1) / RS_TILE;
  const int64_t ncnt = tiles * RS_BINS;
  if (ncnt >= INT32_MAX / 2) return fail(c, F2M_ERR_LIMIT, "spatial order: too many points");
  F2M_HIP(c, B[4].reserve((size_t)(2 * ncnt + 1) * 4));
  int32_t* cnt = B[4].as<int32_t>();
  int32_t* off = cnt + ncnt;
  // ping-pong: keys B[1] <-> B[3], values B[2] <-> perm
  unsigned long long* kin = B[1].as<unsigned long long>();
  unsigned long long* kout = B[3].as<unsigned long long>();
  int64_t* vin = B[2].as<int64_t>();
  int64_t* vout = perm;
  for (int shift = 0; shift < end_bit; shift += RS_BITS) {
    hipLaunchKernelGGL(o_hist, dim3((unsigned)tiles), dim3(RS_T), 0, c->stream, kin, n_total, shift, cnt, (int)tiles);
    F2M_LAUNCH_CHECK(c);
    F2M_HIP(c, scan_excess_i32(c, cnt, 0, off, ncnt));
    hipLaunchKernelGGL(o_scatter, dim3((unsigned)tiles), dim3(RS_T), 0, c->stream, kin, vin, kout, vout, n_total,
                       shift, off, (int)tiles);
    F2M_LAUNCH_CHECK(c);
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  if (vin != perm) F2M_HIP(c, hipMemcpyAsync(perm, vin, n_total * 8ull, hipMemcpyDeviceToDevice, c->stream));
  return F2M_OK;
}

}  // namespace
}  // namespace f2m

using namespace f2m;

extern "C" {

f2m_status f2m_spatial_order(f2m_ctx* c, int32_t n, const double* x, const double* y, int32_t* perm) {
  if (!c || n < 1 || !x || !y || !perm) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  const double* dx = (const double*)stage_in(c, x, n * 8ull, 0);
  const double* dy = (const double*)stage_in(c, y, n * 8ull, 1);
  if (!dx || !dy) return F2M_ERR_HIP;
  F2M_HIP(c, c->stage[2].reserve(16 + n * 8ull + n * 4ull));
  int64_t* doff = c->stage[2].as<int64_t>();
  int64_t* p64 = doff + 2;
  int32_t* p32 = reinterpret_cast<int32_t*>(p64 + n);
  const int64_t hoff[2] = {0, n};
  F2M_HIP(c, hipMemcpyAsync(doff, hoff, 16, hipMemcpyHostToDevice, c->stream));
  f2m_status s = order_into(c, 1, doff, n, dx, dy, p64);
  if (s) return s;
  hipLaunchKernelGGL(o_to_i32, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, (int64_t)n, p64, p32);
  F2M_LAUNCH_CHECK(c);
  F2M_HIP(c, copy_out(c, perm, p32, n * 4ull));
  return F2M_OK;
}

f2m_status f2m_spatial_order_batched(f2m_ctx* c, int32_t n_inst, const int64_t* node_off, const double* x,
                                     const double* y, int64_t* perm) {
  if (!c || n_inst < 1 || !node_off || !x || !y || !perm) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  const int64_t n_total = node_off[n_inst];
  if (node_off[0] != 0) return fail(c, F2M_ERR_ARG, "node_off[0] must be 0");
  if (n_total < 1 || n_total > 2000000000ll) return fail(c, F2M_ERR_LIMIT, "batch size out of range");
  for (int i = 0; i < n_inst; ++i)
    if (node_off[i + 1] < node_off[i]) return fail(c, F2M_ERR_ARG, "node_off must not decrease");
  const double* dx = (const double*)stage_in(c, x, n_total * 8ull, 0);
  const double* dy = (const double*)stage_in(c, y, n_total * 8ull, 1);
  if (!dx || !dy) return F2M_ERR_HIP;
  F2M_HIP(c, c->stage[2].reserve((n_inst + 1) * 8ull + n_total * 8ull));
  int64_t* doff = c->stage[2].as<int64_t>();
  int64_t* p64 = doff + n_inst + 1;
  F2M_HIP(c, hipMemcpyAsync(doff, node_off, (n_inst + 1) * 8ull, hipMemcpyHostToDevice, c->stream));
  f2m_status s = order_into(c, n_inst, doff, n_total, dx, dy, p64);
  if (s) return s;
  F2M_HIP(c, copy_out(c, perm, p64, n_total * 8ull));
  return F2M_OK;
}

}  // extern "C"
