// f2m_common.hip -- context, errors, staging and the device scan used by the
// graph builders.
#include <cstdio>
#include <cstring>

#include "f2m_internal.h"

namespace f2m {

f2m_status fail(f2m_ctx* c, f2m_status s, const std::string& msg) {
  if (c) c->err = msg;
  return s;
}
f2m_status fail(f2m_ctx* c, f2m_status s, const char* msg) {
  if (c) c->err.assign(msg);
  return s;
}

f2m_status hip_fail(f2m_ctx* c, hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what,
           file, line);
  return fail(c, F2M_ERR_HIP, buf);
}

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  std::memset(&a, 0, sizeof a);
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

const void* stage_in(f2m_ctx* c, const void* p, size_t bytes, int slot) {
  if (!p) return nullptr;
  if (is_device_ptr(p)) return p;
  hipError_t e = c->stage[slot].reserve(bytes);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->stage[slot].p, p, bytes, hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) {
    hip_fail(c, e, "stage_in", __FILE__, __LINE__);
    return nullptr;
  }
  return c->stage[slot].p;
}

hipError_t copy_out(f2m_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!dst || !bytes) return hipSuccess;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, c->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(c->stream);
}

// ---------------------------------------------------------------------------
// Exclusive scan: tiles of 2048 (256 threads x 8), tile sums scanned
// recursively, then offsets added.  Sums are carried in int64.
// ---------------------------------------------------------------------------
constexpr int SCAN_T = 256, SCAN_V = 8, SCAN_TILE = SCAN_T * SCAN_V;

template <class Tin>
__global__ void __launch_bounds__(SCAN_T) k_scan_tiles(const Tin* __restrict__ in, int64_t* __restrict__ out,
                                                       int64_t* __restrict__ tile_sum, int64_t n) {
  __shared__ int64_t wsum[SCAN_T / 64];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_V;
  int64_t v[SCAN_V];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_V; ++i) {
    int64_t idx = base + i;
    v[i] = idx < n ? (int64_t)in[idx] : 0;
    s += v[i];
  }
  // inclusive wave scan of per-thread sums
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int64_t inc = s;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int64_t t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int64_t wpre = 0;
  for (int k = 0; k < wid; ++k) wpre += wsum[k];
  int64_t excl = wpre + inc - s;
#pragma unroll
  for (int i = 0; i < SCAN_V; ++i) {
    int64_t idx = base + i;
    if (idx < n) out[idx] = excl;
    excl += v[i];
  }
  if (threadIdx.x == SCAN_T - 1) tile_sum[blockIdx.x] = wpre + inc;
}

template <class Tout>
__global__ void k_scan_finish(const int64_t* __restrict__ loc, const int64_t* __restrict__ tile_off,
                              Tout* __restrict__ out, int64_t n, const int64_t* __restrict__ total) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (Tout)(loc[i] + tile_off[i / SCAN_TILE]);
  if (i == 0) out[n] = (Tout)(*total);
}

// Scans int64 values in place-ish: out[0..n] exclusive, out[n] = total.
static hipError_t scan_i64_rec(f2m_ctx* c, const int64_t* in, int64_t* out, int64_t n,
                               int64_t* scratch, size_t scratch_elems) {
  int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (tiles < 1) tiles = 1;
  int64_t* loc = scratch;
  int64_t* tsum = loc + n;
  int64_t* toff = tsum + tiles;
  int64_t* rest = toff + tiles + 1;
  size_t used = (size_t)(n + 2 * tiles + 1);
  if (used > scratch_elems) return hipErrorOutOfMemory;
  hipLaunchKernelGGL(k_scan_tiles<int64_t>, dim3((unsigned)tiles), dim3(SCAN_T), 0, c->stream, in, loc, tsum, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (tiles == 1) {
    hipLaunchKernelGGL(k_scan_tiles<int64_t>, dim3(1), dim3(SCAN_T), 0, c->stream, tsum, toff, toff + 1, (int64_t)1);
  } else {
    e = scan_i64_rec(c, tsum, toff, tiles, rest, scratch_elems - used);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_scan_finish<int64_t>, dim3(grid_for(n + 1, 256)), dim3(256), 0, c->stream, loc, toff,
                     out, n, toff + tiles);
  return hipGetLastError();
}

template <class Tout>
static hipError_t scan_i32_impl(f2m_ctx* c, const int32_t* in, Tout* out, int64_t n, Tout* total_host) {
  int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (tiles < 1) tiles = 1;
  size_t need = (size_t)(n + 4 * tiles + 64) * 2 + 64;
  hipError_t e = c->k_scan.reserve(need * sizeof(int64_t));
  if (e != hipSuccess) return e;
  int64_t* loc = c->k_scan.as<int64_t>();
  int64_t* tsum = loc + n;
  int64_t* toff = tsum + tiles;
  int64_t* rest = toff + tiles + 1;
  hipLaunchKernelGGL(k_scan_tiles<int32_t>, dim3((unsigned)tiles), dim3(SCAN_T), 0, c->stream, in, loc, tsum, n);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (tiles == 1) {
    hipLaunchKernelGGL(k_scan_tiles<int64_t>, dim3(1), dim3(SCAN_T), 0, c->stream, tsum, toff, toff + 1, (int64_t)1);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  } else {
    e = scan_i64_rec(c, tsum, toff, tiles, rest, need - (size_t)(n + 2 * tiles + 1));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_scan_finish<Tout>, dim3(grid_for(n + 1, 256)), dim3(256), 0, c->stream, loc, toff, out,
                     n, toff + tiles);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (total_host) {
    e = hipMemcpyAsync(total_host, out + n, sizeof(Tout), hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(c->stream);
  }
  return e;
}

// Single-pass exclusive scan (decoupled look-back).  Tile b publishes its
// aggregate, then walks back over tiles b-1, b-2, ... adding aggregates until
// it meets a published prefix, and publishes its own prefix.  A status word
// is (epoch << 34) | (flag << 32) | value, flag 1 = aggregate, 2 = prefix;
// a word of an earlier call carries another epoch and reads as not ready.
// A workgroup takes its tile index from a ticket counter (word 0 of the
// status array; the statuses follow it), not from blockIdx.x: tickets are
// handed out in the order workgroups start running, so every tile a tile
// waits on belongs to a workgroup that already runs and will publish, and
// every wait ends whatever order the dispatcher uses.  The workgroup that
// draws the last ticket resets the counter to 0 for the next launch (all
// other tickets were drawn before it).
constexpr int LB_T = 256, LB_V = 8, LB_TILE = LB_T * LB_V;
constexpr unsigned LB_EPOCH_MAX = (1u << 30) - 1;

__device__ __forceinline__ int32_t wave_sum_i32(int32_t x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

__global__ void k_lb_clear(unsigned long long* st, int64_t m) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) st[i] = 0ull;
}

__global__ void __launch_bounds__(LB_T) k_scan_lb(const int32_t* __restrict__ in, int32_t sub,
                                                  int32_t* __restrict__ out, int64_t n,
                                                  unsigned long long* __restrict__ st, unsigned epoch,
                                                  int32_t* __restrict__ set_ptr, int32_t set_val) {
  __shared__ int32_t wsum[LB_T / 64];
  __shared__ int32_t s_pre;
  __shared__ int32_t s_tile;
  const int tid = threadIdx.x;
  if (tid == 0) {
    const unsigned long long t = atomicAdd(st, 1ull);
    if (t == (unsigned long long)gridDim.x - 1ull) atomicExch(st, 0ull);
    s_tile = (int32_t)t;
  }
  __syncthreads();
  const int b = s_tile;
  st += 1;  // the tile statuses
  const int64_t base = (int64_t)b * LB_TILE + (int64_t)tid * LB_V;
  int32_t v[LB_V];
  int32_t s = 0;
#pragma unroll
  for (int i = 0; i < LB_V; ++i) {
    const int64_t idx = base + i;
    const int32_t x = idx < n ? in[idx] : 0;
    v[i] = x > sub ? x - sub : 0;
    s += v[i];
  }
  const int lane = tid & 63, wid = tid >> 6;
  int32_t inc = s;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t t = __shfl_up(inc, off, 64);
    if (lane >= off) inc += t;
  }
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int32_t wpre = 0, agg = 0;
#pragma unroll
  for (int k = 0; k < LB_T / 64; ++k) {
    if (k < wid) wpre += wsum[k];
    agg += wsum[k];
  }
  if (wid == 0) {
    // the first wave looks back 64 tiles at a time: lane l reads tile b-1-l
    // (below tile 0: a prefix of 0); it stops at the nearest prefix when
    // every tile above it is ready, adding their aggregates and the prefix
    const unsigned long long tag = (unsigned long long)epoch << 34;
    int32_t pre = 0;
    if (b > 0) {
      if (lane == 0)
        __hip_atomic_store(&st[b], tag | (1ull << 32) | (uint32_t)agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      for (int q0 = b - 1;;) {
        const int q = q0 - lane;
        const unsigned long long w =
            q >= 0 ? __hip_atomic_load(&st[q], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) : (tag | (2ull << 32));
        const unsigned flag = (unsigned)(w >> 32) & 3u;
        const bool ready = (w >> 34) == epoch && flag != 0;
        const unsigned long long notready = __ballot(!ready);
        const unsigned long long pmask = __ballot(ready && flag == 2);
        if (pmask) {
          const int lp = __builtin_ctzll(pmask);
          if ((notready & ((2ull << lp) - 1ull)) == 0) {
            pre += wave_sum_i32(lane <= lp ? (int32_t)(uint32_t)w : 0);
            break;
          }
        } else if (notready == 0) {
          pre += wave_sum_i32((int32_t)(uint32_t)w);
          q0 -= 64;
          continue;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    if (lane == 0) {
      __hip_atomic_store(&st[b], tag | (2ull << 32) | (uint32_t)(pre + agg), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
      s_pre = pre;
      if (b == 0 && set_ptr) *set_ptr = set_val;
    }
  }
  __syncthreads();
  int32_t excl = s_pre + wpre + inc - s;
#pragma unroll
  for (int i = 0; i < LB_V; ++i) {
    const int64_t idx = base + i;
    if (idx < n) out[idx] = excl;
    excl += v[i];
  }
  if (b == (int)gridDim.x - 1 && tid == 0) out[n] = s_pre + agg;
}

hipError_t scan_excess_i32(f2m_ctx* c, const int32_t* in, int32_t sub, int32_t* out, int64_t n, int32_t* set_ptr,
                           int32_t set_val) {
  const int64_t tiles = std::max<int64_t>(1, (n + LB_TILE - 1) / LB_TILE);
  if (tiles > INT32_MAX / 2) return hipErrorInvalidValue;
  const void* old = c->k_lbst.p;
  hipError_t e = c->k_lbst.reserve((size_t)(tiles + 1) * 8);  // ticket counter + statuses
  if (e != hipSuccess) return e;
  if (c->k_lbst.p != old || c->k_lb_epoch >= LB_EPOCH_MAX) {
    // fresh memory, or the epochs wrapped: no word may carry a live epoch
    const int64_t m = (int64_t)(c->k_lbst.cap / 8);
    hipLaunchKernelGGL(k_lb_clear, dim3(grid_for(m, 256)), dim3(256), 0, c->stream, c->k_lbst.as<unsigned long long>(),
                       m);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    c->k_lb_epoch = 0;
  }
  const unsigned epoch = ++c->k_lb_epoch;
  hipLaunchKernelGGL(k_scan_lb, dim3((unsigned)tiles), dim3(LB_T), 0, c->stream, in, sub, out, n,
                     c->k_lbst.as<unsigned long long>(), epoch, set_ptr, set_val);
  return hipGetLastError();
}

hipError_t scan_i32_to_i32(f2m_ctx* c, const int32_t* in, int32_t* out, int64_t n, int32_t* total_host) {
  return scan_i32_impl<int32_t>(c, in, out, n, total_host);
}
hipError_t scan_i32_to_i64(f2m_ctx* c, const int32_t* in, int64_t* out, int64_t n, int64_t* total_host) {
  return scan_i32_impl<int64_t>(c, in, out, n, total_host);
}

}  // namespace f2m

// ===========================================================================
extern "C" {

void f2m_default_opts(f2m_opts* o) {
  if (!o) return;
  o->err_grad = 1e-7;
  o->err_comb_mult = 1e2;
  o->glob_k = 5;
  o->glob_rounds = 200;
  o->glob_sweeps = 50;
  o->glob_exit = 0.01;
  o->max_outer = 10000;
  o->factor0 = 100;
  o->factor_step = 300;
  o->factor_max = 5000;
  o->check_every = 50;
  o->early_exit = 0;
  o->repair = 1;
  o->shift0 = 0.5;
  o->shift_max = 32;
  o->fix_rounds = 100;
  o->fix_factor = 500;
  o->frozen_sweeps = 20;
  o->stop_phase = 0;
  o->exact = 1;
  o->exact_alpha = 4;
  o->exact_eps0 = 0.02;
  o->exact_tol = 1e-10;
  o->exact_max_rounds = 1000000;
  o->exact_max_phases = 40;
  o->exact_reserved = 0;
}

const char* f2m_strerror(f2m_status s) {
  switch (s) {
    case F2M_OK: return "ok";
    case F2M_ERR_ARG: return "invalid argument";
    case F2M_ERR_HIP: return "HIP runtime error";
    case F2M_ERR_FORMAT: return "not a TSPLIB coordinate file";
    case F2M_ERR_IO: return "file I/O error";
    case F2M_ERR_DEGENERATE: return "degenerate instance";
    case F2M_ERR_LIMIT: return "size limit exceeded";
    default: return "unknown error";
  }
}

const char* f2m_last_error(const f2m_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

const char* f2m_version(void) { return "f2m_hip gfx950 " __DATE__ " " __TIME__; }

f2m_status f2m_ctx_create(int32_t device, void* stream, f2m_ctx** out) {
  if (!out) return F2M_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) return F2M_ERR_HIP;
  if (device < 0 || device >= ndev) return F2M_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return F2M_ERR_HIP;
  if (f2m::noalloc_init(device) != hipSuccess) return F2M_ERR_HIP;
  f2m_ctx* c = new f2m_ctx;
  c->device = device;
  c->err.reserve(1024);
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      return F2M_ERR_HIP;
    }
    c->own_stream = true;
  }
  if (hipHostMalloc(&c->pinned, 1 << 16, hipHostMallocDefault) != hipSuccess) {
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return F2M_ERR_HIP;
  }
  *out = c;
  return F2M_OK;
}

static void release_ws(f2m::SolveWs& w) {
  f2m::DevBuf* all[] = {&w.nbr, &w.D, &w.orig, &w.F0, &w.F1, &w.g, &w.act, &w.clen, &w.deg,
                        &w.A, &w.T, &w.root, &w.irr, &w.frozen, &w.xw, &w.xc, &w.red, &w.axb,
                        &w.sortk, &w.sortv, &w.sortk2, &w.sortv2, &w.sorttmp, &w.ef, &w.efoff, &w.psl, &w.povo, &w.ptmp,
                        &w.xmir, &w.xf, &w.xp, &w.xe, &w.xact, &w.xcnt, &w.thdr, &w.tnode, &w.tslot, &w.tovf};
  for (auto* b : all) b->release();
}

void f2m_ctx_destroy(f2m_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  f2m::release_batch_plan(c);
  f2m::DevBuf* all[] = {&c->g_rowptr, &c->g_nbr, &c->g_cost, &c->k_x, &c->k_y, &c->k_cell,
                        &c->k_cstart, &c->k_order, &c->k_sx, &c->k_sy, &c->k_knn, &c->k_kcost,
                        &c->k_rev, &c->k_cur, &c->k_deg, &c->k_red, &c->k_scan, &c->k_geom, &c->k_lbst, &c->b_node_off,
                        &c->b_rowptr, &c->b_nbr, &c->b_cost, &c->b_order, &c->b_stats,
                        &c->b_cur, &c->b_knn, &c->b_kc, &c->b_clk, &c->b_ef, &c->b_efoff, &c->b_finit};
  for (auto* b : all) b->release();
  for (auto& s : c->stage) s.release();
  for (auto& s : c->o_buf) s.release();
  release_ws(c->ws);
  release_ws(c->bws);
  for (auto e : c->prof_ev) (void)hipEventDestroy(e);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

f2m_status f2m_ctx_synchronize(f2m_ctx* c) {
  if (!c) return F2M_ERR_ARG;
  F2M_HIP(c, hipStreamSynchronize(c->stream));
  return F2M_OK;
}

}  // extern "C"
