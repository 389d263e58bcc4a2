// f2m_noalloc.hip -- f2m_sweep_noalloc: the Jacobi sweeps of f2m_sweep for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops that forbid hipMalloc).  No counterpart in the reference.
//
// Inputs and outputs are the caller's device arrays.  The only memory beyond
// them is `scratch`, a file-scope device array of fixed size: the potentials
// of the sweeps between the first and the last (ping-pong with F_out) and the
// 32 block-max slots of the last sweep.  Its device address is looked up once
// per device (hipGetSymbolAddress, at f2m_ctx_create with the device current)
// and only that address is handed to kernels and copies.  Every call on a
// device shares that device's scratch, so calls that overlap on one device
// (two streams, two host threads) race: serialise them.
//
// The sweep reads each node's first len[v] slots straight from the CSR rows
// (no compacted copy, which would need memory of the graph's size).  Per node
// the arithmetic is f2m_sweep's: d = (D + 0.5 F[lo]) + 0.5 F[hi], the three
// smallest values with sentinels 1e6, g = -(t2 + t3)/2, F' = F + g
// (graph_cuda_grad.hpp:1-84); the three values do not depend on the order in
// which the slots are visited, so the bits equal f2m_sweep's.
#include <cstring>
#include <mutex>

#include "f2m_device.h"

namespace f2m {
namespace {

constexpr int NA_BS = 256;
constexpr int NA_LPN = 4;
constexpr int NA_SLOTS = 32;              // block-max slots, 128 bytes apart
constexpr int NA_WORD = 8;                // u64 index of the host read-back word (slots use k * 16)
constexpr size_t NA_F_OFFSET = NA_SLOTS * 16 * 8;  // 4 KiB of slots, then the potentials
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NA_F_OFFSET + (size_t)F2M_NOALLOC_MAX_N * 8];

// device address of `scratch` per device (null until looked up)
void* g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

__global__ void na_zero_slots(unsigned long long* slots) {
  const int k = threadIdx.x;
  if (k < NA_SLOTS) slots[16 * k] = 0ull;
}

__global__ void __launch_bounds__(NA_BS) na_sweep(int n, const int32_t* __restrict__ rowptr,
                                                 const int32_t* __restrict__ nbr, const double* __restrict__ D,
                                                 const int32_t* __restrict__ len, const double* __restrict__ Fin,
                                                 double* __restrict__ Fout, double* __restrict__ gout,
                                                 unsigned long long* __restrict__ slots) {
  const int64_t gt = (int64_t)blockIdx.x * NA_BS + threadIdx.x;
  const int v = (int)(gt / NA_LPN), lane = (int)(gt % NA_LPN);
  Top3 t;
  double Fv = 0.0;
  const bool ok = v < n;
  if (ok) {
    const int b = rowptr[v], e = b + len[v];
    Fv = Fin[v];
    for (int k = b + lane; k < e; k += NA_LPN) {
      const int w = nbr[k];
      t.push(dval(D[k], Fv, Fin[w], v, w));
    }
  }
#pragma unroll
  for (int off = 1; off < NA_LPN; off <<= 1) {
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
    Fout[v] = Fv + g;
    gabs = fabs(g);
  }
  if (slots) {
    __shared__ double wm[NA_BS / 64];
    const double m = wave_max_d(gabs);
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double mm = wm[0];
      for (int k = 1; k < NA_BS / 64; ++k) mm = fmax(mm, wm[k]);
      atomicMax(slots + (blockIdx.x & (NA_SLOTS - 1)) * 16, (unsigned long long)__double_as_longlong(mm));
    }
  }
}

// max of the slots (|g| bits, >= 0) into out[0] as a double
__global__ void na_max_out(const unsigned long long* __restrict__ slots, double* __restrict__ out) {
  const int k = threadIdx.x;
  unsigned long long b = k < NA_SLOTS ? slots[16 * k] : 0ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(b, off, 64);
    b = o > b ? o : b;
  }
  if (k == 0) out[0] = __longlong_as_double((long long)b);
}

}  // namespace

hipError_t noalloc_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = p;
  return hipSuccess;
}

}  // namespace f2m

using namespace f2m;

extern "C" f2m_status f2m_sweep_noalloc(f2m_ctx* c, int32_t n, const int32_t* rowptr, const int32_t* nbr,
                                        const double* D, const int32_t* len, const double* F_in, double* F_out,
                                        double* g_out, int32_t nsweeps, double* max_abs_g) {
  if (!c || n < 1 || !rowptr || !nbr || !D || !len || !F_in || !F_out || nsweeps < 1) return F2M_ERR_ARG;
  if (n > F2M_NOALLOC_MAX_N) return fail(c, F2M_ERR_LIMIT, "f2m_sweep_noalloc: n above F2M_NOALLOC_MAX_N");
  if (F_in == F_out) return fail(c, F2M_ERR_ARG, "f2m_sweep_noalloc: F_in and F_out must differ");
  if (!is_device_ptr(rowptr) || !is_device_ptr(nbr) || !is_device_ptr(D) || !is_device_ptr(len) ||
      !is_device_ptr(F_in) || !is_device_ptr(F_out) || (g_out && !is_device_ptr(g_out)))
    return fail(c, F2M_ERR_ARG, "f2m_sweep_noalloc takes device arrays");
  F2M_HIP(c, hipSetDevice(c->device));
  void* base = (c->device >= 0 && c->device < MAX_DEVICES) ? g_scratch[c->device] : nullptr;
  if (!base) {  // not looked up at f2m_ctx_create (another device was current)
    F2M_HIP(c, noalloc_init(c->device));
    base = g_scratch[c->device];
  }
  unsigned long long* slots = static_cast<unsigned long long*>(base);
  double* Fs = reinterpret_cast<double*>(static_cast<unsigned char*>(base) + NA_F_OFFSET);
  const dim3 grid(grid_for((int64_t)n * NA_LPN, NA_BS)), block(NA_BS);
  // sweep i writes F_out when nsweeps - i is even, else the scratch
  // potentials, so the last sweep writes F_out and F_in is only read
  const double* src = F_in;
  for (int i = 1; i <= nsweeps; ++i) {
    double* dst = ((nsweeps - i) % 2 == 0) ? F_out : Fs;
    const bool last = i == nsweeps;
    if (last) hipLaunchKernelGGL(na_zero_slots, dim3(1), dim3(64), 0, c->stream, slots);
    hipLaunchKernelGGL(na_sweep, grid, block, 0, c->stream, n, rowptr, nbr, D, len, src, dst,
                       last ? g_out : nullptr, last ? slots : nullptr);
    F2M_LAUNCH_CHECK(c);
    src = dst;
  }
  if (max_abs_g) {
    if (is_device_ptr(max_abs_g)) {
      hipLaunchKernelGGL(na_max_out, dim3(1), dim3(64), 0, c->stream, slots, max_abs_g);
      F2M_LAUNCH_CHECK(c);
    } else {
      double* word = reinterpret_cast<double*>(slots + NA_WORD);  // between the slots
      hipLaunchKernelGGL(na_max_out, dim3(1), dim3(64), 0, c->stream, slots, word);
      F2M_LAUNCH_CHECK(c);
      F2M_HIP(c, hipMemcpyAsync(max_abs_g, word, 8, hipMemcpyDeviceToHost, c->stream));
      F2M_HIP(c, hipStreamSynchronize(c->stream));
    }
  }
  return F2M_OK;
}
