// f2m_backward.hip -- gradient of the F2M bound w.r.t. the point coordinates
// (new; the reference has no differentiable form).
//
// V(p) = min_x sum_e x_e c_e(p) over the fixed kNN graph.  At a point where
// the optimum x* is unique and the graph does not change, the envelope
// theorem gives dV/dc_e = x*_e, so
//     dV/dp_u = sum_{e = (u,w)} x*_e (p_u - p_w) / |p_u - p_w| .
// One thread per node gathers over its own slots (no atomics, fixed order:
// deterministic), in fp32, scaled by the upstream gradient of its instance.
// Algorithmic bytes: per slot 4 (nbr) + 8 (x) + 8 (p_w gather), per node
// 8 (p_u) + 8 (grad) + rowptr.
#include <vector>

#include "f2m_internal.h"

namespace f2m {
namespace {

template <class RP>
__global__ void g_backward(int64_t n_total, int n_inst, const int64_t* __restrict__ node_off,
                           const RP* __restrict__ rowptr, const int32_t* __restrict__ nbr,
                           const double* __restrict__ x, const float* __restrict__ px,
                           const float* __restrict__ py, const float* __restrict__ scale,
                           float* __restrict__ gx, float* __restrict__ gy) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n_total) return;
  int inst = 0;
  int64_t base = 0;
  if (node_off) {  // largest i with node_off[i] <= u
    int lo = 0, hi = n_inst;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (node_off[mid] <= u) lo = mid; else hi = mid;
    }
    inst = lo;
    base = node_off[lo];
  }
  const float pux = px[u], puy = py[u];
  float ax = 0.f, ay = 0.f;
  for (int64_t s = rowptr[u]; s < rowptr[u + 1]; ++s) {
    const double xe = x[s];
    if (xe == 0.0) continue;
    const int64_t w = base + nbr[s];
    const float dx = pux - px[w], dy = puy - py[w];
    const float r = sqrtf(dx * dx + dy * dy);
    if (r > 0.f) {
      const float xf = (float)xe;
      ax += xf * (dx / r);
      ay += xf * (dy / r);
    }
  }
  const float sc = scale ? scale[inst] : 1.f;
  gx[u] = ax * sc;
  gy[u] = ay * sc;
}

}  // namespace
}  // namespace f2m

using namespace f2m;

extern "C" f2m_status f2m_bound_backward(f2m_ctx* c, const f2m_batch* b, const double* x_slot, const float* px,
                                         const float* py, const float* grad_scale, float* gx, float* gy) {
  if (!c || !x_slot || !px || !py || !gx || !gy) return F2M_ERR_ARG;
  F2M_HIP(c, hipSetDevice(c->device));
  if (!is_device_ptr(x_slot) || !is_device_ptr(px) || !is_device_ptr(py) || !is_device_ptr(gx) ||
      !is_device_ptr(gy) || (grad_scale && !is_device_ptr(grad_scale)))
    return fail(c, F2M_ERR_ARG, "f2m_bound_backward takes device arrays");
  if (b) {
    hipLaunchKernelGGL(g_backward<int64_t>, dim3(grid_for(b->n_total, 256)), dim3(256), 0, c->stream, b->n_total,
                       b->n_inst, b->node_off, b->rowptr, b->nbr, x_slot, px, py, grad_scale, gx, gy);
  } else {
    // the context's graph from f2m_knn_build / f2m_graph_from_csr
    if (c->last_n <= 0) return fail(c, F2M_ERR_ARG, "no graph solved on this context");
    hipLaunchKernelGGL(g_backward<int32_t>, dim3(grid_for(c->last_n, 256)), dim3(256), 0, c->stream, c->last_n, 1,
                       nullptr, c->g_rowptr.as<int32_t>(), c->g_nbr.as<int32_t>(), x_slot, px, py, grad_scale, gx,
                       gy);
  }
  F2M_LAUNCH_CHECK(c);
  return F2M_OK;
}
