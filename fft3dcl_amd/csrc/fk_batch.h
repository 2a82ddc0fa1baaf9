// Batched launches of the serving step (fft3d_step): one launch per kernel
// class over the work of many planes, each plane's share a contiguous run of
// the launch's workgroups described by a BatchItem in device memory. A
// workgroup finds its item by binary search over the items' first
// workgroups, copies the item's arguments and runs the same body as the
// single-plane kernel (analysis_body, analysis_mfma_body, sweep_body,
// combine_body). Every item starts at a multiple of 8 workgroups, so
// workgroup i of an item runs on XCD i % 8, as in its own launch.
//
// The reference enqueues fft2d_fwd, filter, fft2d_inv and synthesize per
// plane and frame (src/cl/gpu.cpp:308-324, :389-458).
#pragma once

#include "fk_sweep.h"

namespace fk {

struct BatchItem {
  PlaneArgs P;
  WindowArgs Wn;   // sweep
  SweepArgs A;     // sweep, combine
  const void* plane;  // analysis: the source plane
  c32* spec;          // analysis: the ring spectrum it writes
  float4* kstate;     // sweep: Kalman state (bm == 0)
  int kfirst;         // sweep: first Kalman frame
  int wg0, nwg;       // the item's workgroups [wg0, wg0 + nwg); wg0 % 8 == 0
};

__device__ __forceinline__ int batch_item(const BatchItem* __restrict__ items, int n, int wg) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].wg0 <= wg) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int BW, int BH, bool H>
__global__ __launch_bounds__(64) void k_analysis_batch(const BatchItem* __restrict__ items,
                                                       int n) {
  const BatchItem* it = items + batch_item(items, n, blockIdx.x);
  const int local = blockIdx.x - it->wg0;
  if (local >= it->nwg) return;  // padding to the next item's multiple of 8
  const PlaneArgs P = it->P;
  analysis_body<BW, BH, H>(P, it->plane, it->spec, local, it->nwg);
}

template <int PR, bool H>
__global__ __launch_bounds__(256) void k_analysis_mfma_batch(const BatchItem* __restrict__ items,
                                                            int n) {
  const BatchItem* it = items + batch_item(items, n, blockIdx.x);
  const int local = blockIdx.x - it->wg0;
  if (local >= it->nwg) return;
  const PlaneArgs P = it->P;
  analysis_mfma_body<PR, H>(P, it->plane, it->spec, local, it->nwg);
}

template <int BW, int BH, int WPW, bool H, int DFT, int MINB = 2>
__global__ __launch_bounds__(64 * WPW, MINB) void k_sweep_batch(
    const BatchItem* __restrict__ items, int n) {
  const BatchItem* it = items + batch_item(items, n, blockIdx.x);
  const int local = blockIdx.x - it->wg0;
  if (local >= it->nwg) return;
  const PlaneArgs P = it->P;
  const WindowArgs Wn = it->Wn;
  const SweepArgs A = it->A;
  sweep_body<BW, BH, WPW, false, H, DFT>(P, Wn, it->kstate, it->kfirst, A, local % A.nstrips,
                                         local / A.nstrips);
}

#ifdef FK_COMMON_KERNELS
__global__ __launch_bounds__(256) void k_combine_batch(const BatchItem* __restrict__ items, int n) {
  const BatchItem* it = items + batch_item(items, n, blockIdx.x);
  const int local = blockIdx.x - it->wg0;
  if (local >= it->nwg) return;
  const PlaneArgs P = it->P;
  const SweepArgs A = it->A;
  combine_body(P, A, local * 4 + (threadIdx.x >> 6));
}
#endif

}  // namespace fk
