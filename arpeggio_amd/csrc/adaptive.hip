// Early-stop certificate of the adaptive detector: a frame whose chosen candidate already explains y within the noise,
// ||y - H x||^2 <= N0 (Nr + c sqrt(Nr)), is final; the others are appended to a frame list that the next, longer anneal
// and decode launches read (see AnnealExtra / DecodeExtra).
//
// The stopping rule is the reference's: a first-pass stop on t = ||y - H x||^2 / N0 at the threshold Nr + c sqrt(Nr)
// (arpeggio/sim/a56an.py:58-67, `stop`), with c chosen so that the paired BER ratio against no stop stays below a
// cut-off (a56an.py:109-118, `pick_c`). Here it gates a second anneal stage instead of further MDI stages.
//
// One thread per frame; each wave reserves its slots with one atomic add and writes its uncertified frames in lane
// order (the order across waves is not fixed; every later stage is keyed by the frame index, not by the slot).
#include "adaptive.h"

namespace arp {

__global__ void certify_reset_kernel(int32_t* count) { *count = 0; }

__global__ __launch_bounds__(256) void certify_kernel(const float* __restrict__ best_d, const float* __restrict__ N0v,
                                                      float N0s, float thr_scale, int B, int32_t* __restrict__ list,
                                                      int32_t* __restrict__ count) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  bool open = false;
  if (b < B) {
    const float n0 = N0v ? N0v[b] : N0s;
    open = !(best_d[b] <= n0 * thr_scale);  // NaN distances stay open
  }
  const uint64_t m = ballot(open);
  if (!m) return;
  const int lane = lane_id();
  int base = 0;
  if (lane == 0) base = atomicAdd(count, __builtin_popcountll(m));
  base = __shfl(base, 0);
  if (open) list[base + __builtin_popcountll(m & ((1ull << lane) - 1ull))] = b;
}

hipError_t launch_certify(const float* best_d, const float* N0v, float N0s, float thr_scale, int B, int32_t* list,
                          int32_t* count, hipStream_t st) {
  hipLaunchKernelGGL(certify_reset_kernel, dim3(1), dim3(1), 0, st, count);
  if (B > 0)
    hipLaunchKernelGGL(certify_kernel, dim3((B + 255) / 256), dim3(256), 0, st, best_d, N0v, N0s, thr_scale, B, list,
                       count);
  return hipGetLastError();
}

}  // namespace arp
