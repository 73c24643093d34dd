// Measurement probes (no reference counterpart): the chip facts the kernels' bounds are stated against.
//   arp_probe_lds: LDS read bandwidth. Every workgroup of 256 threads fills 16 KiB of LDS, then each thread reads
//     16-byte words at a stride that keeps the 64 lanes of a read on distinct banks, `iters` times; one float per
//     thread goes out so that the reads are not removed.
//   arp_probe_issue: whether FP32 MFMA and independent vector FMAs overlap on a SIMD. mode 0: 8 independent
//     v_mfma_f32_16x16x4f32 per iteration; mode 1: 64 independent FMAs per iteration; mode 2: both in each wave.
#include "kernels.h"
#include "adaptive.h"

namespace arp {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void probe_lds_kernel(float* __restrict__ out, int iters) {
  __shared__ __align__(16) float buf[4096];
  const int t = threadIdx.x;
  for (int i = t; i < 4096; i += 256) buf[i] = (float)(i & 63);
  __syncthreads();
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  const floatx4* b4 = reinterpret_cast<const floatx4*>(buf);
  int idx = t;  // 1,024 16-byte words: thread t starts at word t and steps by 256 words (wraps within the buffer)
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const floatx4 v = b4[(idx + 128 * u) & 1023];
      acc += v;
    }
    idx = (idx + 64) & 1023;
  }
  out[blockIdx.x * 256 + t] = acc[0] + acc[1] + acc[2] + acc[3];
}

template <int MODE>
__global__ __launch_bounds__(256) void probe_issue_kernel(float* __restrict__ out, int iters, float a, float b) {
  floatx4 acc[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) acc[u] = floatx4{0.f, 0.f, 0.f, 0.f};
  float f[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) f[u] = a * (float)(threadIdx.x + u);
  for (int it = 0; it < iters; ++it) {
    if (MODE != 1) {
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[u], 0, 0, 0);
    }
    if (MODE != 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int u = 0; u < 16; ++u) f[u] = __builtin_fmaf(f[u], b, a);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int u = 0; u < 8; ++u) s += acc[u][0] + acc[u][1] + acc[u][2] + acc[u][3];
#pragma unroll
  for (int u = 0; u < 16; ++u) s += f[u];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace
}  // namespace arp

// out: workgroups * 256 floats
extern "C" int arp_probe_lds(float* out, int workgroups, int iters, void* stream) {
  if (!out || workgroups < 1 || iters < 1) return arp::set_error(ARP_E_ARG, "probe_lds: out, workgroups or iters");
  hipLaunchKernelGGL(arp::probe_lds_kernel, dim3(workgroups), dim3(256), 0, (hipStream_t)stream, out, iters);
  return arp::hip_status(hipGetLastError(), "probe_lds launch");
}

// out: workgroups * 256 floats
extern "C" int arp_probe_issue(int mode, float* out, int workgroups, int iters, void* stream) {
  if (!out || workgroups < 1 || iters < 1 || mode < 0 || mode > 2)
    return arp::set_error(ARP_E_ARG, "probe_issue: mode, out, workgroups or iters");
  const hipStream_t st = (hipStream_t)stream;
  if (mode == 0) hipLaunchKernelGGL(arp::probe_issue_kernel<0>, dim3(workgroups), dim3(256), 0, st, out, iters, 1.0001f, 0.9999f);
  else if (mode == 1) hipLaunchKernelGGL(arp::probe_issue_kernel<1>, dim3(workgroups), dim3(256), 0, st, out, iters, 1.0001f, 0.9999f);
  else hipLaunchKernelGGL(arp::probe_issue_kernel<2>, dim3(workgroups), dim3(256), 0, st, out, iters, 1.0001f, 0.9999f);
  return arp::hip_status(hipGetLastError(), "probe_issue launch");
}
