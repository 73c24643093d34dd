// RI-MIMO Ising build on MI355X: the regularised-Ising mapping of the MMSE-initialised problem and its fixed point.
//
// Reference semantics: mimo.ising (arpeggio/mimo/mimo.py:132-149) followed by chordfx.to_fixed's J, h, h0
// (arpeggio/mimo/chordfx.py:60-76), as fxdetect.detect uses them for its 'ri' schedules (arpeggio/mimo/fxdetect.py:17-
// 18, 28-30). Levels are l = T s: T = I for QPSK (N = D spins), T = [2 I, I] for 16-QAM (N = 2 D spins, the MSB block
// first). With G = H^T H and c = H^T y:
//   J = -offdiag(T^T G T) / scale            (blocks t_p t_p' G, t = (2, 1); the diagonal alone is zeroed)
//   h0 = 2 T^T c / scale,  h = h0 + 2 r s_P  (s_P: the spins of the MMSE output's nearest levels, ties to the lower)
//   scale = max(max|h|, max|J|), clamped at 1e-12  (over T^T G T's off-diagonal: 4 max_{i != j}|G_ij| and
//           2 max_j G_jj for 16-QAM, max_{i != j}|G_ij| for QPSK)
// and the integers are the half-even rounding of J Q (int16), h Q and h0 Q (int32).
//
// One workgroup of 256 threads per frame: H and y staged in LDS (row stride D + 1), the Gram's upper triangle by
// per-thread dot products over the receive rows (mirrored into the full D x D in LDS), the maxima by a workgroup
// reduction, then the N x N couplings written row-major, one entry per thread per step.
#include "kernels.h"
#include "adaptive.h"

#include <cstdint>

namespace arp {

__global__ __launch_bounds__(256) void ising_ri_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                       const float* __restrict__ z, int Mr, int D, int M, float r,
                                                       float Q, int16_t* __restrict__ J, int32_t* __restrict__ h,
                                                       int32_t* __restrict__ h0, float* __restrict__ scale_out) {
  extern __shared__ float sm[];
  const int b = blockIdx.x, t = threadIdx.x, nt = blockDim.x, lane = t & 63, wv = t >> 6, W = nt >> 6;
  const int ld = D + 1;
  float* Hs = sm;               // Mr x (D + 1)
  float* ys = Hs + Mr * ld;     // Mr
  float* Gs = ys + Mr;          // D x D
  float* cs = Gs + D * D;       // D
  float* red = cs + D;          // 3 W: per-wave maxima
  const float* Hb = H + (size_t)b * Mr * D;
  for (int x = t; x < Mr * D; x += nt) {
    const int rr = x / D, cc = x - rr * D;
    Hs[rr * ld + cc] = Hb[x];
  }
  for (int x = t; x < Mr; x += nt) ys[x] = y[(size_t)b * Mr + x];
  __syncthreads();
  float offmax = 0.f, diagmax = 0.f, cmax = 0.f;
  for (int x = t; x < D * D; x += nt) {
    const int i = x / D, j = x - i * D;
    if (i > j) continue;
    float s = 0.f;
    for (int rr = 0; rr < Mr; ++rr) s = __builtin_fmaf(Hs[rr * ld + i], Hs[rr * ld + j], s);
    Gs[i * D + j] = s;
    Gs[j * D + i] = s;
    if (i == j) diagmax = fmaxf(diagmax, fabsf(s));
    else offmax = fmaxf(offmax, fabsf(s));
  }
  for (int j = t; j < D; j += nt) {
    float s = 0.f;
    for (int rr = 0; rr < Mr; ++rr) s = __builtin_fmaf(Hs[rr * ld + j], ys[rr], s);
    cs[j] = s;
    cmax = fmaxf(cmax, fabsf(s));
  }
  for (int o = 32; o; o >>= 1) {
    offmax = fmaxf(offmax, __shfl_xor(offmax, o));
    diagmax = fmaxf(diagmax, __shfl_xor(diagmax, o));
    cmax = fmaxf(cmax, __shfl_xor(cmax, o));
  }
  if (lane == 0) {
    red[3 * wv] = offmax;
    red[3 * wv + 1] = diagmax;
    red[3 * wv + 2] = cmax;
  }
  __syncthreads();
  offmax = diagmax = cmax = 0.f;
  for (int w = 0; w < W; ++w) {
    offmax = fmaxf(offmax, red[3 * w]);
    diagmax = fmaxf(diagmax, red[3 * w + 1]);
    cmax = fmaxf(cmax, red[3 * w + 2]);
  }
  const bool qam = M == 16;
  const int N = qam ? 2 * D : D;
  const float tmax = qam ? 2.f : 1.f;
  const float jmax = qam ? fmaxf(4.f * offmax, 2.f * diagmax) : offmax;
  const float scale = fmaxf(fmaxf(2.f * tmax * cmax, jmax), 1e-12f);
  if (t == 0 && scale_out) scale_out[b] = scale;
  // fields: spin a = p D + i (p = 0: MSB or the QPSK spin, p = 1: LSB), t_p = 2 or 1
  const float lmax = qam ? 3.f : 1.f;
  for (int a = t; a < N; a += nt) {
    const int p = a / D, i = a - p * D;
    const float tp = qam ? (p == 0 ? 2.f : 1.f) : 1.f;
    const float hf = 2.f * tp * cs[i] / scale;
    const float zz = z[(size_t)b * D + i];
    float lq = fminf(fmaxf(2.f * ceilf(zz * 0.5f) - 1.f, -lmax), lmax);
    if (zz != zz) lq = -lmax;
    const float sp = !qam ? lq : p == 0 ? (lq > 0.f ? 1.f : -1.f) : ((lq == 3.f || lq == -1.f) ? 1.f : -1.f);
    h0[(size_t)b * N + a] = (int32_t)rintf(hf * Q);
    h[(size_t)b * N + a] = (int32_t)rintf((hf + 2.f * r * sp) * Q);
  }
  int16_t* Jb = J + (size_t)b * N * N;
  for (int x = t; x < N * N; x += nt) {
    const int a = x / N, c = x - a * N;
    const int p = a / D, i = a - p * D, p2 = c / D, j = c - p2 * D;
    const float tp = qam ? (p == 0 ? 2.f : 1.f) : 1.f, tq = qam ? (p2 == 0 ? 2.f : 1.f) : 1.f;
    const float v = a == c ? 0.f : -(tp * tq * Gs[i * D + j]) / scale;
    Jb[x] = (int16_t)rintf(v * Q);
  }
}

size_t ising_ri_lds_bytes(int Mr, int D) {
  return sizeof(float) * ((size_t)Mr * (D + 1) + Mr + (size_t)D * D + D + 12);
}

}  // namespace arp

extern "C" int arp_ising_ri(const float* H, const float* y, const float* z, int B, int Mr, int D, int M, float r,
                            float Q, int16_t* J, int32_t* h, int32_t* h0, float* scale, void* stream) {
  if (D < 1 || D > 128 || Mr < 1 || Mr > 128) return arp::set_error(ARP_E_RANGE, "ising_ri: D or Mr outside 1..128");
  if (M != 4 && M != 16) return arp::set_error(ARP_E_ARG, "ising_ri: M must be 4 or 16");
  if (M == 16 && 2 * D > 256) return arp::set_error(ARP_E_RANGE, "ising_ri: N = 2D beyond 256 spins");
  if (!(Q > 0.f) || Q > 16384.f) return arp::set_error(ARP_E_ARG, "ising_ri: Q outside (0, 2^14]");
  if (B < 0) return arp::set_error(ARP_E_ARG, "ising_ri: B < 0");
  if (B == 0) return arp::set_error(0, "");
  if (!H || !y || !z || !J || !h || !h0) return arp::set_error(ARP_E_ARG, "ising_ri: null pointer");
  const size_t lds = arp::ising_ri_lds_bytes(Mr, D);
  if (lds > 160 * 1024) return arp::set_error(ARP_E_RANGE, "ising_ri: frame does not fit the LDS");
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)arp::ising_ri_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(arp::ising_ri_kernel, dim3(B), dim3(256), lds, (hipStream_t)stream, H, y, z, Mr, D, M, r, Q, J, h,
                     h0, scale);
  return arp::hip_status(hipGetLastError(), "ising_ri launch");
}
