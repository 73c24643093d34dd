// Decode and select on MI355X: anneal spins -> candidate levels -> lowest ||y - H x||^2 (candidates, then the MMSE
// estimate) -> Gray bit errors.
//
// Reference semantics: mimo.spins_to_levels_delta / spins_to_levels (arpeggio/mimo/mimo.py:185-190, 39-42),
// mimo.dist / select (mimo.py:416-425, first minimum, MMSE estimate last), mimo.levels_to_bits / bit_errors
// (mimo.py:45-51, 428-431), arpx.biterr_per_use (arpeggio/sim/arpx.py:63-66).
// One workgroup per frame. H is staged in LDS with row stride D+1, so the lanes of a wave (one receive row each)
// read distinct banks while the candidate's level is a broadcast.
#include "kernels.h"

namespace arp {

__device__ __forceinline__ int gray_bits(int l, int M) {
  // QPSK: -1 -> 0, +1 -> 1.  16-QAM (MSB first): -3 00, -1 01, +1 11, +3 10
  if (M == 4) return l > 0 ? 1 : 0;
  return ((l > 0) ? 2 : 0) | ((l == 1 || l == -1) ? 1 : 0);
}

// delta = 1: levels = clip(x_g + sum over the P copies, +-lmax) (dDI). delta = 0: RI spins, levels = s (QPSK) or
// 2 s_MSB + s_LSB with spins [MSB (D); LSB (D)] (16-QAM).
__global__ __launch_bounds__(256) void decode_select_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                            const int8_t* __restrict__ xg,
                                                            const int8_t* __restrict__ spins, int R, int P, int D,
                                                            int Mr, int M, int delta, const int8_t* __restrict__ ltrue,
                                                            int8_t* __restrict__ lhat, int32_t* __restrict__ biterr,
                                                            float* __restrict__ dist_out) {
  extern __shared__ float sm[];
  const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int ld = D + 1, C = R + 1;
  float* Hs = sm;                 // Mr x (D+1)
  float* ys = Hs + Mr * ld;       // Mr
  float* Lc = ys + Mr;            // C x D candidate levels
  float* dv = Lc + C * D;         // C distances
  const float* Hb = H + (size_t)b * Mr * D;
  for (int x = tid; x < Mr * D; x += nt) {
    const int r = x / D, c = x - r * D;
    Hs[r * ld + c] = Hb[x];
  }
  for (int x = tid; x < Mr; x += nt) ys[x] = y[(size_t)b * Mr + x];
  const float lmax = M == 4 ? 1.f : 3.f;
  const int N = delta ? P * D : (M == 4 ? D : 2 * D);
  for (int x = tid; x < C * D; x += nt) {
    const int r = x / D, j = x - r * D;
    float l;
    if (r == R) {
      l = (float)xg[(size_t)b * D + j];
    } else {
      const int8_t* s = spins + ((size_t)b * R + r) * N;
      if (delta) {
        int d = 0;
        for (int p = 0; p < P; ++p) d += s[p * D + j];
        l = fminf(fmaxf((float)(xg[(size_t)b * D + j] + d), -lmax), lmax);
      } else {
        l = M == 4 ? (float)s[j] : (float)(2 * s[j] + s[D + j]);
      }
    }
    Lc[x] = l;
  }
  __syncthreads();
  const int lane = tid & 63, wv = tid >> 6, W = nt >> 6;
  for (int c = wv; c < C; c += W) {
    const float* lc = Lc + c * D;
    float part = 0.f;
    for (int i = lane; i < Mr; i += 64) {
      float s = 0.f;
      for (int j = 0; j < D; ++j) s += Hs[i * ld + j] * lc[j];
      const float e = ys[i] - s;
      part += e * e;
    }
    for (int o = 32; o; o >>= 1) part += __shfl_xor(part, o);
    if (lane == 0) dv[c] = part;
  }
  __syncthreads();
  int best = 0;
  float bd = dv[0];
  for (int c = 1; c < C; ++c)
    if (dv[c] < bd) {
      bd = dv[c];
      best = c;
    }
  if (dist_out)
    for (int c = tid; c < C; c += nt) dist_out[(size_t)b * C + c] = dv[c];
  int errs = 0;
  for (int j = tid; j < D; j += nt) {
    const int l = (int)Lc[best * D + j];
    if (lhat) lhat[(size_t)b * D + j] = (int8_t)l;
    if (ltrue) errs += __builtin_popcount(gray_bits(l, M) ^ gray_bits(ltrue[(size_t)b * D + j], M));
  }
  if (biterr && ltrue) {
    for (int o = 32; o; o >>= 1) errs += __shfl_xor(errs, o);
    __syncthreads();
    if (lane == 0) dv[wv] = (float)errs;
    __syncthreads();
    if (tid == 0) {
      int t = 0;
      for (int v = 0; v < W; ++v) t += (int)dv[v];
      biterr[b] = t;
    }
  }
}

size_t decode_lds_bytes(int R, int D, int Mr) {
  return sizeof(float) * ((size_t)Mr * (D + 1) + Mr + (size_t)(R + 1) * D + (R + 1 > 4 ? R + 1 : 4));
}

hipError_t launch_decode_select(const float* H, const float* y, const int8_t* xg, const int8_t* spins, int B, int R,
                                int P, int D, int Mr, int M, int delta, const int8_t* ltrue, int8_t* lhat,
                                int32_t* biterr, float* dist_out, hipStream_t st) {
  const size_t lds = decode_lds_bytes(R, D, Mr);
  if (lds > 65536)
    hipFuncSetAttribute((const void*)decode_select_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(decode_select_kernel, dim3(B), dim3(256), lds, st, H, y, xg, spins, R, P, D, Mr, M, delta, ltrue,
                     lhat, biterr, dist_out);
  return hipGetLastError();
}

}  // namespace arp
