// Fused per-frame front end for MI355X: LMMSE hard decision + delta-Ising (dDI) build + fixed-point quantisation.
//
// Reference semantics (float, then rounded): mimo.mmse / _mu / mmse_hard / quantize_levels (arpeggio/mimo/mimo.py:86-109,
// 176-182), mimo.ising_delta_at (mimo.py:506-523), chordfx.to_fixed (arpeggio/mimo/chordfx.py:60-76) and
// chord_ref.compress (arpeggio/sim/x3model/chord_ref.py:47-59). One workgroup (4 waves) per frame:
//   1. G = H^T H on the matrix cores: v_mfma_f32_16x16x4f32 over the upper 16x16 tiles of G, H staged in LDS and
//      the lower triangle mirrored (G exactly symmetric); c = H^T y. The Gram is formed again in step 4 rather than
//      kept, so LDS holds H and one D x D matrix (D = 128: 130 KiB).
//   2. A = G + sigma I, sigma = (N0/2)/E[l^2]; Cholesky A = L L^T and X = L^-1 in place in LDS.
//      z = A^-1 c = X^T (X c); mu_i = diag(A^-1 G)_i = 1 - sigma (A^-1)_ii = 1 - sigma sum_k X_ki^2.
//   3. x_g = nearest level of z / mu (ties to the lower level, as argmin over the level list).
//   4. ybar = y - H x_g, h = 2 H^T ybar, scale = max(max|h|, max|G|) (the dDI coupling matrix holds every G_ij,
//      its diagonal included, on the cross-copy blocks), clamped at 1e-12.
//   5. G_int = rne(G / scale * Q) (G symmetric by construction), h_int = rne(h / scale * Q); optionally the float
//      scaled G and h for the FP32 annealers, and the expanded Chord inputs J (N x N int16, N = 2K D) and h (N).
// Sizes: D <= 128 (2 Nt), Mr <= 128 (2 Nr); both are zero-padded inside LDS to multiples of 16 and 4.
#include "kernels.h"

namespace arp {

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float block_max(float v, float* red) {
  for (int o = 32; o; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[wv] = v;
  __syncthreads();
  float m = red[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fmaxf(m, red[i]);
  __syncthreads();
  return m;
}

// Gram tile loop: G = H^T H into dst (row stride ld) for the Dp x Dp padded problem, then the upper triangle is
// mirrored so that G is exactly symmetric. Tile (I, J) is 16 x 16; lane l feeds A[i = l%16][k = l/16] = H[k][16I + i]
// and B[k = l/16][j = l%16] = H[k][16J + j]; the accumulator holds G[16I + 4(l/16) + v][16J + l%16].
__device__ __forceinline__ void gram_mfma(const float* Hs, int Mp, int Dp, float* dst, int ld) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, W = blockDim.x >> 6;
  const int nT = Dp / 16;
  const int li = lane & 15, lk = lane >> 4;
  for (int tIdx = wv; tIdx < nT * nT; tIdx += W) {
    const int I = tIdx / nT, J = tIdx - I * nT;
    if (J < I) continue;  // lower tiles come from the mirror
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < Mp; kk += 4) {
      const float a = Hs[(kk + lk) * Dp + 16 * I + li];
      const float bb = Hs[(kk + lk) * Dp + 16 * J + li];
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bb, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) dst[(16 * I + 4 * lk + v) * ld + 16 * J + li] = acc[v];
  }
  __syncthreads();
  for (int x = tid; x < Dp * Dp; x += blockDim.x) {
    const int i = x / Dp, j = x - i * Dp;
    if (i > j) dst[i * ld + j] = dst[j * ld + i];
  }
  __syncthreads();
}

// LDS (floats): Hs [Mp x Dp] | As [Dp x (Dp+1)] | ys [Mp] | 6 vectors of Dp | red 8
__global__ __launch_bounds__(256) void front_end_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                        const float* __restrict__ N0v, float N0s, int Mr, int D, int M,
                                                        int K, float Q, FrontOut o) {
  extern __shared__ float sm[];
  const int b = blockIdx.x;
  const int Dp = (D + 15) & ~15, Mp = (Mr + 3) & ~3, Ap = Dp + 1;
  float* Hs = sm;
  float* As = Hs + Mp * Dp;
  float* ys = As + Dp * Ap;
  float* cv = ys + Mp;   // H^T y
  float* uv = cv + Dp;   // X c
  float* xv = uv + Dp;   // x_g (float levels)
  float* hv = xv + Dp;   // h
  float* tv = hv + Dp;   // saved column of L
  float* red = tv + 2 * Dp;
  const int tid = threadIdx.x, nt = blockDim.x;
  const float* Hb = H + (size_t)b * Mr * D;
  for (int x = tid; x < Mp * Dp; x += nt) {
    const int r = x / Dp, c = x - r * Dp;
    Hs[x] = (r < Mr && c < D) ? Hb[(size_t)r * D + c] : 0.f;
  }
  for (int x = tid; x < Mp; x += nt) ys[x] = x < Mr ? y[(size_t)b * Mr + x] : 0.f;
  __syncthreads();
  // ---- 1. Gram on the matrix cores, c = H^T y
  gram_mfma(Hs, Mp, Dp, As, Ap);
  for (int j = tid; j < Dp; j += nt) {
    float s = 0.f;
    for (int k = 0; k < Mp; ++k) s += Hs[k * Dp + j] * ys[k];
    cv[j] = s;
  }
  // ---- 2. A = G + sigma I (identity on the padding), Cholesky A = L L^T, X = L^-1
  const float el2 = M == 4 ? 1.f : 5.f;
  const float n0 = N0v ? N0v[b] : N0s;
  const float sigma = (n0 / 2.f) / el2;
  for (int i = tid; i < Dp; i += nt) As[i * Ap + i] = i < D ? As[i * Ap + i] + sigma : 1.f;
  for (int k = 0; k < D; ++k) {
    __syncthreads();
    const float dk = sqrtf(As[k * Ap + k]);
    for (int i = k + 1 + tid; i < D; i += nt) As[i * Ap + k] = As[i * Ap + k] / dk;
    __syncthreads();
    if (tid == 0) As[k * Ap + k] = dk;
    const int m = D - k - 1;
    for (int x = tid; x < m * m; x += nt) {
      const int i = k + 1 + x / m, j = k + 1 + x % m;
      if (j <= i) As[i * Ap + j] -= As[i * Ap + k] * As[j * Ap + k];
    }
  }
  __syncthreads();
  // columns right to left: X_jj = 1/L_jj, X_ij = -(sum_{j<k<=i} X_ik L_kj) / L_jj for i > j (from X L = I)
  for (int j = D - 1; j >= 0; --j) {
    for (int i = j + 1 + tid; i < D; i += nt) tv[i] = As[i * Ap + j];
    __syncthreads();
    const float ljj = As[j * Ap + j];
    for (int i = j + 1 + tid; i < D; i += nt) {
      float s = 0.f;
      for (int k = j + 1; k <= i; ++k) s += As[i * Ap + k] * tv[k];
      As[i * Ap + j] = -s / ljj;
    }
    __syncthreads();
    if (tid == 0) As[j * Ap + j] = 1.f / ljj;
  }
  __syncthreads();
  for (int i = tid; i < D; i += nt) {
    float s = 0.f;
    for (int k = 0; k <= i; ++k) s += As[i * Ap + k] * cv[k];
    uv[i] = s;
  }
  __syncthreads();
  // ---- 3. z = X^T u, mu_i = 1 - sigma sum_k X_ki^2, x_g = nearest level of z/mu (ties to the lower level)
  const float lmax = M == 4 ? 1.f : 3.f;
  for (int i = tid; i < D; i += nt) {
    float s = 0.f, d = 0.f;
    for (int k = i; k < D; ++k) {
      const float xk = As[k * Ap + i];
      s += xk * uv[k];
      d += xk * xk;
    }
    const float mu = 1.f - sigma * d;
    const float v = s / mu;
    float l = 2.f * ceilf(v * 0.5f) - 1.f;
    l = fminf(fmaxf(l, -lmax), lmax);
    if (v != v) l = -lmax;
    xv[i] = l;
    if (o.z) o.z[(size_t)b * D + i] = s;
    if (o.mu) o.mu[(size_t)b * D + i] = mu;
    if (o.xg) o.xg[(size_t)b * D + i] = (int8_t)l;
  }
  __syncthreads();
  // ---- 4. ybar = y - H x_g, h = 2 H^T ybar; the Gram again (As is free now) for scale and the fixed point
  for (int k = tid; k < Mr; k += nt) {
    float s = 0.f;
    for (int j = 0; j < D; ++j) s += Hs[k * Dp + j] * xv[j];
    ys[k] = ys[k] - s;
  }
  __syncthreads();
  gram_mfma(Hs, Mp, Dp, As, Ap);
  float mloc = 0.f;
  for (int j = tid; j < D; j += nt) {
    float s = 0.f;
    for (int k = 0; k < Mr; ++k) s += Hs[k * Dp + j] * ys[k];
    hv[j] = 2.f * s;
    mloc = fmaxf(mloc, fabsf(2.f * s));
  }
  for (int x = tid; x < D * D; x += nt) {
    const int i = x / D, j = x - i * D;
    mloc = fmaxf(mloc, fabsf(As[i * Ap + j]));
  }
  const float scale = fmaxf(block_max(mloc, red), 1e-12f);
  if (tid == 0 && o.scale) o.scale[b] = scale;
  // ---- 5. fixed point and float outputs
  const int P = 2 * K, N = P * D;
  for (int x = tid; x < D * D; x += nt) {
    const int i = x / D, j = x - i * D;
    const float gs = As[i * Ap + j] / scale;
    if (o.Gi) o.Gi[(size_t)b * D * D + x] = (int32_t)rintf(gs * Q);
    if (o.Gf) o.Gf[(size_t)b * D * D + x] = gs;
  }
  for (int j = tid; j < D; j += nt) {
    const float hs = hv[j] / scale;
    const int32_t h32 = (int32_t)rintf(hs * Q);
    if (o.hi) o.hi[(size_t)b * D + j] = h32;
    if (o.hf) o.hf[(size_t)b * D + j] = hs;
    if (o.hN)
      for (int p = 0; p < P; ++p) o.hN[(size_t)b * N + p * D + j] = h32;
    if (o.hNf)
      for (int p = 0; p < P; ++p) o.hNf[(size_t)b * N + p * D + j] = hs;
  }
  if (o.Jf) {
    for (int x = tid; x < N * N; x += nt) {
      const int r = x / N, c = x - r * N;
      o.Jf[(size_t)b * N * N + x] = r == c ? 0.f : -(As[(r % D) * Ap + (c % D)] / scale);
    }
  }
  if (o.J) {
    // J[(p,i),(p',j)] = -G_int[i][j] off the diagonal, 0 on it (chord_ref.expand)
    for (int x = tid; x < N * N; x += nt) {
      const int r = x / N, c = x - r * N;
      const int i = r % D, j = c % D;
      const int32_t gi = (int32_t)rintf(As[i * Ap + j] / scale * Q);
      o.J[(size_t)b * N * N + x] = r == c ? (int16_t)0 : (int16_t)(-gi);
    }
  }
}

size_t front_end_lds_bytes(int Mr, int D) {
  const int Dp = (D + 15) & ~15, Mp = (Mr + 3) & ~3;
  return sizeof(float) * ((size_t)Mp * Dp + (size_t)Dp * (Dp + 1) + Mp + 6 * Dp + 8);
}

hipError_t launch_front_end(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M,
                            int K, float Q, const FrontOut& o, hipStream_t st) {
  const size_t lds = front_end_lds_bytes(Mr, D);
  if (lds > 65536) hipFuncSetAttribute((const void*)front_end_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(front_end_kernel, dim3(B), dim3(256), lds, st, H, y, N0v, N0s, Mr, D, M, K, Q, o);
  return hipGetLastError();
}

}  // namespace arp
