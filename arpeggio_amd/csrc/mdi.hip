// Kernels of the multi-stage delta-Ising (MDI) wrapper on MI355X: ordered MMSE-SIC guesses, the search radius, and
// the fixed-point dDI sub-problem build around a guess. The stages run as kernels on device buffers; the host only
// reads the per-frame radii once per stage to group the sub-problems as the reference does.
//
// Reference semantics (paths relative to the reference repository):
//   mmse_sic      arpeggio/mimo/mimo.py:466-503  (stream with the smallest MMSE error variance first, per real stream or
//                                                 per user; quantise z/mu, cancel, repeat)
//   mdi_radius    arpeggio/mimo/mimo.py:526-543  ('mean' heuristic: lambda = mean of diag(H^T H))
//   sub-problem   mimo.ising_delta_at (mimo.py:506-523) -> chordfx.to_fixed (chordfx.py:60-76) -> chord_ref.compress
//                 (arpeggio/sim/x3model/chord_ref.py:47-59), in the compressed (D x D) form the first-accept kernel takes
//
// MMSE-SIC here inverts A = H^T H + sigma I once (symmetric sweep in LDS) and then removes each detected stream from the
// inverse with a rank-1 downdate, P <- P - P[:, j] P[j, :] / P[j, j] (the inverse of the remaining streams' A), instead
// of the reference's fresh inverse per stage; mu_j = (P G)_jj = 1 - sigma P_jj. In the real lift of a complex channel
// the two streams of a user start with equal error variances, so which of them goes first is decided by rounding, in the
// reference as here.
#include "mdi.h"
#include "kernels.h"

namespace arp {

__device__ __forceinline__ float quantize_level(float v, float lmax) {
  // nearest PAM level (odd integers within +-lmax), the lower one on a tie (the reference's first argmin)
  float l = fminf(fmaxf(2.f * ceilf(v * 0.5f) - 1.f, -lmax), lmax);
  return v != v ? -lmax : l;
}

// Workgroup-wide min with the lowest index among equal minima (threads past `n` pass +inf).
__device__ __forceinline__ int block_argmin(float v, int idx, float* red_v, int* red_i) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, W = blockDim.x >> 6;
  for (int o = 32; o; o >>= 1) {
    const float ov = __shfl_xor(v, o);
    const int oi = __shfl_xor(idx, o);
    if (ov < v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  if (lane == 0) {
    red_v[wv] = v;
    red_i[wv] = idx;
  }
  __syncthreads();
  float bv = red_v[0];
  int bi = red_i[0];
  for (int w = 1; w < W; ++w)
    if (red_v[w] < bv || (red_v[w] == bv && red_i[w] < bi)) {
      bv = red_v[w];
      bi = red_i[w];
    }
  __syncthreads();
  return bi;
}

// One workgroup per frame, thread i owns row i (blockDim a multiple of 64, >= D). LDS: H (Mr x D), P (D x (D+1)),
// c (D), the pivot row buffer (D), activity flags, reductions.
__global__ __launch_bounds__(128) void mmse_sic_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                       const float* __restrict__ N0v, float N0s, int Mr, int D, int M,
                                                       int per_user, int8_t* __restrict__ xs) {
  extern __shared__ float sm[];
  const int b = blockIdx.x, i = threadIdx.x, nt = blockDim.x;
  const int ld = D + 1;
  float* Hs = sm;                  // Mr x D
  float* Pm = Hs + Mr * D;         // D x ld
  float* c = Pm + D * ld;          // D: H^T y of the remaining signal
  float* buf = c + D;              // D: pivot row
  float* red_v = buf + D;          // 2
  int* red_i = (int*)(red_v + 2);  // 2
  int* act = red_i + 2;            // D
  const float* Hb = H + (size_t)b * Mr * D;
  for (int x = i; x < Mr * D; x += nt) Hs[x] = Hb[x];
  const float el2 = M == 4 ? 1.f : 5.f, lmax = M == 4 ? 1.f : 3.f;
  const float sigma = ((N0v ? N0v[b] : N0s) / 2.f) / el2;
  __syncthreads();
  const bool own = i < D;
  if (own) {
    float cs = 0.f;
    for (int r = 0; r < Mr; ++r) cs = __builtin_fmaf(Hs[r * D + i], y[(size_t)b * Mr + r], cs);
    c[i] = cs;
    for (int j = 0; j < D; ++j) {
      float g = 0.f;
      for (int r = 0; r < Mr; ++r) g = __builtin_fmaf(Hs[r * D + i], Hs[r * D + j], g);
      Pm[i * ld + j] = j == i ? g + sigma : g;
    }
    act[i] = 1;
  }
  // in-place symmetric sweep: after D pivots Pm holds -A^-1
  for (int k = 0; k < D; ++k) {
    __syncthreads();
    if (own) buf[i] = Pm[k * ld + i];
    __syncthreads();
    if (own) {
      const float piv = buf[k];
      if (i == k) {
        for (int j = 0; j < D; ++j) Pm[k * ld + j] = j == k ? -1.f / piv : buf[j] / piv;
      } else {
        const float f = Pm[i * ld + k];
        const float fr = f / piv;
        for (int j = 0; j < D; ++j)
          if (j != k) Pm[i * ld + j] = __builtin_fmaf(-fr, buf[j], Pm[i * ld + j]);
        Pm[i * ld + k] = fr;
      }
    }
  }
  __syncthreads();
  if (own)
    for (int j = 0; j < D; ++j) Pm[i * ld + j] = -Pm[i * ld + j];   // P = A^-1 (of the active streams)
  const int Nt = D / 2;
  const int nst = per_user ? Nt : D;
  for (int s = 0; s < nst; ++s) {
    __syncthreads();
    int js[2];
    int nj = 1;
    if (per_user) {
      const float v = (i < Nt && act[i]) ? Pm[i * ld + i] + Pm[(i + Nt) * ld + (i + Nt)] : __builtin_inff();
      const int u = block_argmin(v, i < Nt ? i : 0x7FFFFFFF, red_v, red_i);
      js[0] = u;
      js[1] = u + Nt;
      nj = 2;
    } else {
      const float v = (own && act[i]) ? Pm[i * ld + i] : __builtin_inff();
      js[0] = block_argmin(v, own ? i : 0x7FFFFFFF, red_v, red_i);
    }
    // z_j and mu_j of every selected stream from the same P and c (as the reference forms z once per stage)
    float xj[2];
    for (int q = 0; q < nj; ++q) {
      const int j = js[q];
      float zp = (own && act[i]) ? Pm[j * ld + i] * c[i] : 0.f;
      for (int o = 32; o; o >>= 1) zp += __shfl_xor(zp, o);
      if ((i & 63) == 0) red_v[i >> 6] = zp;
      __syncthreads();
      float z = 0.f;
      for (int w = 0; w < (nt >> 6); ++w) z += red_v[w];
      __syncthreads();
      const float mu = 1.f - sigma * Pm[j * ld + j];
      xj[q] = quantize_level(z / mu, lmax);
    }
    for (int q = 0; q < nj; ++q) {
      const int j = js[q];
      if (i == j) xs[(size_t)b * D + j] = (int8_t)xj[q];
      // cancel stream j from c: c_k -= G_kj x_j
      if (own && act[i]) {
        float g = 0.f;
        for (int r = 0; r < Mr; ++r) g = __builtin_fmaf(Hs[r * D + i], Hs[r * D + j], g);
        c[i] = __builtin_fmaf(-g, xj[q], c[i]);
      }
      // remove j from the inverse: P <- P - P[:, j] P[j, :] / P[j, j]
      __syncthreads();
      if (own) buf[i] = Pm[j * ld + i];
      __syncthreads();
      if (own && act[i] && i != j) {
        const float f = Pm[i * ld + j] / buf[j];
        for (int k = 0; k < D; ++k)
          if (act[k]) Pm[i * ld + k] = __builtin_fmaf(-f, buf[k], Pm[i * ld + k]);
      }
      __syncthreads();
      if (i == j) act[j] = 0;
    }
  }
}

// Search radius and the guess's own distance, one wave per frame: yb = y - H x_g, g = ||H^T yb||, lambda = mean
// diag(H^T H), Gamma = 2 g / lambda, R = omega if 2 omega <= Gamma else floor(Gamma / 2), clipped to [0, Rmax].
__global__ __launch_bounds__(64) void mdi_radius_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                        const int8_t* __restrict__ xg, int Mr, int D, int M, int Rmax,
                                                        int32_t* __restrict__ Rout, float* __restrict__ Gam,
                                                        float* __restrict__ dist) {
  __shared__ float yb[256];
  __shared__ float xs[256];
  const int b = blockIdx.x, lane = threadIdx.x;
  const float* Hb = H + (size_t)b * Mr * D;
  for (int j = lane; j < D; j += 64) xs[j] = (float)xg[(size_t)b * D + j];
  __syncthreads();
  float d = 0.f, fro = 0.f;
  for (int r = lane; r < Mr; r += 64) {
    float s = 0.f;
    for (int j = 0; j < D; ++j) {
      const float h = Hb[(size_t)r * D + j];
      s = __builtin_fmaf(h, xs[j], s);
      fro = __builtin_fmaf(h, h, fro);
    }
    const float e = y[(size_t)b * Mr + r] - s;
    yb[r] = e;
    d = __builtin_fmaf(e, e, d);
  }
  __syncthreads();
  float g2 = 0.f;
  for (int j = lane; j < D; j += 64) {
    float s = 0.f;
    for (int r = 0; r < Mr; ++r) s = __builtin_fmaf(Hb[(size_t)r * D + j], yb[r], s);
    g2 = __builtin_fmaf(s, s, g2);
  }
  for (int o = 32; o; o >>= 1) {
    d += __shfl_xor(d, o);
    fro += __shfl_xor(fro, o);
    g2 += __shfl_xor(g2, o);
  }
  if (lane == 0) {
    const float omega = M == 4 ? 1.f : 3.f;
    const float G = 2.f * sqrtf(g2) / (fro / (float)D);
    float R = (2.f * omega <= G) ? omega : floorf(G / 2.f);
    R = fminf(fmaxf(R, 0.f), (float)Rmax);
    if (Rout) Rout[b] = (int32_t)R;
    if (Gam) Gam[b] = G;
    if (dist) dist[b] = d;
  }
}

// Fixed-point dDI sub-problem around x_g for the frames of a list (slot i: frame list[i]), written at the frame's
// index: Gi (D x D) = rne(G / scale Q), hi = rne(2 H^T (y - H x_g) / scale Q), scale = max(max|h|, max|G|) (for 2K >= 2
// copies the expanded couplings hold every entry of G). Thread i owns row i; H is staged in LDS.
__global__ __launch_bounds__(128) void mdi_build_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                        const int8_t* __restrict__ xg,
                                                        const int32_t* __restrict__ list, int n, int Mr, int D,
                                                        float Q, int32_t* __restrict__ Gi, int32_t* __restrict__ hi) {
  extern __shared__ float sm[];
  const int slot = blockIdx.x, i = threadIdx.x, nt = blockDim.x;
  if (slot >= n) return;
  const int b = list[slot];
  float* Hs = sm;               // Mr x D
  float* yb = Hs + Mr * D;      // Mr
  float* xs = yb + Mr;          // D
  float* red = xs + D;          // 2
  const float* Hb = H + (size_t)b * Mr * D;
  for (int x = i; x < Mr * D; x += nt) Hs[x] = Hb[x];
  for (int j = i; j < D; j += nt) xs[j] = (float)xg[(size_t)b * D + j];
  __syncthreads();
  for (int r = i; r < Mr; r += nt) {
    float s = 0.f;
    for (int j = 0; j < D; ++j) s = __builtin_fmaf(Hs[r * D + j], xs[j], s);
    yb[r] = y[(size_t)b * Mr + r] - s;
  }
  __syncthreads();
  const bool own = i < D;
  float hv = 0.f, gm = 0.f;
  if (own) {
    float s = 0.f;
    for (int r = 0; r < Mr; ++r) s = __builtin_fmaf(Hs[r * D + i], yb[r], s);
    hv = 2.f * s;
    for (int j = 0; j < D; ++j) {
      float g = 0.f;
      for (int r = 0; r < Mr; ++r) g = __builtin_fmaf(Hs[r * D + i], Hs[r * D + j], g);
      gm = fmaxf(gm, fabsf(g));
    }
  }
  float m = fmaxf(gm, fabsf(hv));
  for (int o = 32; o; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if ((i & 63) == 0) red[i >> 6] = m;
  __syncthreads();
  for (int w = 0; w < (nt >> 6); ++w) m = fmaxf(m, red[w]);
  const float scale = fmaxf(m, 1e-12f);
  if (own) {
    hi[(size_t)b * D + i] = (int32_t)rintf(hv / scale * Q);
    for (int j = 0; j < D; ++j) {
      float g = 0.f;
      for (int r = 0; r < Mr; ++r) g = __builtin_fmaf(Hs[r * D + i], Hs[r * D + j], g);
      Gi[(size_t)b * D * D + (size_t)i * D + j] = (int32_t)rintf(g / scale * Q);
    }
  }
}

static int threads_for(int D) { return D <= 64 ? 64 : 128; }

hipError_t launch_mmse_sic(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M,
                           int per_user, int8_t* xs, hipStream_t st) {
  const size_t lds = sizeof(float) * ((size_t)Mr * D + (size_t)D * (D + 1) + 2 * D + 4) + sizeof(int) * D;
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)mmse_sic_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(mmse_sic_kernel, dim3(B), dim3(threads_for(D)), lds, st, H, y, N0v, N0s, Mr, D, M, per_user, xs);
  return hipGetLastError();
}

hipError_t launch_mdi_radius(const float* H, const float* y, const int8_t* xg, int B, int Mr, int D, int M, int Rmax,
                             int32_t* R, float* Gam, float* dist, hipStream_t st) {
  hipLaunchKernelGGL(mdi_radius_kernel, dim3(B), dim3(64), 0, st, H, y, xg, Mr, D, M, Rmax, R, Gam, dist);
  return hipGetLastError();
}

hipError_t launch_mdi_build(const float* H, const float* y, const int8_t* xg, const int32_t* list, int n, int Mr, int D,
                            float Q, int32_t* Gi, int32_t* hi, hipStream_t st) {
  // up to 32 users: the matrix-core front end in its given-decision form (Gram on MFMA, no inversion); 64 users with
  // the complex-lift form selected: the Hermitian Gram from half of H, likewise
  if (mdi_build_mfma(H, y, xg, list, n, Mr, D, Q, Gi, hi, st)) return hipGetLastError();
  if (D == 128 && Mr == 128 && front_end_form() == 2 && mdi_build_cpx128(H, y, xg, list, n, Q, Gi, hi, st))
    return hipGetLastError();
  const size_t lds = sizeof(float) * ((size_t)Mr * D + Mr + D + 2);
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)mdi_build_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(mdi_build_kernel, dim3(n), dim3(threads_for(D)), lds, st, H, y, xg, list, n, Mr, D, Q, Gi, hi);
  return hipGetLastError();
}

}  // namespace arp
