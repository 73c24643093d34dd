// Decode and select on MI355X: anneal spins -> candidate levels -> lowest ||y - H x||^2 (candidates, then the MMSE
// estimate) -> Gray bit errors.
//
// Reference semantics: mimo.spins_to_levels_delta / spins_to_levels (arpeggio/mimo/mimo.py:185-190, 39-42),
// mimo.dist / select (mimo.py:416-425, first minimum, MMSE estimate last), mimo.levels_to_bits / bit_errors
// (mimo.py:45-51, 428-431), arpx.biterr_per_use (arpeggio/sim/arpx.py:63-66).
// One workgroup per frame. H is staged in LDS with row stride D+1, so the lanes of a wave (one receive row each)
// read distinct banks while the candidate's level is a broadcast.
#include "kernels.h"
#include "adaptive.h"

#include <algorithm>
#include <cstdint>

namespace arp {

typedef float floatx4 __attribute__((ext_vector_type(4)));

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
                                                            float* __restrict__ dist_out, DecodeExtra ex) {
  extern __shared__ float sm[];
  const int slot = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  int b = slot;
  if (ex.list) {
    if (slot >= *ex.count) return;
    b = ex.list[slot];
  }
  // read before any thread of the workgroup writes best_d, which may alias prior_d
  const float pdist = ex.prior_d ? ex.prior_d[b] : 0.f;
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
      const int8_t* s = spins + ((size_t)slot * R + r) * N;
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
    for (int c = tid; c < C; c += nt) dist_out[(size_t)slot * C + c] = dv[c];
  const bool keep = ex.prior_d && !(bd < pdist);   // the prior candidate wins ties
  if (ex.best_d && tid == 0) ex.best_d[b] = keep ? pdist : bd;
  int errs = 0;
  for (int j = tid; j < D; j += nt) {
    const int l = keep ? (int)ex.prior_l[(size_t)b * D + j] : (int)Lc[best * D + j];
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

// One-wave form for D <= 64 and Mr <= 64 (16 and 32 users): lane i keeps receive row i of H in registers (16-byte
// loads when rows are aligned), the candidates' levels are formed into LDS (from 32-bit spin words when D % 4 == 0),
// and each candidate's residual row sums read the levels back as broadcasts. The per-row sum runs over j in order
// with fused multiply-adds; the row terms meet in a butterfly.
template <int DP>
__global__ __launch_bounds__(64) void decode_w64_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                        const int8_t* __restrict__ xg, const int8_t* __restrict__ spins,
                                                        int R, int P, int D, int Mr, int M, int delta, int vec4,
                                                        int vecs,
                                                        const int8_t* __restrict__ ltrue, int8_t* __restrict__ lhat,
                                                        int32_t* __restrict__ biterr, float* __restrict__ dist_out,
                                                        DecodeExtra ex) {
  extern __shared__ float sm[];
  const int slot = blockIdx.x, lane = threadIdx.x;
  int b = slot;
  if (ex.list) {
    if (slot >= *ex.count) return;
    b = ex.list[slot];
  }
  // read before any thread of the workgroup writes best_d, which may alias prior_d
  const float pdist = ex.prior_d ? ex.prior_d[b] : 0.f;
  const int C = R + 1;
  float* Lc = sm;  // C x DP candidate levels (columns beyond D are never read)
  const float* Hb = H + (size_t)b * Mr * D;
  const bool row = lane < Mr;
  float hr[DP];
  if (vec4) {
#pragma unroll
    for (int q = 0; q < DP / 4; ++q) {
      floatx4 t = {0.f, 0.f, 0.f, 0.f};
      if (row && 4 * q < D) t = *reinterpret_cast<const floatx4*>(Hb + (size_t)lane * D + 4 * q);
#pragma unroll
      for (int e = 0; e < 4; ++e) hr[4 * q + e] = t[e];
    }
  } else {
#pragma unroll
    for (int j = 0; j < DP; ++j) hr[j] = (row && j < D) ? Hb[(size_t)lane * D + j] : 0.f;
  }
  const float yi = row ? y[(size_t)b * Mr + lane] : 0.f;
  const float lmax = M == 4 ? 1.f : 3.f;
  const int N = delta ? P * D : (M == 4 ? D : 2 * D);
  const bool col = lane < D;
  if (vecs) {
    // spins as 32-bit words: lane (g, q) forms coordinates 4q..4q+3 of candidate r0 + t G + g from word q of each copy
    // (delta) or of each bit plane (RI); every word a lane needs is loaded before the first is used
    const int QW = D / 4, G = 64 / QW;       // words per copy row, candidates per pass
    const int q = lane % QW, g = lane / QW;
    const int planes = delta ? P : (M == 4 ? 1 : 2);
    const bool live = g < G;
    const int xw = live ? *reinterpret_cast<const int*>(xg + (size_t)b * D + 4 * q) : 0;
    for (int r0 = 0; r0 < C; r0 += 4 * G) {
      int w[4][8];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int r = r0 + t * G + g;
#pragma unroll
        for (int p = 0; p < 8; ++p)
          w[t][p] = (live && r < R && p < planes)
                        ? *reinterpret_cast<const int*>(spins + ((size_t)slot * R + r) * N + (size_t)p * D + 4 * q)
                        : 0;
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int r = r0 + t * G + g;
        if (!live || r >= C) continue;
        floatx4 lv;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int x = (int)(int8_t)(xw >> (8 * e));
          float l;
          if (r == R) {
            l = (float)x;
          } else if (delta) {
            int d = 0;
#pragma unroll
            for (int p = 0; p < 8; ++p)
              if (p < planes) d += (int)(int8_t)(w[t][p] >> (8 * e));
            l = fminf(fmaxf((float)(x + d), -lmax), lmax);
          } else {
            const int s0 = (int)(int8_t)(w[t][0] >> (8 * e));
            l = M == 4 ? (float)s0 : (float)(2 * s0 + (int)(int8_t)(w[t][1] >> (8 * e)));
          }
          lv[e] = l;
        }
        *reinterpret_cast<floatx4*>(Lc + r * DP + 4 * q) = lv;
      }
    }
  } else {
    const int x0 = col ? (int)xg[(size_t)b * D + lane] : 0;
    for (int r = 0; r < C; ++r) {
      float l = 0.f;
      if (col) {
        if (r == R) {
          l = (float)x0;
        } else {
          const int8_t* s = spins + ((size_t)slot * R + r) * N;
          if (delta) {
            int d = 0;
            for (int p = 0; p < P; ++p) d += s[p * D + lane];
            l = fminf(fmaxf((float)(x0 + d), -lmax), lmax);
          } else {
            l = M == 4 ? (float)s[lane] : (float)(2 * s[lane] + s[D + lane]);
          }
        }
      }
      if (lane < DP) Lc[r * DP + lane] = l;
    }
  }
  __syncthreads();
  int best = 0;
  float bd = 0.f;
  for (int r = 0; r < C; ++r) {
    const float* lc = Lc + r * DP;
    float sacc = 0.f;
#pragma unroll
    for (int j = 0; j < DP; j += 4) {
      const floatx4 lv = *reinterpret_cast<const floatx4*>(lc + j);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (j + e < D) sacc = __builtin_fmaf(hr[j + e], lv[e], sacc);
    }
    const float e = yi - sacc;
    float part = row ? e * e : 0.f;
    for (int o = 32; o; o >>= 1) part += __shfl_xor(part, o);
    if (dist_out && lane == 0) dist_out[(size_t)slot * C + r] = part;
    if (r == 0 || part < bd) {
      bd = part;
      best = r;
    }
  }
  const bool keep = ex.prior_d && !(bd < pdist);   // the prior candidate wins ties
  if (ex.best_d && lane == 0) ex.best_d[b] = keep ? pdist : bd;
  int errs = 0;
  if (col) {
    const int l = keep ? (int)ex.prior_l[(size_t)b * D + lane] : (int)Lc[best * DP + lane];
    if (lhat) lhat[(size_t)b * D + lane] = (int8_t)l;
    if (ltrue) errs = __builtin_popcount(gray_bits(l, M) ^ gray_bits(ltrue[(size_t)b * D + lane], M));
  }
  if (biterr && ltrue) {
    for (int o = 32; o; o >>= 1) errs += __shfl_xor(errs, o);
    if (lane == 0) biterr[b] = errs;
  }
}

// Sum over the wave, the same value in every lane: within each row of 16 lanes by DPP (quad xor 1, quad xor 2, the
// half-row mirror, the row mirror: each step adds two equal partial sums, so all 16 lanes hold the same row sum), then
// the four row sums through scalar registers in a fixed order.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_mov<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_mov<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_mov<0x141>(v);  // row_half_mirror
  v += dpp_mov<0x140>(v);  // row_mirror
  const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0));
  const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16));
  const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
  const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48));
  return (r0 + r1) + (r2 + r3);
}

// The dDI decode of the detector's common shape: two copies, square (Mr = D), D = 32 or 64 as a compile-time
// constant, H rows 16-byte aligned. H passes through LDS: coalesced 16-byte loads (each instruction reads 1 KiB of
// consecutive rows), then lane r reads receive row r into registers (row stride D + 4 words: no bank conflicts). Lane
// j forms coordinate j of every candidate into LDS (over the staged H, once it is in registers); each candidate's
// residual rows are dot products over the row (no bound tests) and meet in wave_sum. Same selection rule and extras
// as decode_w64_kernel.
// LIFT: H is the real lift of a complex channel, [[Re, -Im], [Im, Re]] (n = D / 2): only its first n columns
// C1 = [Re; Im] are read, and lane r rebuilds its row as [C1[r], -C1[r + n]] (r < n) or [C1[r], C1[r - n]]: the same
// values as the stored row, so the distances are the full-H ones bit for bit, from half the bytes.
template <int D, bool LIFT>
__global__ __launch_bounds__(64) void decode_dd_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                       const int8_t* __restrict__ xg,
                                                       const int8_t* __restrict__ spins, int R, int M,
                                                       const int8_t* __restrict__ ltrue, int8_t* __restrict__ lhat,
                                                       int32_t* __restrict__ biterr, float* __restrict__ dist_out,
                                                       DecodeExtra ex) {
  extern __shared__ float sm[];
  const int slot = blockIdx.x, lane = threadIdx.x;
  int b = slot;
  if (ex.list) {
    if (slot >= *ex.count) return;
    b = ex.list[slot];
  }
  // read before any lane writes best_d, which may alias prior_d
  const float pdist = ex.prior_d ? ex.prior_d[b] : 0.f;
  const int C = R + 1;
  constexpr int W = LIFT ? D / 2 : D;   // staged columns per row
  constexpr int HS = W + 4;             // LDS row stride of the staged H, in words
  float* Lc = sm;                       // C x D candidate levels (after H has left the LDS)
  const float* Hb = H + (size_t)b * D * D;
  const bool row = lane < D;
  constexpr int NV = D * W / 4 / 64;    // 16-byte pieces per lane
  floatx4 stage[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const int f = 64 * k + lane, r = f / (W / 4), c4 = f % (W / 4);
    stage[k] = *reinterpret_cast<const floatx4*>(Hb + (size_t)r * D + 4 * c4);
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const int f = 64 * k + lane, r = f / (W / 4), c4 = f % (W / 4);
    *reinterpret_cast<floatx4*>(sm + r * HS + 4 * c4) = stage[k];
  }
  __syncthreads();
  float hr[D];
#pragma unroll
  for (int q = 0; q < W / 4; ++q) {
    floatx4 t = {0.f, 0.f, 0.f, 0.f};
    if (row) t = *reinterpret_cast<const floatx4*>(sm + lane * HS + 4 * q);
#pragma unroll
    for (int e = 0; e < 4; ++e) hr[4 * q + e] = t[e];
  }
  if constexpr (LIFT) {
    const int r2 = lane < W ? lane + W : lane - W;
    const float sg = lane < W ? -1.f : 1.f;
#pragma unroll
    for (int q = 0; q < W / 4; ++q) {
      floatx4 t = {0.f, 0.f, 0.f, 0.f};
      if (row) t = *reinterpret_cast<const floatx4*>(sm + r2 * HS + 4 * q);
#pragma unroll
      for (int e = 0; e < 4; ++e) hr[W + 4 * q + e] = sg * t[e];
    }
  }
  const float yi = row ? y[(size_t)b * D + lane] : 0.f;
  const float lmax = M == 4 ? 1.f : 3.f;
  __syncthreads();   // every lane has its row: the LDS now takes the candidates
  if (row) {
    const int x0 = (int)xg[(size_t)b * D + lane];
    const int8_t* s = spins + (size_t)slot * R * 2 * D + lane;
    for (int c = 0; c < R; ++c)
      Lc[c * D + lane] = fminf(fmaxf((float)(x0 + (int)s[(size_t)c * 2 * D] + (int)s[(size_t)c * 2 * D + D]), -lmax),
                               lmax);
    Lc[R * D + lane] = (float)x0;
  }
  __syncthreads();
  int best = 0;
  float bd = 0.f;
  for (int c = 0; c < C; ++c) {
    const float* lc = Lc + c * D;
    float sacc = 0.f;
#pragma unroll
    for (int j = 0; j < D; j += 4) {
      const floatx4 lv = *reinterpret_cast<const floatx4*>(lc + j);
#pragma unroll
      for (int e = 0; e < 4; ++e) sacc = __builtin_fmaf(hr[j + e], lv[e], sacc);
    }
    const float e = yi - sacc;
    const float part = wave_sum(row ? e * e : 0.f);
    if (dist_out && lane == 0) dist_out[(size_t)slot * C + c] = part;
    if (c == 0 || part < bd) {
      bd = part;
      best = c;
    }
  }
  const bool keep = ex.prior_d && !(bd < pdist);   // the prior candidate wins ties
  if (ex.best_d && lane == 0) ex.best_d[b] = keep ? pdist : bd;
  int errs = 0;
  if (row) {
    const int l = keep ? (int)ex.prior_l[(size_t)b * D + lane] : (int)Lc[best * D + lane];
    if (lhat) lhat[(size_t)b * D + lane] = (int8_t)l;
    if (ltrue) errs = __builtin_popcount(gray_bits(l, M) ^ gray_bits(ltrue[(size_t)b * D + lane], M));
  }
  if (biterr && ltrue) {
    for (int o = 32; o; o >>= 1) errs += __shfl_xor(errs, o);
    if (lane == 0) biterr[b] = errs;
  }
}

// Two-wave form for D, Mr <= 128 (up to 64 users): thread i keeps receive row i of H in registers, thread j forms
// coordinate j of every candidate into LDS, and each candidate's row terms meet in a butterfly per wave and then
// across the two waves through LDS.
__global__ __launch_bounds__(128) void decode_w128_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                          const int8_t* __restrict__ xg,
                                                          const int8_t* __restrict__ spins, int R, int P, int D,
                                                          int Mr, int M, int delta, int vec4,
                                                          const int8_t* __restrict__ ltrue, int8_t* __restrict__ lhat,
                                                          int32_t* __restrict__ biterr, float* __restrict__ dist_out,
                                                          DecodeExtra ex) {
  constexpr int DP = 128;
  extern __shared__ float sm[];
  const int slot = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  int b = slot;
  if (ex.list) {
    if (slot >= *ex.count) return;
    b = ex.list[slot];
  }
  // read before any thread of the workgroup writes best_d, which may alias prior_d
  const float pdist = ex.prior_d ? ex.prior_d[b] : 0.f;
  const int C = R + 1;
  float* Lc = sm;              // C x DP candidate levels
  float* dsum = Lc + C * DP;   // C x 2 per-wave partial distances
  const float* Hb = H + (size_t)b * Mr * D;
  const bool row = t < Mr;
  float hr[DP];
  if (vec4 == 2) {
    // the real lift of a complex channel (D = Mr = 128): only the first 64 columns are read, and row t's last 64
    // entries are row t + 64's first (negated) for t < 64, row t - 64's first for t >= 64: the stored values
    const int t2 = t < 64 ? t + 64 : t - 64;
    const float sg = t < 64 ? -1.f : 1.f;
#pragma unroll
    for (int q = 0; q < DP / 4; ++q) {
      const floatx4 v = *reinterpret_cast<const floatx4*>(Hb + (size_t)(q < 16 ? t : t2) * D + 4 * (q & 15));
#pragma unroll
      for (int e = 0; e < 4; ++e) hr[4 * q + e] = q < 16 ? v[e] : sg * v[e];
    }
  } else if (vec4) {
#pragma unroll
    for (int q = 0; q < DP / 4; ++q) {
      floatx4 v = {0.f, 0.f, 0.f, 0.f};
      if (row && 4 * q < D) v = *reinterpret_cast<const floatx4*>(Hb + (size_t)t * D + 4 * q);
#pragma unroll
      for (int e = 0; e < 4; ++e) hr[4 * q + e] = v[e];
    }
  } else {
#pragma unroll
    for (int jj = 0; jj < DP; ++jj) hr[jj] = (row && jj < D) ? Hb[(size_t)t * D + jj] : 0.f;
  }
  const float yi = row ? y[(size_t)b * Mr + t] : 0.f;
  const float lmax = M == 4 ? 1.f : 3.f;
  const int N = delta ? P * D : (M == 4 ? D : 2 * D);
  const bool col = t < D;
  const int x0 = col ? (int)xg[(size_t)b * D + t] : 0;
  for (int r = 0; r < C; ++r) {
    float l = 0.f;
    if (col) {
      if (r == R) {
        l = (float)x0;
      } else {
        const int8_t* s = spins + ((size_t)slot * R + r) * N;
        if (delta) {
          int d = 0;
          for (int p = 0; p < P; ++p) d += s[p * D + t];
          l = fminf(fmaxf((float)(x0 + d), -lmax), lmax);
        } else {
          l = M == 4 ? (float)s[t] : (float)(2 * s[t] + s[D + t]);
        }
      }
    }
    Lc[r * DP + t] = l;
  }
  __syncthreads();
  // all DP terms: past D both hr (loaded as zero) and the levels (written as zero) are zero, and fma(0, 0, s) = s, so
  // the sum is the D-term sum bit for bit, without a per-term bound test
  for (int r = 0; r < C; ++r) {
    const float* lc = Lc + r * DP;
    float sacc = 0.f;
#pragma unroll
    for (int jj = 0; jj < DP; jj += 4) {
      const floatx4 lv = *reinterpret_cast<const floatx4*>(lc + jj);
#pragma unroll
      for (int e = 0; e < 4; ++e) sacc = __builtin_fmaf(hr[jj + e], lv[e], sacc);
    }
    const float e = yi - sacc;
    float part = row ? e * e : 0.f;
    for (int o = 32; o; o >>= 1) part += __shfl_xor(part, o);
    if (lane == 0) dsum[2 * r + wv] = part;
  }
  __syncthreads();
  int best = 0;
  float bd = 0.f;
  for (int r = 0; r < C; ++r) {
    const float dr = dsum[2 * r] + dsum[2 * r + 1];
    if (r == 0 || dr < bd) {
      bd = dr;
      best = r;
    }
  }
  if (dist_out)
    for (int r = t; r < C; r += 128) dist_out[(size_t)slot * C + r] = dsum[2 * r] + dsum[2 * r + 1];
  const bool keep = ex.prior_d && !(bd < pdist);   // the prior candidate wins ties
  if (ex.best_d && t == 0) ex.best_d[b] = keep ? pdist : bd;
  int errs = 0;
  if (col) {
    const int l = keep ? (int)ex.prior_l[(size_t)b * D + t] : (int)Lc[best * DP + t];
    if (lhat) lhat[(size_t)b * D + t] = (int8_t)l;
    if (ltrue) errs = __builtin_popcount(gray_bits(l, M) ^ gray_bits(ltrue[(size_t)b * D + t], M));
  }
  if (biterr && ltrue) {
    for (int o = 32; o; o >>= 1) errs += __shfl_xor(errs, o);
    __syncthreads();
    if (lane == 0) dsum[wv] = (float)errs;
    __syncthreads();
    if (t == 0) biterr[b] = (int)dsum[0] + (int)dsum[1];
  }
}

size_t decode_lds_bytes(int R, int D, int Mr) {
  return sizeof(float) * ((size_t)Mr * (D + 1) + Mr + (size_t)(R + 1) * D + (R + 1 > 4 ? R + 1 : 4));
}

hipError_t launch_decode_select(const float* H, const float* y, const int8_t* xg, const int8_t* spins, int B, int R,
                                int P, int D, int Mr, int M, int delta, const int8_t* ltrue, int8_t* lhat,
                                int32_t* biterr, float* dist_out, hipStream_t st) {
  return launch_decode_ex(H, y, xg, spins, B, R, P, D, Mr, M, delta, ltrue, lhat, biterr, dist_out,
                          DecodeExtra{nullptr, nullptr, nullptr, nullptr, nullptr}, st);
}

hipError_t launch_decode_ex(const float* H, const float* y, const int8_t* xg, const int8_t* spins, int B, int R, int P,
                            int D, int Mr, int M, int delta, const int8_t* ltrue, int8_t* lhat, int32_t* biterr,
                            float* dist_out, const DecodeExtra& ex, hipStream_t st) {
  const bool aligned = ((reinterpret_cast<uintptr_t>(H) & 15) == 0);
  if (delta && P == 2 && Mr == D && (D == 64 || D == 32) && aligned) {
    const bool lift = front_end_form() == 2;   // the complex-lift form: H is the real lift of a complex channel
    const size_t lds = sizeof(float) * std::max((size_t)(R + 1) * D, (size_t)D * ((lift ? D / 2 : D) + 4));
#define ARP_DD(DV, LV)                                                                                                \
  hipLaunchKernelGGL((decode_dd_kernel<DV, LV>), dim3(B), dim3(64), lds, st, H, y, xg, spins, R, M, ltrue, lhat,     \
                     biterr, dist_out, ex)
    if (D == 64 && lift) ARP_DD(64, true);
    else if (D == 64) ARP_DD(64, false);
    else if (lift) ARP_DD(32, true);
    else ARP_DD(32, false);
#undef ARP_DD
    return hipGetLastError();
  }
  if (D <= 64 && Mr <= 64) {
    const int DP = (D + 15) & ~15;
    const size_t l64 = sizeof(float) * (size_t)(R + 1) * DP;
    const int vec4 = (D % 4 == 0) && ((reinterpret_cast<uintptr_t>(H) & 15) == 0);
    const int planes = delta ? P : (M == 4 ? 1 : 2);
    const int vecs = (D % 4 == 0) && planes <= 8 && ((reinterpret_cast<uintptr_t>(spins) & 3) == 0) &&
                     ((reinterpret_cast<uintptr_t>(xg) & 3) == 0);
#define ARP_DECODE_W64(DPv)                                                                                        \
  hipLaunchKernelGGL((decode_w64_kernel<DPv>), dim3(B), dim3(64), l64, st, H, y, xg, spins, R, P, D, Mr, M, delta,   \
                     vec4, vecs, ltrue, lhat, biterr, dist_out, ex)
    if (DP == 16) ARP_DECODE_W64(16);
    else if (DP == 32) ARP_DECODE_W64(32);
    else if (DP == 48) ARP_DECODE_W64(48);
    else ARP_DECODE_W64(64);
#undef ARP_DECODE_W64
    return hipGetLastError();
  }
  if (D <= 128 && Mr <= 128) {
    const size_t l128 = sizeof(float) * ((size_t)(R + 1) * 128 + 2 * (size_t)(R + 1) + 2);
    if (l128 > 65536)
      hipFuncSetAttribute((const void*)decode_w128_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l128);
    const int vec4 = (D % 4 == 0) && ((reinterpret_cast<uintptr_t>(H) & 15) == 0);
    // vec4 = 2: the complex-lift form reads half of H
    const int lift = vec4 && D == 128 && Mr == 128 && front_end_form() == 2 ? 2 : vec4;
    hipLaunchKernelGGL(decode_w128_kernel, dim3(B), dim3(128), l128, st, H, y, xg, spins, R, P, D, Mr, M, delta, lift,
                       ltrue, lhat, biterr, dist_out, ex);
    return hipGetLastError();
  }
  const size_t lds = decode_lds_bytes(R, D, Mr);
  if (lds > 65536)
    hipFuncSetAttribute((const void*)decode_select_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(decode_select_kernel, dim3(B), dim3(256), lds, st, H, y, xg, spins, R, P, D, Mr, M, delta, ltrue,
                     lhat, biterr, dist_out, ex);
  return hipGetLastError();
}

}  // namespace arp
