// Fused per-frame front end for MI355X: LMMSE hard decision + delta-Ising (dDI) build + fixed-point quantisation.
//
// Reference semantics (float, then rounded): mimo.mmse / _mu / mmse_hard / quantize_levels (arpeggio/mimo/mimo.py:86-109,
// 176-182), mimo.ising_delta_at (mimo.py:506-523), chordfx.to_fixed (arpeggio/mimo/chordfx.py:60-76) and
// chord_ref.compress (arpeggio/sim/x3model/chord_ref.py:47-59). Per frame:
//   1. G = H^T H on the matrix cores (v_mfma_f32_16x16x4f32 over the upper 16x16 tiles, H staged in LDS, G exactly
//      symmetric); c = H^T y.
//   2. A = G + sigma I, sigma = (N0/2)/E[l^2], inverted in registers by the symmetric sweep (below);
//      z = A^-1 c; mu_i = diag(A^-1 G)_i = 1 - sigma (A^-1)_ii.
//   3. x_g = nearest level of z / mu (ties to the lower level, as argmin over the level list).
//   4. ybar = y - H x_g, h = 2 H^T ybar, scale = max(max|h|, max|G|) (the dDI coupling matrix holds every G_ij,
//      its diagonal included, on the cross-copy blocks), clamped at 1e-12.
//   5. G_int = rne(G / scale * Q) (G symmetric by construction), h_int = rne(h / scale * Q); optionally the float
//      scaled G and h for the FP32 annealers, and the expanded Chord inputs J (N x N int16, N = 2K D) and h (N).
// Sizes: D <= 128 (2 Nt), Mr <= 128 (2 Nr), zero-padded to multiples of 16 and 4. Systems with D, Mr <= 64 take the
// one-wave form, larger ones the two-wave form.
#include "kernels.h"

namespace arp {

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// ---------------------------------------------------------------------------------------------------------------
// One-wave form for D <= 64 and Mr <= 64 (16 and 32 users): lane j owns column j of A = G + sigma I in registers and
// inverts it in place with the symmetric sweep operator (A is symmetric positive definite, no pivoting): sweeping
// pivot k maps a_kk -> -1/a_kk, a_ik -> a_ik/a_kk, a_kj -> a_kj/a_kk and a_ij -> a_ij - a_ik a_kj / a_kk, so every
// intermediate matrix stays symmetric and, after all pivots, holds -A^-1. Symmetry means the pivot column k that
// every lane needs is the k-th entry of each lane's own column: one all-lane LDS store publishes it and the update
// reads it back as broadcasts, D^2 FMAs per lane in all. Lane i's column of -A^-1 is row i, so z_i and (A^-1)_ii stay
// in lane i. h is formed as 2 (c - G x_g) from the LDS-resident Gram. LDS holds H (Mp x (DP+1)) until the Gram tiles
// and c are formed, then G (DP x (DP+1)) in the same place: 17 KiB per frame at D = 64, rows padded against bank
// conflicts. vec4: H rows are 16-byte aligned (D % 4 == 0 and an aligned base), so H is fetched with 16-byte loads,
// all of a lane's loads in flight before the first LDS store.
template <int DP>
__global__ __launch_bounds__(64) void front_end_w64_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                           const float* __restrict__ N0v, float N0s, int Mr, int D,
                                                           int M, int K, float Q, int vec4, FrontOut o) {
  extern __shared__ float sm[];
  const int b = blockIdx.x, lane = threadIdx.x;
  const int Mp = (Mr + 3) & ~3, ld = DP + 1;
  const int rows = Mp > DP ? Mp : DP;
  float* Hs = sm;             // Mp x ld, later G (DP x ld) in the same place
  float* Gs = sm;
  float* ys = sm + rows * ld; // Mp
  float* colb = ys + Mp;      // 64 words: broadcast buffer
  const float* Hb = H + (size_t)b * Mr * D;
  if (vec4) {
    constexpr int CPR = DP / 4;                   // 16-byte chunks per row
    constexpr int RPP = 64 / CPR;                 // rows per pass of the wave
    constexpr int NPASS = (64 + RPP - 1) / RPP;   // passes covering 64 rows
    const int cr = lane % CPR, rr = lane / CPR;
    const bool live = rr < RPP;
    floatx4 v[NPASS];
#pragma unroll
    for (int t = 0; t < NPASS; ++t) {
      const int r = t * RPP + rr;
      v[t] = floatx4{0.f, 0.f, 0.f, 0.f};
      if (live && r < Mr && 4 * cr < D) v[t] = *reinterpret_cast<const floatx4*>(Hb + (size_t)r * D + 4 * cr);
    }
#pragma unroll
    for (int t = 0; t < NPASS; ++t) {
      const int r = t * RPP + rr;
      if (live && r < Mp) {
#pragma unroll
        for (int e = 0; e < 4; ++e) Hs[r * ld + 4 * cr + e] = v[t][e];
      }
    }
  } else {
    for (int x = lane; x < Mp * DP; x += 64) {
      const int r = x / DP, cc = x - r * DP;
      Hs[r * ld + cc] = (r < Mr && cc < D) ? Hb[(size_t)r * D + cc] : 0.f;
    }
  }
  if (lane < Mp) ys[lane] = lane < Mr ? y[(size_t)b * Mr + lane] : 0.f;
  __syncthreads();
  // Gram on the matrix cores: per 4-row slice of H each lane reads one operand per 16-column block, shared by the
  // nU upper tiles; the tiles stay in the accumulators until H is no longer needed
  constexpr int nT = DP / 16, nU = nT * (nT + 1) / 2;
  floatx4 acc[nU];
#pragma unroll
  for (int u = 0; u < nU; ++u) acc[u] = floatx4{0.f, 0.f, 0.f, 0.f};
  {
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll 2
    for (int kk = 0; kk < Mp; kk += 4) {
      float op[nT];
#pragma unroll
      for (int I = 0; I < nT; ++I) op[I] = Hs[(kk + lk) * ld + 16 * I + li];
      int u = 0;
#pragma unroll
      for (int I = 0; I < nT; ++I)
#pragma unroll
        for (int J = I; J < nT; ++J, ++u) acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(op[I], op[J], acc[u], 0, 0, 0);
    }
  }
  // c_j = (H^T y)_j, four partial sums
  float c;
  {
    float c4[4] = {0.f, 0.f, 0.f, 0.f};
    const int kq = Mp & ~15;
    for (int k = 0; k < kq; k += 16)
#pragma unroll
      for (int e = 0; e < 16; ++e) c4[e & 3] += Hs[(k + e) * ld + lane] * ys[k + e];
    for (int k = kq; k < Mp; ++k) c4[0] += Hs[k * ld + lane] * ys[k];
    c = (c4[0] + c4[1]) + (c4[2] + c4[3]);
  }
  __syncthreads();
  {
    const int li = lane & 15, lk = lane >> 4;
    int u = 0;
#pragma unroll
    for (int I = 0; I < nT; ++I)
#pragma unroll
      for (int J = I; J < nT; ++J, ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          Gs[(16 * I + 4 * lk + v) * ld + 16 * J + li] = acc[u][v];
          Gs[(16 * J + li) * ld + 16 * I + 4 * lk + v] = acc[u][v];
        }
  }
  __syncthreads();
  const float el2 = M == 4 ? 1.f : 5.f;
  const float n0 = N0v ? N0v[b] : N0s;
  const float sigma = (n0 / 2.f) / el2;
  const bool valid = lane < D;
  // column j of A (identity on the padding)
  float a[DP];
#pragma unroll
  for (int i = 0; i < DP; ++i) {
    const float g = lane < DP ? Gs[i * ld + lane] : 0.f;
    a[i] = valid ? (i == lane ? g + sigma : (i < D ? g : 0.f)) : (i == lane ? 1.f : 0.f);
  }
  // Sweep, in place. Step k: every lane publishes a_kj (its own a[k]); lane j != k forms r = a_kj / a_kk and
  // subtracts a_ik r from a_ij, then keeps r as a_kj; lane k takes r = 1 - 1/a_kk, which turns the same update into
  // a_ik -> a_ik / a_kk, and keeps -1/a_kk. The broadcast buffer alternates between two halves so that a step's
  // store never meets the previous step's reads. All DP/4 broadcast loads of a step are issued before its first FMA
  // (the scheduling fence keeps the compiler from sinking them to their uses, which serialises one LDS round trip per
  // four FMAs); both loops are fully unrolled so that a[k] and a[i] are fixed registers.
#pragma clang loop unroll(full)
  for (int k = 0; k < DP; ++k) {
    float* cb = colb + (k & 1) * 64;
    cb[lane] = a[k];
    __syncthreads();
    floatx4 f[DP / 4];
#pragma unroll
    for (int i = 0; i < DP / 4; ++i) f[i] = *reinterpret_cast<const floatx4*>(cb + 4 * i);
    __builtin_amdgcn_sched_barrier(0);
    const float p = __builtin_amdgcn_rcpf(f[k / 4][k % 4]);
    const bool is_k = lane == k;
    const float rk = is_k ? 1.f - p : a[k] * p;
#pragma clang loop unroll(full)
    for (int i = 0; i < DP; ++i)
      if (i != k) a[i] = __builtin_fmaf(-f[i / 4][i % 4], rk, a[i]);
    a[k] = is_k ? -p : rk;
  }
  __syncthreads();
  // z_i = sum_j (A^-1)_ij c_j, d_i = (A^-1)_ii (the array holds -A^-1); c and, below, x_g are broadcast through LDS
  // the same way
  colb[lane] = c;
  __syncthreads();
  float z = 0.f, d = 0.f;
#pragma unroll
  for (int j = 0; j < DP; j += 4) {
    const floatx4 cj = *reinterpret_cast<const floatx4*>(colb + j);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      z = __builtin_fmaf(-a[j + v], cj[v], z);
      d = j + v == lane ? -a[j + v] : d;
    }
  }
  __syncthreads();
  const float mu = 1.f - sigma * d;
  const float v = z / mu;
  const float lmax = M == 4 ? 1.f : 3.f;
  float l = fminf(fmaxf(2.f * ceilf(v * 0.5f) - 1.f, -lmax), lmax);
  if (v != v) l = -lmax;
  if (!valid) l = 0.f;
  if (valid) {
    if (o.z) o.z[(size_t)b * D + lane] = z;
    if (o.mu) o.mu[(size_t)b * D + lane] = mu;
    if (o.xg) o.xg[(size_t)b * D + lane] = (int8_t)l;
  }
  if (o.feat) {
    float gd = valid ? Gs[lane * ld + lane] : 0.f;
    for (int off = 32; off; off >>= 1) gd += __shfl_xor(gd, off);
    const float gbar = gd / (float)D;
    uint16_t* f = o.feat + (size_t)b * o.feat_ld;
    if (valid) {
      f[lane] = to_bf16(c / gbar);
      f[D + lane] = to_bf16(Gs[lane * ld + lane] / gbar - 1.f);
      f[2 * D + lane] = to_bf16(l / lmax);
      f[3 * D + lane] = to_bf16(v - l);
    }
    for (int x = 4 * D + lane; x < o.feat_ld; x += 64) f[x] = to_bf16(x == 4 * D ? sigma / gbar : 0.f);
  }
  // h_j = 2 (c_j - (G x_g)_j); scale = max(max|h|, max|G|)
  colb[lane] = l;
  __syncthreads();
  float gx = 0.f, gm = 0.f;
  for (int i = 0; i < D; ++i) {
    const float gij = Gs[i * ld + lane];
    gx = __builtin_fmaf(gij, colb[i], gx);
    gm = fmaxf(gm, fabsf(gij));
  }
  const float hval = valid ? 2.f * (c - gx) : 0.f;
  float m = valid ? fmaxf(gm, fabsf(hval)) : 0.f;
  for (int off = 32; off; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  const float scale = fmaxf(m, 1e-12f);
  if (lane == 0 && o.scale) o.scale[b] = scale;
  const int P = 2 * K, N = P * D;
  if (valid) {
    const float hs = hval / scale;
    const int32_t h32 = (int32_t)rintf(hs * Q);
    if (o.hi) o.hi[(size_t)b * D + lane] = h32;
    if (o.hf) o.hf[(size_t)b * D + lane] = hs;
    for (int p = 0; p < P; ++p) {
      if (o.hN) o.hN[(size_t)b * N + p * D + lane] = h32;
      if (o.hNf) o.hNf[(size_t)b * N + p * D + lane] = hs;
    }
    for (int i = 0; i < D; ++i) {
      const float gs = Gs[i * ld + lane] / scale;
      const int32_t gi = (int32_t)rintf(gs * Q);
      if (o.Gi) o.Gi[(size_t)b * D * D + i * D + lane] = gi;
      if (o.Gh) o.Gh[(size_t)b * D * D + i * D + lane] = (int16_t)gi;
      if (o.Gf) o.Gf[(size_t)b * D * D + i * D + lane] = gs;
      if (o.J || o.Jf)
        for (int pr = 0; pr < P; ++pr)
          for (int pc = 0; pc < P; ++pc) {
            const size_t x = (size_t)b * N * N + (size_t)(pr * D + i) * N + pc * D + lane;
            const bool diag = pr == pc && i == lane;
            if (o.J) o.J[x] = diag ? (int16_t)0 : (int16_t)(-gi);
            if (o.Jf) o.Jf[x] = diag ? 0.f : -gs;
          }
    }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Two-wave form for 64 < D <= 128 and Mr <= 128 (33..64 users): thread j (of 128) owns column j of A = G + sigma I
// (identity on the padding) in registers and runs the symmetric sweep of the one-wave form over 128 pivots, in
// blocks of 16 so that the unrolled code stays small: every thread publishes its a_kj (= a_jk), one barrier, and the
// update reads the 128 entries back as broadcasts; the broadcast buffer alternates between two halves so one barrier
// per pivot suffices. The Gram's 36 upper 16x16 tiles
// are split over the two waves' matrix cores and kept in accumulators until H is no longer needed; LDS then holds G
// (128 x 129 floats) in H's place: 66 KiB per frame.
constexpr int W128_DP = 128;

__device__ __forceinline__ float block_reduce128(float v, float* red, bool is_max) {
  for (int o = 32; o; o >>= 1) {
    const float t = __shfl_xor(v, o);
    v = is_max ? fmaxf(v, t) : v + t;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const float r = is_max ? fmaxf(red[0], red[1]) : red[0] + red[1];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(128) void front_end_w128_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                             const float* __restrict__ N0v, float N0s, int Mr, int D,
                                                             int M, int K, float Q, int vec4, FrontOut o) {
  constexpr int DP = W128_DP;
  extern __shared__ float sm[];
  const int b = blockIdx.x, j = threadIdx.x, lane = j & 63, wv = j >> 6;
  const int Mp = (Mr + 3) & ~3, ld = DP + 1;
  float* Hs = sm;                 // Mp x ld (Mp <= 128): H, later G
  float* Gs = sm;
  float* ys = sm + DP * ld;       // 128
  float* cb = ys + DP;            // 2 x 128 broadcast halves
  float* red = cb + 2 * DP;       // 2
  const float* Hb = H + (size_t)b * Mr * D;
  if (vec4) {
    constexpr int CPR = DP / 4, RPP = 128 / CPR, NPASS = DP / RPP;
    const int cr = j % CPR, rr = j / CPR;
    floatx4 v[NPASS];
#pragma unroll
    for (int t = 0; t < NPASS; ++t) {
      const int r = t * RPP + rr;
      v[t] = floatx4{0.f, 0.f, 0.f, 0.f};
      if (r < Mr && 4 * cr < D) v[t] = *reinterpret_cast<const floatx4*>(Hb + (size_t)r * D + 4 * cr);
    }
#pragma unroll
    for (int t = 0; t < NPASS; ++t) {
      const int r = t * RPP + rr;
      if (r < Mp) {
#pragma unroll
        for (int e = 0; e < 4; ++e) Hs[r * ld + 4 * cr + e] = v[t][e];
      }
    }
  } else {
    for (int x = j; x < Mp * DP; x += 128) {
      const int r = x / DP, cc = x - r * DP;
      Hs[r * ld + cc] = (r < Mr && cc < D) ? Hb[(size_t)r * D + cc] : 0.f;
    }
  }
  if (j < Mp) ys[j] = j < Mr ? y[(size_t)b * Mr + j] : 0.f;
  __syncthreads();
  // Gram: wave w takes the upper tiles u = 2 q + w (q = 0..17) of the row-major enumeration of I <= J
  constexpr int nT = DP / 16, nU = nT * (nT + 1) / 2, PER = nU / 2;
  floatx4 acc[PER];
  int tI[PER], tJ[PER];
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    int I = 0, rem = 2 * q + wv;
    for (int it = 0; it < nT; ++it)
      if (rem >= nT - I) {
        rem -= nT - I;
        ++I;
      }
    tI[q] = I;
    tJ[q] = I + rem;
    acc[q] = floatx4{0.f, 0.f, 0.f, 0.f};
  }
  for (int kk = 0; kk < Mp; kk += 4) {
    const float* hr = Hs + (kk + lk) * ld + li;
#pragma unroll
    for (int q = 0; q < PER; ++q)
      acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(hr[16 * tI[q]], hr[16 * tJ[q]], acc[q], 0, 0, 0);
  }
  float c;
  {
    float c4[4] = {0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < Mp; ++k) c4[k & 3] += Hs[k * ld + j] * ys[k];
    c = (c4[0] + c4[1]) + (c4[2] + c4[3]);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      Gs[(16 * tI[q] + 4 * lk + v) * ld + 16 * tJ[q] + li] = acc[q][v];
      Gs[(16 * tJ[q] + li) * ld + 16 * tI[q] + 4 * lk + v] = acc[q][v];
    }
  __syncthreads();
  const float el2 = M == 4 ? 1.f : 5.f;
  const float n0 = N0v ? N0v[b] : N0s;
  const float sigma = (n0 / 2.f) / el2;
  const bool valid = j < D;
  float a[DP];
#pragma unroll
  for (int i = 0; i < DP; ++i) {
    const float g = Gs[i * ld + j];
    a[i] = valid ? (i == j ? g + sigma : (i < D ? g : 0.f)) : (i == j ? 1.f : 0.f);
  }
  // Sweep in 8 blocks of 16 pivots with every register index static: the block's 16 rows are selected into blk by
  // the block index, its pivots update all 128 rows of a (the block's own rows there are stale and are refreshed from
  // blk at the end of the block) and blk; pivot row k of blk is published for the update.
  for (int kb = 0; kb < DP / 16; ++kb) {
    float blk[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      float v = a[t];
#pragma unroll
      for (int q = 1; q < DP / 16; ++q) v = kb == q ? a[16 * q + t] : v;
      blk[t] = v;
    }
    for (int kk = 0; kk < 16; ++kk) {
      const int k = 16 * kb + kk;
      float* cbk = cb + (kk & 1) * DP;
      float bk = blk[0];
#pragma unroll
      for (int t = 1; t < 16; ++t) bk = kk == t ? blk[t] : bk;
      cbk[j] = bk;
      __syncthreads();
      const float p = __builtin_amdgcn_rcpf(cbk[k]);
      const bool is_k = j == k;
      const float rk = is_k ? 1.f - p : bk * p;
      // The 128 broadcasts in groups of 8 loads, software-pipelined: group g + 1 is in flight while group g's FMAs
      // issue (the scheduling fences keep the compiler from sinking each load to its four FMAs).
      const floatx4* c4 = reinterpret_cast<const floatx4*>(cbk);
      floatx4 f0[8], f1[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) f0[u] = c4[u];
#pragma unroll
      for (int g = 0; g < DP / 32; g += 2) {
#pragma unroll
        for (int u = 0; u < 8; ++u) f1[u] = c4[8 * (g + 1) + u];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) a[32 * g + 4 * u + v] = __builtin_fmaf(-f0[u][v], rk, a[32 * g + 4 * u + v]);
        __builtin_amdgcn_sched_barrier(0);
        if (g + 2 < DP / 32) {
#pragma unroll
          for (int u = 0; u < 8; ++u) f0[u] = c4[8 * (g + 2) + u];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            a[32 * (g + 1) + 4 * u + v] = __builtin_fmaf(-f1[u][v], rk, a[32 * (g + 1) + 4 * u + v]);
        __builtin_amdgcn_sched_barrier(0);
      }
      const floatx4* fb = reinterpret_cast<const floatx4*>(cbk + 16 * kb);
#pragma unroll
      for (int t4 = 0; t4 < 4; ++t4) {
        const floatx4 f = fb[t4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int t = 4 * t4 + v;
          const float upd = __builtin_fmaf(-f[v], rk, blk[t]);
          blk[t] = t == kk ? (is_k ? -p : rk) : upd;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < DP / 16; ++q)
#pragma unroll
      for (int t = 0; t < 16; ++t) a[16 * q + t] = kb == q ? blk[t] : a[16 * q + t];
  }
  __syncthreads();
  // z_j = sum_i (A^-1)_ji c_i, d_j = (A^-1)_jj (the registers hold -A^-1); c and x_g through LDS
  cb[j] = c;
  __syncthreads();
  float z = 0.f, d = 0.f;
#pragma unroll
  for (int i = 0; i < DP; i += 4) {
    const floatx4 c4 = *reinterpret_cast<const floatx4*>(cb + i);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      z = __builtin_fmaf(-a[i + v], c4[v], z);
      d = i + v == j ? -a[i + v] : d;
    }
  }
  __syncthreads();
  const float mu = 1.f - sigma * d;
  const float vv = z / mu;
  const float lmax = M == 4 ? 1.f : 3.f;
  float l = fminf(fmaxf(2.f * ceilf(vv * 0.5f) - 1.f, -lmax), lmax);
  if (vv != vv) l = -lmax;
  if (!valid) l = 0.f;
  if (valid) {
    if (o.z) o.z[(size_t)b * D + j] = z;
    if (o.mu) o.mu[(size_t)b * D + j] = mu;
    if (o.xg) o.xg[(size_t)b * D + j] = (int8_t)l;
  }
  if (o.feat) {
    const float gbar = block_reduce128(valid ? Gs[j * ld + j] : 0.f, red, false) / (float)D;
    uint16_t* f = o.feat + (size_t)b * o.feat_ld;
    if (valid) {
      f[j] = to_bf16(c / gbar);
      f[D + j] = to_bf16(Gs[j * ld + j] / gbar - 1.f);
      f[2 * D + j] = to_bf16(l / lmax);
      f[3 * D + j] = to_bf16(vv - l);
    }
    for (int x = 4 * D + j; x < o.feat_ld; x += 128) f[x] = to_bf16(x == 4 * D ? sigma / gbar : 0.f);
  }
  // h_j = 2 (c_j - (G x_g)_j); scale = max(max|h|, max|G|)
  cb[j] = l;
  __syncthreads();
  float gx = 0.f, gm = 0.f;
  for (int i = 0; i < D; ++i) {
    const float gij = Gs[i * ld + j];
    gx = __builtin_fmaf(gij, cb[i], gx);
    gm = fmaxf(gm, fabsf(gij));
  }
  const float hval = valid ? 2.f * (c - gx) : 0.f;
  const float scale = fmaxf(block_reduce128(valid ? fmaxf(gm, fabsf(hval)) : 0.f, red, true), 1e-12f);
  if (j == 0 && o.scale) o.scale[b] = scale;
  const int P = 2 * K, N = P * D;
  if (valid) {
    const float hs = hval / scale;
    const int32_t h32 = (int32_t)rintf(hs * Q);
    if (o.hi) o.hi[(size_t)b * D + j] = h32;
    if (o.hf) o.hf[(size_t)b * D + j] = hs;
    for (int pp = 0; pp < P; ++pp) {
      if (o.hN) o.hN[(size_t)b * N + pp * D + j] = h32;
      if (o.hNf) o.hNf[(size_t)b * N + pp * D + j] = hs;
    }
    for (int i = 0; i < D; ++i) {
      const float gs = Gs[i * ld + j] / scale;
      const int32_t gi = (int32_t)rintf(gs * Q);
      if (o.Gi) o.Gi[(size_t)b * D * D + (size_t)i * D + j] = gi;
      if (o.Gh) o.Gh[(size_t)b * D * D + (size_t)i * D + j] = (int16_t)gi;
      if (o.Gf) o.Gf[(size_t)b * D * D + (size_t)i * D + j] = gs;
      if (o.J || o.Jf)
        for (int pr = 0; pr < P; ++pr)
          for (int pc = 0; pc < P; ++pc) {
            const size_t x = (size_t)b * N * N + (size_t)(pr * D + i) * N + pc * D + j;
            const bool diag = pr == pc && i == j;
            if (o.J) o.J[x] = diag ? (int16_t)0 : (int16_t)(-gi);
            if (o.Jf) o.Jf[x] = diag ? 0.f : -gs;
          }
    }
  }
}

size_t front_end_w128_lds_bytes() { return sizeof(float) * ((size_t)W128_DP * (W128_DP + 1) + 3 * W128_DP + 2); }

template <int DP>
static void launch_w64(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M, int K,
                       float Q, const FrontOut& o, hipStream_t st) {
  const int Mp = (Mr + 3) & ~3;
  const size_t lds = sizeof(float) * ((size_t)(Mp > DP ? Mp : DP) * (DP + 1) + Mp + 128);
  const int vec4 = (D % 4 == 0) && ((reinterpret_cast<uintptr_t>(H) & 15) == 0);
  hipLaunchKernelGGL((front_end_w64_kernel<DP>), dim3(B), dim3(64), lds, st, H, y, N0v, N0s, Mr, D, M, K, Q, vec4, o);
}

hipError_t launch_front_end_sweep(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D,
                                  int M, int K, float Q, const FrontOut& o, hipStream_t st) {
  if (D <= 64 && Mr <= 64) {
    const int DP = (D + 15) & ~15;
    if (DP == 16) launch_w64<16>(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
    else if (DP == 32) launch_w64<32>(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
    else if (DP == 48) launch_w64<48>(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
    else launch_w64<64>(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
    return hipGetLastError();
  }
  if (D <= W128_DP && Mr <= W128_DP) {
    const size_t l128 = front_end_w128_lds_bytes();
    hipFuncSetAttribute((const void*)front_end_w128_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l128);
    const int vec4 = (D % 4 == 0) && ((reinterpret_cast<uintptr_t>(H) & 15) == 0);
    hipLaunchKernelGGL(front_end_w128_kernel, dim3(B), dim3(128), l128, st, H, y, N0v, N0s, Mr, D, M, K, Q, vec4, o);
    return hipGetLastError();
  }
  return hipErrorInvalidValue;  // the C ABI admits D, Mr <= 128 only
}

}  // namespace arp
