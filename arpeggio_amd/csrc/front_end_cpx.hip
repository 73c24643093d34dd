// Front end for the real lift of a complex channel on the matrix cores: 16, 32 or 64 users, as many receive antennas.
//
// The real lift H = [[Re, -Im], [Im, Re]] (mimo.complex_to_real, arpeggio/mimo/mimo.py:55-58) has a Gram of the same
// form, G = [[A, -B], [B, A]] with A = Re^T Re + Im^T Im symmetric and B = Re^T Im - Im^T Re antisymmetric: the real
// lift of the Hermitian matrix A + iB. This kernel computes the front end's outputs (identical quantities to
// front_end_mfma.hip, to float32 rounding) from the first n = D / 2 columns of H only, C1 = [Re; Im]:
//   1. A = C1^T C1 and B = C1^T D1, D1 = [Im; -Re] (row r of D1 is +-row r -+ Nr of C1: the same lane's operand of
//      another k-step), in v_mfma_f32_16x16x4f32 accumulators over the upper 16x16 tiles of the n x n tiling;
//      c = H^T y as c_r = C1^T y and c_i = C1^T [y_bot; -y_top]. A and B wait in LDS for the outputs.
//   2. A + sigma I + iB is swept in place, two complex pivots S = {s0, s1} per block: with V = (A + iB)_.S - E and
//      U = V (A + iB)_SS^-1 (a 2 x 2 Hermitian inverse), the block sweep is the rank-2 complex update
//      A + iB -= U V^H, i.e. A -= [U_r | U_i] [V_r | V_i]^T and B -= [U_i | -U_r] [V_r | V_i]^T: one K = 4 MFMA per
//      upper tile of A and of B (6 tiles at n = 32, against 10 for the real 64 x 64 form). The right-hand side
//      c_r + i c_i rides along (c -= V (A + iB)_SS^-1 c_S), so it ends as z = (A + iB)^-1 c.
//   3. mu_i = 1 - sigma Re((A + iB)^-1)_jj (the real lift's (A^-1)_ii for both parts of coordinate j), x_g.
//   4. h = 2 (c - G x_g) from the LDS copies of A and B; scale; fixed point.
//   5. G / scale written row-major: row (part, j) is [A_j, -B_j] for the real part and [B_j, A_j] for the imaginary.
// Tile order: position p = 16 I + m holds complex coordinate j = nC m + I (nC = n / 16 tiles), so a lane's operand
// columns nC col .. nC col + nC - 1 are adjacent in memory.
// Reference semantics as front_end_mfma.hip: mimo.mmse / _mu / quantize_levels (arpeggio/mimo/mimo.py:86-109,
// 176-182), mimo.ising_delta_at (mimo.py:506-523), chordfx.to_fixed (arpeggio/mimo/chordfx.py:60-76).
#include "kernels.h"

#include <cmath>

namespace arp {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float pick4c(floatx4 v, int i) {
  const float a = i & 1 ? v[1] : v[0];
  const float b = i & 1 ? v[3] : v[2];
  return i & 2 ? b : a;
}

__device__ __forceinline__ float lane_val(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

__device__ __forceinline__ uint16_t bf16_bits(float f) {
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

template <int nC>
__device__ constexpr int utile(int I, int J) {
  return I * nC - I * (I - 1) / 2 + (J - I);
}

// rne(x) for |x| < 2^22 from the bits of x + 1.5 2^23
__device__ __forceinline__ uint32_t rne_bits(float x) { return __builtin_bit_cast(uint32_t, x + 12582912.f) - 0x4B400000u; }

}  // namespace

// Four waves per SIMD (128 registers, no accumulator registers; at 16 users the kernel needs fewer anyway).
template <int DP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void front_end_cpx_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                           const float* __restrict__ N0v, float N0s, int M, int K,
                                                           float Q, FrontOut o) {
  constexpr int n = DP / 2, nC = n / 16, nU = nC * (nC + 1) / 2, Mr = DP, Nr = n, S = Mr / 4, SH = Nr / 4;
  constexpr int TS = 260;  // tile stride in words: 16 x 16 entries + 4
  __shared__ __align__(16) float At[nU * TS];  // A's upper tiles, each column-major: entry (r, c) at 16 c + r
  __shared__ __align__(16) float Bt[nU * TS];  // B's upper tiles
  __shared__ __align__(16) float pan[n * 4];   // pan[p * 4 + t]: A[s0][p], A[s1][p], B[s0][p], B[s1][p]
  __shared__ __align__(16) float cz[2 * n];    // c_r, c_i (H^T y) by position
  __shared__ __align__(16) float xs[4 * n];    // x_g by position: [part][p] as x1 = (x_r, x_i), x2 = (-x_i, x_r)
  const int b = blockIdx.x, lane = threadIdx.x, col = lane & 15, grp = lane >> 4;
  const float* Hb = H + (size_t)b * Mr * DP;
  const float* yb = y + (size_t)b * Mr;

  // 1. Gram: at k-step s, lane (col, grp) supplies C1[4 s + grp][nC col + I].
  float opv[S][nC];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const float* hr = Hb + (size_t)(4 * s + grp) * DP + nC * col;
    if constexpr (nC == 2) {
      const float2 x = *reinterpret_cast<const float2*>(hr);
      opv[s][0] = x.x;
      opv[s][1] = x.y;
    } else {
      opv[s][0] = hr[0];
    }
  }
  const float ylane = lane < Mr ? yb[lane] : 0.f;
  floatx4 TA[nU], TB[nU];
  float cpr[nC], cpi[nC];
#pragma unroll
  for (int u = 0; u < nU; ++u) {
    TA[u] = floatx4{0.f, 0.f, 0.f, 0.f};
    TB[u] = floatx4{0.f, 0.f, 0.f, 0.f};
  }
#pragma unroll
  for (int I = 0; I < nC; ++I) cpr[I] = cpi[I] = 0.f;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const float yr = __shfl(ylane, 4 * s + grp);
    // [y_bot; -y_top] at row 4 s + grp; D1's row: C1's row 4 s + grp + Nr (top half) or -(4 s + grp - Nr)
    const float y2 = s < SH ? __shfl(ylane, 4 * s + grp + Nr) : -__shfl(ylane, 4 * s + grp - Nr);
    float d1[nC];
#pragma unroll
    for (int J = 0; J < nC; ++J) d1[J] = s < SH ? opv[s + SH][J] : -opv[s - SH][J];
#pragma unroll
    for (int I = 0; I < nC; ++I) {
      cpr[I] = __builtin_fmaf(opv[s][I], yr, cpr[I]);
      cpi[I] = __builtin_fmaf(opv[s][I], y2, cpi[I]);
    }
#pragma unroll
    for (int I = 0; I < nC; ++I)
#pragma unroll
      for (int J = I; J < nC; ++J) {
        TA[utile<nC>(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(opv[s][I], opv[s][J], TA[utile<nC>(I, J)], 0, 0, 0);
        TB[utile<nC>(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(opv[s][I], d1[J], TB[utile<nC>(I, J)], 0, 0, 0);
      }
  }
  float cr[nC], ci[nC];  // c at position 16 I + col, the same in all four groups
#pragma unroll
  for (int I = 0; I < nC; ++I) {
    float v = cpr[I], w = cpi[I];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    w += __shfl_xor(w, 16);
    w += __shfl_xor(w, 32);
    cr[I] = v;
    ci[I] = w;
    if (grp == 0) {
      cz[16 * I + col] = v;
      cz[n + 16 * I + col] = w;
    }
  }
#pragma unroll
  for (int u = 0; u < nU; ++u) {
    *reinterpret_cast<floatx4*>(At + u * TS + 16 * col + 4 * grp) = TA[u];
    *reinterpret_cast<floatx4*>(Bt + u * TS + 16 * col + 4 * grp) = TB[u];
  }
  // B's diagonal tiles: the two triangles accumulate their products in different orders, so B is made exactly
  // antisymmetric there, (B - B^T) / 2, in the registers and the LDS copy (the output Gram is then exactly symmetric)
  __syncthreads();
#pragma unroll
  for (int I = 0; I < nC; ++I) {
    const float* bt = Bt + utile<nC>(I, I) * TS + col;
#pragma unroll
    for (int v = 0; v < 4; ++v) TB[utile<nC>(I, I)][v] = (TB[utile<nC>(I, I)][v] - bt[16 * (4 * grp + v)]) * 0.5f;
  }
  __syncthreads();
#pragma unroll
  for (int I = 0; I < nC; ++I)
    *reinterpret_cast<floatx4*>(Bt + utile<nC>(I, I) * TS + 16 * col + 4 * grp) = TB[utile<nC>(I, I)];
  const float el2 = M == 4 ? 1.f : 5.f;
  const float n0 = N0v ? N0v[b] : N0s;
  const float sigma = (n0 / 2.f) / el2;
#pragma unroll
  for (int I = 0; I < nC; ++I)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (4 * grp + v == col) TA[utile<nC>(I, I)][v] += sigma;

  // 2. Complex block sweep: pivots s0 = 2 blk, s1 = 2 blk + 1 (positions), tile J0, in-tile row q0 (group g0,
  // registers v0, v0 + 1).
  const int kk = grp & 1;  // this lane's pivot index in the update operands
#pragma unroll
  for (int blk = 0; blk < n / 2; ++blk) {
    const int p0 = 2 * blk, J0 = p0 / 16, q0 = p0 % 16, g0 = q0 / 4, v0 = q0 % 4;
    // panel rows s0, s1 of A and B: tiles (J0, I >= J0) hold them in group g0, registers v0, v0 + 1; for I < J0 the
    // entries are columns q0, q0 + 1 of tile (I, J0) (A symmetric, B antisymmetric)
    if (grp == g0) {
#pragma unroll
      for (int I = J0; I < nC; ++I)
        *reinterpret_cast<floatx4*>(pan + (16 * I + col) * 4) =
            floatx4{TA[utile<nC>(J0, I)][v0], TA[utile<nC>(J0, I)][v0 + 1], TB[utile<nC>(J0, I)][v0],
                    TB[utile<nC>(J0, I)][v0 + 1]};
    }
    if ((col >> 1) == (q0 >> 1)) {
      const int e = col - q0;
#pragma unroll
      for (int I = 0; I < J0; ++I)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          pan[(16 * I + 4 * grp + v) * 4 + e] = TA[utile<nC>(I, J0)][v];
          pan[(16 * I + 4 * grp + v) * 4 + 2 + e] = -TB[utile<nC>(I, J0)][v];
        }
    }
    __syncthreads();
    const floatx4 r0 = *reinterpret_cast<const floatx4*>(pan + p0 * 4);        // A[s0][s0], A[s1][s0], B[s0][s0], B[s1][s0]
    const floatx4 r1 = *reinterpret_cast<const floatx4*>(pan + (p0 + 1) * 4);  // A[s0][s1], A[s1][s1], B[s0][s1], B[s1][s1]
    const float a = r0[0], d = r1[1], al = r1[0], be = r1[2];
    // (A + iB)_SS = [[a, al + i be], [al - i be, d]]; inverse = [[d, -(al + i be)], [-(al - i be), a]] / det
    const float rdet = __builtin_amdgcn_rcpf(__builtin_fmaf(a, d, -__builtin_fmaf(al, al, be * be)));
    const float m00 = d * rdet, m11 = a * rdet, m01r = -al * rdet, m01i = -be * rdet;
    // column kk of the inverse: rows 0 and 1 (complex)
    const float Mr0 = kk ? m01r : m00, Mi0 = kk ? m01i : 0.f, Mr1 = kk ? m11 : m01r, Mi1 = kk ? 0.f : -m01i;
    // q = (A + iB)_SS^-1 c_S
    const float c0r = lane_val(cr[J0], q0), c1r = lane_val(cr[J0], q0 + 1);
    const float c0i = lane_val(ci[J0], q0), c1i = lane_val(ci[J0], q0 + 1);
    const float q0r = __builtin_fmaf(m00, c0r, __builtin_fmaf(m01r, c1r, -m01i * c1i));
    const float q0i = __builtin_fmaf(m00, c0i, __builtin_fmaf(m01r, c1i, m01i * c1r));
    const float q1r = __builtin_fmaf(m01r, c0r, __builtin_fmaf(m01i, c0i, m11 * c1r));
    const float q1i = __builtin_fmaf(m01r, c0i, __builtin_fmaf(-m01i, c0r, m11 * c1i));
    float nuA[nC], nuB[nC], vt[nC];
#pragma unroll
    for (int I = 0; I < nC; ++I) {
      // row p = 16 I + col of V = (A + iB)_.S - E: A[p][s_e] = A[s_e][p], B[p][s_e] = -B[s_e][p]
      const floatx4 x = *reinterpret_cast<const floatx4*>(pan + (16 * I + col) * 4);
      const float vr0 = x[0] - (I == J0 && col == q0 ? 1.f : 0.f), vr1 = x[1] - (I == J0 && col == q0 + 1 ? 1.f : 0.f);
      const float vi0 = -x[2], vi1 = -x[3];
      const float ur = __builtin_fmaf(vr0, Mr0, __builtin_fmaf(-vi0, Mi0, __builtin_fmaf(vr1, Mr1, -vi1 * Mi1)));
      const float ui = __builtin_fmaf(vr0, Mi0, __builtin_fmaf(vi0, Mr0, __builtin_fmaf(vr1, Mi1, vi1 * Mr1)));
      nuA[I] = grp < 2 ? -ur : -ui;
      nuB[I] = grp < 2 ? -ui : ur;
      cr[I] -= __builtin_fmaf(vr0, q0r, __builtin_fmaf(-vi0, q0i, __builtin_fmaf(vr1, q1r, -vi1 * q1i)));
      ci[I] -= __builtin_fmaf(vr0, q0i, __builtin_fmaf(vi0, q0r, __builtin_fmaf(vr1, q1i, vi1 * q1r)));
      // B operand: [V_r | V_i] at row 16 I + col, column grp
      const float w = pan[(16 * I + col) * 4 + grp];
      vt[I] = grp < 2 ? w - (I == J0 && col == q0 + grp ? 1.f : 0.f) : -w;
    }
#pragma unroll
    for (int I = 0; I < nC; ++I)
#pragma unroll
      for (int J = I; J < nC; ++J) {
        TA[utile<nC>(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(nuA[I], vt[J], TA[utile<nC>(I, J)], 0, 0, 0);
        TB[utile<nC>(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(nuB[I], vt[J], TB[utile<nC>(I, J)], 0, 0, 0);
      }
  }

  // 3. Lane i = real coordinate (part = i / n, complex coordinate j = i % n at position p = 16 (j % nC) + j / nC).
  const bool valid = lane < DP;
  const int part = lane / n, j = lane % n, Ii = j % nC, mi = j / nC, pp = 16 * Ii + mi;
  const int src = mi + 16 * (mi >> 2);
  float z = 0.f, dgi = 0.f;
#pragma unroll
  for (int I = 0; I < nC; ++I) {
    const float zr = __shfl(cr[I], mi), zi = __shfl(ci[I], mi);
    const float td = __shfl(pick4c(TA[utile<nC>(I, I)], col & 3), src);
    if (Ii == I) {
      z = part ? zi : zr;
      dgi = 2.f - td;
    }
  }
  const float gii = valid ? At[(Ii * nC - Ii * (Ii - 1) / 2) * TS + 17 * mi] : 0.f;
  const float cc0 = valid ? cz[part * n + pp] : 0.f;
  const float mu = 1.f - sigma * dgi;
  const float vq = z / mu;
  const float lmax = M == 4 ? 1.f : 3.f;
  float l = fminf(fmaxf(2.f * ceilf(vq * 0.5f) - 1.f, -lmax), lmax);
  if (vq != vq) l = -lmax;
  if (!valid) l = 0.f;
  if (valid) {
    if (o.z) o.z[(size_t)b * DP + lane] = z;
    if (o.mu) o.mu[(size_t)b * DP + lane] = mu;
    if (o.xg) o.xg[(size_t)b * DP + lane] = (int8_t)l;
  }
  if (o.feat) {
    float gs = gii;
    for (int off = 32; off; off >>= 1) gs += __shfl_xor(gs, off);
    const float gbar = gs / (float)DP;
    uint16_t* f = o.feat + (size_t)b * o.feat_ld;
    if (valid) {
      f[lane] = bf16_bits(cc0 / gbar);
      f[DP + lane] = bf16_bits(gii / gbar - 1.f);
      f[2 * DP + lane] = bf16_bits(l / lmax);
      f[3 * DP + lane] = bf16_bits(vq - l);
    }
    for (int x = 4 * DP + lane; x < o.feat_ld; x += 64) f[x] = bf16_bits(x == 4 * DP ? sigma / gbar : 0.f);
  }

  // 4. h = 2 (c - G x_g): row (part, j) of G x is A_j x_r - B_j x_i (part 0) or B_j x_r + A_j x_i (part 1), i.e.
  // A_j x1 + B_j x2 with x1 = (x_r, x_i)[part], x2 = (-x_i, x_r)[part]. Row j of A and B: row mi of tiles
  // (Ii, J >= Ii) at 16 c + mi, column mi of tiles (J < Ii, Ii) at 16 mi + c (B's negated).
  if (valid) {
    xs[part * n + pp] = l;                      // x1 of part `part` is x_part
    xs[2 * n + (1 - part) * n + pp] = part ? -l : l;  // x2 of the other part: -x_i for part 0, x_r for part 1
  }
  __syncthreads();
  float gx = 0.f, gm = 0.f;
  if (valid) {
    const float* x1 = xs + part * n;
    const float* x2 = xs + 2 * n + part * n;
#pragma unroll
    for (int J = 0; J < nC; ++J) {
      const bool up = J >= Ii;
      const int tl = up ? Ii * nC - Ii * (Ii - 1) / 2 + (J - Ii) : J * nC - J * (J - 1) / 2 + (Ii - J);
      const float* ga = At + tl * TS + (up ? mi : 16 * mi);
      const float* gb = Bt + tl * TS + (up ? mi : 16 * mi);
      const int st = up ? 16 : 1;
      const float sb = up ? 1.f : -1.f;
#pragma unroll
      for (int c4 = 0; c4 < 16; c4 += 4) {
        const floatx4 u1 = *reinterpret_cast<const floatx4*>(x1 + 16 * J + c4);
        const floatx4 u2 = *reinterpret_cast<const floatx4*>(x2 + 16 * J + c4);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float ga_ = ga[st * (c4 + e)], gb_ = sb * gb[st * (c4 + e)];
          gx = __builtin_fmaf(ga_, u1[e], gx);
          gx = __builtin_fmaf(gb_, u2[e], gx);
          gm = fmaxf(gm, fmaxf(fabsf(ga_), fabsf(gb_)));
        }
      }
    }
  }
  const float hval = valid ? 2.f * (cc0 - gx) : 0.f;
  float mx = fmaxf(gm, fabsf(hval));
  for (int off = 32; off; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  const float scale = fmaxf(mx, 1e-12f);
  const float inv = 1.f / scale;
  if (lane == 0 && o.scale) o.scale[b] = scale;
  const int P = 2 * K, N = P * DP;
  if (valid) {
    const float hs = hval * inv;
    const int32_t h32 = (int32_t)rintf(hs * Q);
    if (o.hi) o.hi[(size_t)b * DP + lane] = h32;
    if (o.hf) o.hf[(size_t)b * DP + lane] = hs;
    for (int p = 0; p < P; ++p) {
      if (o.hN) o.hN[(size_t)b * N + p * DP + lane] = h32;
      if (o.hNf) o.hNf[(size_t)b * N + p * DP + lane] = hs;
    }
  }

  // 5. G / scale row-major. For tile row I and k = 0..3, lane (grp, col) takes row j = nC (4 grp + k) + I of A and B
  // at columns nC col .. nC col + nC - 1 (entries (4 grp + k, col) of tiles (I, J'), or (col, 4 grp + k) of (J', I)
  // for J' < I, B's negated) and writes rows j (A | -B) and n + j (B | A).
  if (o.Gi || o.Gf || o.Gh) {
    int32_t* Gib = o.Gi ? o.Gi + (size_t)b * DP * DP : nullptr;
    int16_t* Ghb = o.Gh ? o.Gh + (size_t)b * DP * DP : nullptr;
    float* Gfb = o.Gf ? o.Gf + (size_t)b * DP * DP : nullptr;
    const float invQ = inv * Q;  // rne(G invQ) = rne((G / scale) Q): Q is a power of two
#pragma unroll
    for (int I = 0; I < nC; ++I) {
      floatx4 ba[nC], bb[nC];
#pragma unroll
      for (int J = 0; J < nC; ++J) {
        if (J >= I) {
          ba[J] = *reinterpret_cast<const floatx4*>(At + utile<nC>(I, J) * TS + 16 * col + 4 * grp);
          bb[J] = *reinterpret_cast<const floatx4*>(Bt + utile<nC>(I, J) * TS + 16 * col + 4 * grp);
        } else {
          const float* a0 = At + utile<nC>(J, I) * TS + 64 * grp + col;
          const float* b0 = Bt + utile<nC>(J, I) * TS + 64 * grp + col;
          ba[J] = floatx4{a0[0], a0[16], a0[32], a0[48]};
          bb[J] = floatx4{-b0[0], -b0[16], -b0[32], -b0[48]};
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int jr = nC * (4 * grp + k) + I;
        const size_t r0 = (size_t)jr * DP + nC * col, r1 = (size_t)(n + jr) * DP + nC * col;
        uint32_t qa[nC], qb[nC], qn[nC];
#pragma unroll
        for (int J = 0; J < nC; ++J) {
          qa[J] = rne_bits(ba[J][k] * invQ);
          qb[J] = rne_bits(bb[J][k] * invQ);
          qn[J] = 0u - qb[J];  // rne(-x) = -rne(x)
        }
        if (Ghb) {
          if constexpr (nC == 2) {
            *reinterpret_cast<uint32_t*>(Ghb + r0) = __builtin_amdgcn_perm(qa[1], qa[0], 0x05040100u);
            *reinterpret_cast<uint32_t*>(Ghb + r0 + n) = __builtin_amdgcn_perm(qn[1], qn[0], 0x05040100u);
            *reinterpret_cast<uint32_t*>(Ghb + r1) = __builtin_amdgcn_perm(qb[1], qb[0], 0x05040100u);
            *reinterpret_cast<uint32_t*>(Ghb + r1 + n) = __builtin_amdgcn_perm(qa[1], qa[0], 0x05040100u);
          } else {
            Ghb[r0] = (int16_t)qa[0];
            Ghb[r0 + n] = (int16_t)qn[0];
            Ghb[r1] = (int16_t)qb[0];
            Ghb[r1 + n] = (int16_t)qa[0];
          }
        }
        if (Gib) {
#pragma unroll
          for (int J = 0; J < nC; ++J) {
            Gib[r0 + J] = (int32_t)qa[J];
            Gib[r0 + n + J] = (int32_t)qn[J];
            Gib[r1 + J] = (int32_t)qb[J];
            Gib[r1 + n + J] = (int32_t)qa[J];
          }
        }
        if (Gfb) {
#pragma unroll
          for (int J = 0; J < nC; ++J) {
            Gfb[r0 + J] = ba[J][k] * inv;
            Gfb[r0 + n + J] = -(bb[J][k] * inv);
            Gfb[r1 + J] = bb[J][k] * inv;
            Gfb[r1 + n + J] = ba[J][k] * inv;
          }
        }
      }
    }
  }
}


// 64 users (D = Mr = 128: n = 64 complex coordinates, 4 x 4 tiles of A and of B): the same algorithm with one lane per
// complex coordinate in the per-coordinate phases (lane j owns both real coordinates j and n + j), and the Gram's
// k-steps taken in pairs (s, s + 16): rows r = 4 s + grp (top) and r + 64 (bottom) of C1, whose D1 rows are each other
// (D1[r] = C1[r + 64], D1[r + 64] = -C1[r]). GIVEN: the MDI sub-problem build around given decisions (mimo.mdi's dDI
// step, arpeggio/mimo/mimo.py:570-590): x_g read from xg_in, no inversion, workgroup i builds frame list[i].
// Two waves per SIMD (256 registers, no accumulator registers): the sweep's serial pivot steps of one frame overlap
// another frame's; the few values that do not fit are spilled once per frame, in the Gram phase.
template <bool GIVEN>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void front_end_cpx128_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                              const float* __restrict__ N0v, float N0s, int M, int K,
                                                              float Q, FrontOut o, const int8_t* __restrict__ xg_in,
                                                              const int32_t* __restrict__ list) {
  constexpr int DP = 128, n = 64, nC = 4, nU = 10, Mr = 128, SH = 16, TS = 260;
  __shared__ __align__(16) float At[nU * TS];
  __shared__ __align__(16) float Bt[nU * TS];
  __shared__ __align__(16) float pan[n * 4];
  __shared__ __align__(16) float cz[2 * n];
  __shared__ __align__(16) float xs[4 * n];
  const int b = GIVEN ? list[blockIdx.x] : (int)blockIdx.x, lane = threadIdx.x, col = lane & 15, grp = lane >> 4;
  const float* Hb = H + (size_t)b * Mr * DP;
  const float* yb = y + (size_t)b * Mr;
  const float ylo = yb[lane], yhi = yb[n + lane];

  // 1. Gram and c over the row pairs (4 s + grp, 4 s + grp + 64)
  floatx4 TA[nU], TB[nU];
  float cpr[nC], cpi[nC];
#pragma unroll
  for (int u = 0; u < nU; ++u) {
    TA[u] = floatx4{0.f, 0.f, 0.f, 0.f};
    TB[u] = floatx4{0.f, 0.f, 0.f, 0.f};
  }
#pragma unroll
  for (int I = 0; I < nC; ++I) cpr[I] = cpi[I] = 0.f;
#pragma unroll
  for (int s = 0; s < SH; ++s) {
    if (s % 4 == 0) __builtin_amdgcn_sched_barrier(0);  // H loads in groups of four row pairs: bounded registers
    const int ra = 4 * s + grp;
    const floatx4 a = *reinterpret_cast<const floatx4*>(Hb + (size_t)ra * DP + nC * col);
    const floatx4 c = *reinterpret_cast<const floatx4*>(Hb + (size_t)(ra + n) * DP + nC * col);
    const float ya = __shfl(ylo, ra), yb2 = __shfl(yhi, ra);  // y[ra], y[ra + 64]
    // c_r = C1^T y; c_i = C1^T [y_bot; -y_top]: row ra takes y[ra + 64], row ra + 64 takes -y[ra]
#pragma unroll
    for (int I = 0; I < nC; ++I) {
      cpr[I] = __builtin_fmaf(c[I], yb2, __builtin_fmaf(a[I], ya, cpr[I]));
      cpi[I] = __builtin_fmaf(c[I], -ya, __builtin_fmaf(a[I], yb2, cpi[I]));
    }
#pragma unroll
    for (int I = 0; I < nC; ++I)
#pragma unroll
      for (int J = I; J < nC; ++J) {
        const int u = utile<nC>(I, J);
        TA[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[I], a[J], TA[u], 0, 0, 0);
        TA[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(c[I], c[J], TA[u], 0, 0, 0);
        TB[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[I], c[J], TB[u], 0, 0, 0);
        TB[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(c[I], -a[J], TB[u], 0, 0, 0);
      }
  }
  float cr[nC], ci[nC];
#pragma unroll
  for (int I = 0; I < nC; ++I) {
    float v = cpr[I], w = cpi[I];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    w += __shfl_xor(w, 16);
    w += __shfl_xor(w, 32);
    cr[I] = v;
    ci[I] = w;
    if (grp == 0) {
      cz[16 * I + col] = v;
      cz[n + 16 * I + col] = w;
    }
  }
#pragma unroll
  for (int u = 0; u < nU; ++u) {
    *reinterpret_cast<floatx4*>(At + u * TS + 16 * col + 4 * grp) = TA[u];
    *reinterpret_cast<floatx4*>(Bt + u * TS + 16 * col + 4 * grp) = TB[u];
  }
  __syncthreads();
#pragma unroll
  for (int I = 0; I < nC; ++I) {
    const float* bt = Bt + utile<nC>(I, I) * TS + col;
#pragma unroll
    for (int v = 0; v < 4; ++v) TB[utile<nC>(I, I)][v] = (TB[utile<nC>(I, I)][v] - bt[16 * (4 * grp + v)]) * 0.5f;
  }
  __syncthreads();
#pragma unroll
  for (int I = 0; I < nC; ++I)
    *reinterpret_cast<floatx4*>(Bt + utile<nC>(I, I) * TS + 16 * col + 4 * grp) = TB[utile<nC>(I, I)];
  const float el2 = M == 4 ? 1.f : 5.f;
  const float n0 = GIVEN ? 0.f : (N0v ? N0v[b] : N0s);
  const float sigma = (n0 / 2.f) / el2;
#pragma unroll
  for (int I = 0; I < nC; ++I)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (4 * grp + v == col) TA[utile<nC>(I, I)][v] += sigma;

  // 2. Complex block sweep (as front_end_cpx_kernel); the eight pivot pairs of a tile column run as a loop, so that
  // the 32 blocks do not all unroll (in-tile row q0 at run time: the pair's registers v0, v0 + 1 by selects)
  const int kk = grp & 1;
#pragma unroll
  for (int J0 = 0; J0 < (GIVEN ? 0 : nC); ++J0)
#pragma unroll 1
  for (int bq = 0; bq < 8; ++bq) {
    const int q0 = 2 * bq, g0 = q0 >> 2, v0 = q0 & 3, p0 = 16 * J0 + q0;
    if (grp == g0) {
#pragma unroll
      for (int I = J0; I < nC; ++I) {
        const floatx4 ta = TA[utile<nC>(J0, I)], tb = TB[utile<nC>(J0, I)];
        *reinterpret_cast<floatx4*>(pan + (16 * I + col) * 4) =
            floatx4{v0 ? ta[2] : ta[0], v0 ? ta[3] : ta[1], v0 ? tb[2] : tb[0], v0 ? tb[3] : tb[1]};
      }
    }
    if ((col >> 1) == (q0 >> 1)) {
      const int e = col - q0;
#pragma unroll
      for (int I = 0; I < J0; ++I)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          pan[(16 * I + 4 * grp + v) * 4 + e] = TA[utile<nC>(I, J0)][v];
          pan[(16 * I + 4 * grp + v) * 4 + 2 + e] = -TB[utile<nC>(I, J0)][v];
        }
    }
    __syncthreads();
    const floatx4 r0 = *reinterpret_cast<const floatx4*>(pan + p0 * 4);
    const floatx4 r1 = *reinterpret_cast<const floatx4*>(pan + (p0 + 1) * 4);
    const float a = r0[0], d = r1[1], al = r1[0], be = r1[2];
    const float rdet = __builtin_amdgcn_rcpf(__builtin_fmaf(a, d, -__builtin_fmaf(al, al, be * be)));
    const float m00 = d * rdet, m11 = a * rdet, m01r = -al * rdet, m01i = -be * rdet;
    const float Mr0 = kk ? m01r : m00, Mi0 = kk ? m01i : 0.f, Mr1 = kk ? m11 : m01r, Mi1 = kk ? 0.f : -m01i;
    const float c0r = lane_val(cr[J0], q0), c1r = lane_val(cr[J0], q0 + 1);
    const float c0i = lane_val(ci[J0], q0), c1i = lane_val(ci[J0], q0 + 1);
    const float q0r = __builtin_fmaf(m00, c0r, __builtin_fmaf(m01r, c1r, -m01i * c1i));
    const float q0i = __builtin_fmaf(m00, c0i, __builtin_fmaf(m01r, c1i, m01i * c1r));
    const float q1r = __builtin_fmaf(m01r, c0r, __builtin_fmaf(m01i, c0i, m11 * c1r));
    const float q1i = __builtin_fmaf(m01r, c0i, __builtin_fmaf(-m01i, c0r, m11 * c1i));
    float nuA[nC], nuB[nC], vt[nC];
#pragma unroll
    for (int I = 0; I < nC; ++I) {
      const floatx4 x = *reinterpret_cast<const floatx4*>(pan + (16 * I + col) * 4);
      const float vr0 = x[0] - (I == J0 && col == q0 ? 1.f : 0.f), vr1 = x[1] - (I == J0 && col == q0 + 1 ? 1.f : 0.f);
      const float vi0 = -x[2], vi1 = -x[3];
      const float ur = __builtin_fmaf(vr0, Mr0, __builtin_fmaf(-vi0, Mi0, __builtin_fmaf(vr1, Mr1, -vi1 * Mi1)));
      const float ui = __builtin_fmaf(vr0, Mi0, __builtin_fmaf(vi0, Mr0, __builtin_fmaf(vr1, Mi1, vi1 * Mr1)));
      nuA[I] = grp < 2 ? -ur : -ui;
      nuB[I] = grp < 2 ? -ui : ur;
      cr[I] -= __builtin_fmaf(vr0, q0r, __builtin_fmaf(-vi0, q0i, __builtin_fmaf(vr1, q1r, -vi1 * q1i)));
      ci[I] -= __builtin_fmaf(vr0, q0i, __builtin_fmaf(vi0, q0r, __builtin_fmaf(vr1, q1i, vi1 * q1r)));
      const float w = pan[(16 * I + col) * 4 + grp];
      vt[I] = grp < 2 ? w - (I == J0 && col == q0 + grp ? 1.f : 0.f) : -w;
    }
#pragma unroll
    for (int I = 0; I < nC; ++I)
#pragma unroll
      for (int J = I; J < nC; ++J) {
        TA[utile<nC>(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(nuA[I], vt[J], TA[utile<nC>(I, J)], 0, 0, 0);
        TB[utile<nC>(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(nuB[I], vt[J], TB[utile<nC>(I, J)], 0, 0, 0);
      }
  }

  // 3. Lane j = complex coordinate (position p = 16 (j % 4) + j / 4): both real coordinates j and n + j.
  const int j = lane, Ii = j % nC, mi = j / nC, pp = 16 * Ii + mi;
  const int src = mi + 16 * (mi >> 2);
  float zr = 0.f, zi = 0.f, dgi = 0.f;
#pragma unroll
  for (int I = 0; I < nC; ++I) {
    const float a = __shfl(cr[I], mi), c = __shfl(ci[I], mi);
    const float td = __shfl(pick4c(TA[utile<nC>(I, I)], col & 3), src);
    if (Ii == I) {
      zr = a;
      zi = c;
      dgi = 2.f - td;
    }
  }
  const float gii = At[(Ii * nC - Ii * (Ii - 1) / 2) * TS + 17 * mi];
  const float c0r = cz[pp], c0i = cz[n + pp];
  const float lmax = M == 4 ? 1.f : 3.f;
  float lr, li, vr = 0.f, vi = 0.f, mu = 1.f;
  if (GIVEN) {
    lr = (float)xg_in[(size_t)b * DP + j];
    li = (float)xg_in[(size_t)b * DP + n + j];
  } else {
    mu = 1.f - sigma * dgi;
    vr = zr / mu;
    vi = zi / mu;
    lr = fminf(fmaxf(2.f * ceilf(vr * 0.5f) - 1.f, -lmax), lmax);
    li = fminf(fmaxf(2.f * ceilf(vi * 0.5f) - 1.f, -lmax), lmax);
    if (vr != vr) lr = -lmax;
    if (vi != vi) li = -lmax;
    if (o.z) {
      o.z[(size_t)b * DP + j] = zr;
      o.z[(size_t)b * DP + n + j] = zi;
    }
    if (o.mu) {
      o.mu[(size_t)b * DP + j] = mu;
      o.mu[(size_t)b * DP + n + j] = mu;
    }
    if (o.xg) {
      o.xg[(size_t)b * DP + j] = (int8_t)lr;
      o.xg[(size_t)b * DP + n + j] = (int8_t)li;
    }
    if (o.feat) {
      float gs = gii;
      for (int off = 32; off; off >>= 1) gs += __shfl_xor(gs, off);
      const float gbar = gs / (float)n;   // each A_jj is G_ii for two real coordinates
      uint16_t* f = o.feat + (size_t)b * o.feat_ld;
      f[j] = bf16_bits(c0r / gbar);
      f[n + j] = bf16_bits(c0i / gbar);
      f[DP + j] = bf16_bits(gii / gbar - 1.f);
      f[DP + n + j] = bf16_bits(gii / gbar - 1.f);
      f[2 * DP + j] = bf16_bits(lr / lmax);
      f[2 * DP + n + j] = bf16_bits(li / lmax);
      f[3 * DP + j] = bf16_bits(vr - lr);
      f[3 * DP + n + j] = bf16_bits(vi - li);
      for (int x = 4 * DP + lane; x < o.feat_ld; x += 64) f[x] = bf16_bits(x == 4 * DP ? sigma / gbar : 0.f);
    }
  }

  // 4. h = 2 (c - G x_g) for both parts of row j: A_j x_r - B_j x_i and B_j x_r + A_j x_i
  xs[pp] = lr;
  xs[n + pp] = li;
  __syncthreads();
  float g0 = 0.f, g1 = 0.f, gm = 0.f;
#pragma unroll
  for (int J = 0; J < nC; ++J) {
    const bool up = J >= Ii;
    const int tl = up ? Ii * nC - Ii * (Ii - 1) / 2 + (J - Ii) : J * nC - J * (J - 1) / 2 + (Ii - J);
    const float* ga = At + tl * TS + (up ? mi : 16 * mi);
    const float* gb = Bt + tl * TS + (up ? mi : 16 * mi);
    const int st = up ? 16 : 1;
    const float sb = up ? 1.f : -1.f;
#pragma unroll
    for (int c4 = 0; c4 < 16; c4 += 4) {
      const floatx4 ur = *reinterpret_cast<const floatx4*>(xs + 16 * J + c4);
      const floatx4 ui = *reinterpret_cast<const floatx4*>(xs + n + 16 * J + c4);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float ga_ = ga[st * (c4 + e)], gb_ = sb * gb[st * (c4 + e)];
        g0 = __builtin_fmaf(-gb_, ui[e], __builtin_fmaf(ga_, ur[e], g0));
        g1 = __builtin_fmaf(ga_, ui[e], __builtin_fmaf(gb_, ur[e], g1));
        gm = fmaxf(gm, fmaxf(fabsf(ga_), fabsf(gb_)));
      }
    }
  }
  const float h0 = 2.f * (c0r - g0), h1 = 2.f * (c0i - g1);
  float mx = fmaxf(gm, fmaxf(fabsf(h0), fabsf(h1)));
  for (int off = 32; off; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  const float scale = fmaxf(mx, 1e-12f);
  const float inv = 1.f / scale;
  if (lane == 0 && o.scale) o.scale[b] = scale;
  {
    const int P = 2 * K, N = P * DP;
    const float hs0 = h0 * inv, hs1 = h1 * inv;
    const int32_t q0 = (int32_t)rintf(hs0 * Q), q1 = (int32_t)rintf(hs1 * Q);
    if (o.hi) {
      o.hi[(size_t)b * DP + j] = q0;
      o.hi[(size_t)b * DP + n + j] = q1;
    }
    if (o.hf) {
      o.hf[(size_t)b * DP + j] = hs0;
      o.hf[(size_t)b * DP + n + j] = hs1;
    }
    for (int p = 0; p < P; ++p) {
      if (o.hN) {
        o.hN[(size_t)b * N + p * DP + j] = q0;
        o.hN[(size_t)b * N + p * DP + n + j] = q1;
      }
      if (o.hNf) {
        o.hNf[(size_t)b * N + p * DP + j] = hs0;
        o.hNf[(size_t)b * N + p * DP + n + j] = hs1;
      }
    }
  }

  // 5. G / scale row-major, rows (part, jr), jr = 4 (4 grp + k) + I, columns 4 col .. 4 col + 3 of each half
  if (o.Gi || o.Gf || o.Gh) {
    int32_t* Gib = o.Gi ? o.Gi + (size_t)b * DP * DP : nullptr;
    int16_t* Ghb = o.Gh ? o.Gh + (size_t)b * DP * DP : nullptr;
    float* Gfb = o.Gf ? o.Gf + (size_t)b * DP * DP : nullptr;
    const float invQ = inv * Q;
#pragma unroll
    for (int I = 0; I < nC; ++I) {
      floatx4 ba[nC], bb[nC];
#pragma unroll
      for (int J = 0; J < nC; ++J) {
        if (J >= I) {
          ba[J] = *reinterpret_cast<const floatx4*>(At + utile<nC>(I, J) * TS + 16 * col + 4 * grp);
          bb[J] = *reinterpret_cast<const floatx4*>(Bt + utile<nC>(I, J) * TS + 16 * col + 4 * grp);
        } else {
          const float* a0 = At + utile<nC>(J, I) * TS + 64 * grp + col;
          const float* b0 = Bt + utile<nC>(J, I) * TS + 64 * grp + col;
          ba[J] = floatx4{a0[0], a0[16], a0[32], a0[48]};
          bb[J] = floatx4{-b0[0], -b0[16], -b0[32], -b0[48]};
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int jr = nC * (4 * grp + k) + I;
        const size_t r0 = (size_t)jr * DP + nC * col, r1 = (size_t)(n + jr) * DP + nC * col;
        uint32_t qa[nC], qb[nC], qn[nC];
#pragma unroll
        for (int J = 0; J < nC; ++J) {
          qa[J] = rne_bits(ba[J][k] * invQ);
          qb[J] = rne_bits(bb[J][k] * invQ);
          qn[J] = 0u - qb[J];
        }
        if (Ghb) {
          const uint2 ua = make_uint2(__builtin_amdgcn_perm(qa[1], qa[0], 0x05040100u),
                                      __builtin_amdgcn_perm(qa[3], qa[2], 0x05040100u));
          *reinterpret_cast<uint2*>(Ghb + r0) = ua;
          *reinterpret_cast<uint2*>(Ghb + r0 + n) = make_uint2(__builtin_amdgcn_perm(qn[1], qn[0], 0x05040100u),
                                                               __builtin_amdgcn_perm(qn[3], qn[2], 0x05040100u));
          *reinterpret_cast<uint2*>(Ghb + r1) = make_uint2(__builtin_amdgcn_perm(qb[1], qb[0], 0x05040100u),
                                                           __builtin_amdgcn_perm(qb[3], qb[2], 0x05040100u));
          *reinterpret_cast<uint2*>(Ghb + r1 + n) = ua;
        }
        if (Gib) {
          *reinterpret_cast<int4*>(Gib + r0) = make_int4((int)qa[0], (int)qa[1], (int)qa[2], (int)qa[3]);
          *reinterpret_cast<int4*>(Gib + r0 + n) = make_int4((int)qn[0], (int)qn[1], (int)qn[2], (int)qn[3]);
          *reinterpret_cast<int4*>(Gib + r1) = make_int4((int)qb[0], (int)qb[1], (int)qb[2], (int)qb[3]);
          *reinterpret_cast<int4*>(Gib + r1 + n) = make_int4((int)qa[0], (int)qa[1], (int)qa[2], (int)qa[3]);
        }
        if (Gfb) {
#pragma unroll
          for (int J = 0; J < nC; ++J) {
            Gfb[r0 + J] = ba[J][k] * inv;
            Gfb[r0 + n + J] = -(bb[J][k] * inv);
            Gfb[r1 + J] = bb[J][k] * inv;
            Gfb[r1 + n + J] = ba[J][k] * inv;
          }
        }
      }
    }
  }
}

bool front_end_cpx(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M, int K,
                   float Q, const FrontOut& o, hipStream_t st) {
  int qe = 0;
  if (Mr != D || (D != 32 && D != 64 && D != 128) || o.J || o.Jf || !(Q > 0.f) || std::frexp(Q, &qe) != 0.5f)
    return false;
  if (D == 64 && reinterpret_cast<uintptr_t>(H) % 8) return false;    // float2 loads of a lane's two columns
  if (D == 128 && reinterpret_cast<uintptr_t>(H) % 16) return false;  // floatx4 loads of a lane's four columns
  if (D == 128)
    hipLaunchKernelGGL((front_end_cpx128_kernel<false>), dim3(B), dim3(64), 0, st, H, y, N0v, N0s, M, K, Q, o, nullptr,
                       nullptr);
  else if (D == 64)
    hipLaunchKernelGGL((front_end_cpx_kernel<64>), dim3(B), dim3(64), 0, st, H, y, N0v, N0s, M, K, Q, o);
  else
    hipLaunchKernelGGL((front_end_cpx_kernel<32>), dim3(B), dim3(64), 0, st, H, y, N0v, N0s, M, K, Q, o);
  return true;
}

}  // namespace arp

namespace arp {

bool mdi_build_cpx128(const float* H, const float* y, const int8_t* xg, const int32_t* list, int n, float Q,
                      int32_t* Gi, int32_t* hi, hipStream_t st) {
  int qe = 0;
  if (!(Q > 0.f) || std::frexp(Q, &qe) != 0.5f || reinterpret_cast<uintptr_t>(H) % 16) return false;
  FrontOut o{};
  o.Gi = Gi;
  o.hi = hi;
  hipLaunchKernelGGL((front_end_cpx128_kernel<true>), dim3(n), dim3(64), 0, st, H, y, nullptr, 1.f, 16, 1, Q, o, xg,
                     list);
  return true;
}

}  // namespace arp
