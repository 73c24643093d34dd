// Front end for 16 and 32 users (D, Mr <= 64) with the whole MMSE inverse on the matrix cores.
//
// The same outputs as the register-sweep front end (front_end.hip) from a different schedule. One wave per frame:
//   1. G = H^T H in v_mfma_f32_16x16x4f32 accumulators (the upper 16x16 tiles of the coordinates in tile order,
//      position 16 I + m = coordinate nT m + I, so that a lane's operand columns are adjacent in memory), H read once
//      from global memory straight into the operands; c = H^T y by the VALU beside it. A copy of G waits in LDS
//      (10.2 KiB at D = 64: the upper tiles, each column-major) for the outputs.
//   2. A = G + sigma I (identity on the padding) is swept in place in the accumulators, four pivots at a time. For the
//      pivot set S = {k0 .. k0+3} the block sweep is
//          A_SS -> -A_SS^-1,  A_RS -> A_RS A_SS^-1,  A_SR -> A_SS^-1 A_SR,  A_RR -> A_RR - A_RS A_SS^-1 A_SR,
//      the composition of the four single-pivot sweeps. With V = A_.S - E_.S (E: ones at (k, k), k in S), all four
//      parts are one symmetric rank-4 update A <- A - (V A_SS^-1) V^T on every tile, except that the diagonal of A_SS
//      comes out 2 too high; no later block reads those entries, so the 2 comes off once, at the end. The update is
//      one MFMA per upper tile (A operand V A_SS^-1, B operand V^T); the panel (rows S, every column) goes through a
//      1 KiB LDS buffer. A_SS (symmetric positive definite: the Schur complement of the pivots swept so far) is
//      factored as L D L^T in registers, and each lane solves for the one column of A_SS^-1 it feeds to the update
//      (its group's) and for A_SS^-1 c_S. The right-hand side c rides along as an extra column
//      (c <- c - V A_SS^-1 c_S), so it ends as z = A^-1 c.
//   3. mu_i = 1 - sigma (A^-1)_ii, x_g = nearest level of z / mu (ties to the lower level).
//   4. h = 2 H^T (y - H x_g) (mimo.ising_delta_at) as 2 (c - G x_g), one row of the LDS copy per lane;
//      scale = max(max|h|, max|G|), clamped at 1e-12; G_int = rne(G / scale * Q), h_int = rne(h / scale * Q),
//      G written row-major: 16-byte loads of four rows' entries from the LDS copy, whole rows per store instruction;
//      rne by the bits of x + 1.5 2^23 (exact for |x| < 2^22), so two int16 entries pack with one byte permute.
// Reference semantics: mimo.mmse / _mu / quantize_levels (arpeggio/mimo/mimo.py:86-109, 176-182), mimo.ising_delta_at
// (mimo.py:506-523), chordfx.to_fixed (arpeggio/mimo/chordfx.py:60-76), chord_ref.compress
// (arpeggio/sim/x3model/chord_ref.py:47-59). The expanded Chord inputs (J, Jf) are the register-sweep kernel's.
#include "kernels.h"

#include <cmath>

namespace arp {

typedef float floatx4 __attribute__((ext_vector_type(4)));

static int g_front_end_form = 0;
void set_front_end_form(int form) { g_front_end_form = form; }
int front_end_form() { return g_front_end_form; }

__device__ __forceinline__ uint16_t bf16_of(float f) {
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// entry i (runtime, 0..3) of a 4-vector
__device__ __forceinline__ float pick4(floatx4 v, int i) {
  const float a = i & 1 ? v[1] : v[0];
  const float b = i & 1 ? v[3] : v[2];
  return i & 2 ? b : a;
}

__device__ __forceinline__ float lane_value(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// index of upper tile (I, J), I <= J, of an nT x nT tiling
template <int nT>
__device__ constexpr int tile(int I, int J) {
  return I * nT - I * (I - 1) / 2 + (J - I);
}

// Register budget: three waves per SIMD at D = 64 (161 VGPRs, 12.7 KiB of LDS), four below.
// GIVEN: the MDI sub-problem build around a given decision (mimo.mdi's dDI step, arpeggio/mimo/mimo.py:570-590):
// x_g is read from xg_in instead of computed, the inversion (phases 2-3) is skipped, and workgroup i builds frame
// list[i] (written at the frame's index).
template <int DP, bool GIVEN>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(DP >= 64 ? 3 : 4))) void front_end_mfma_kernel(
    const float* __restrict__ H, const float* __restrict__ y, const float* __restrict__ N0v, float N0s, int Mr, int D,
    int M, int K, float Q, int vec, FrontOut o, const int8_t* __restrict__ xg_in, const int32_t* __restrict__ list) {
  constexpr int nT = DP / 16, nU = nT * (nT + 1) / 2;
  constexpr int TS = 260;                               // tile stride in words: 16 x 16 entries + 4 (bank offset)
  __shared__ __align__(16) float Gt[nU * TS];           // the Gram's upper tiles, each column-major
  __shared__ __align__(16) float pan[2][DP * 4];        // panel rows S: pan[p * 4 + t] = A[k0 + t][p]
  __shared__ __align__(16) float bc[64];                // x_g in tile order
  const int b = GIVEN ? list[blockIdx.x] : (int)blockIdx.x, lane = threadIdx.x, col = lane & 15, grp = lane >> 4;
  const float* Hb = H + (size_t)b * Mr * D;
  const float* yb = y + (size_t)b * Mr;

  // Tile order: position p = 16 I + m of the 16 x 16 tiling is coordinate j = nT m + I, so the nT columns a lane
  // supplies to the tile rows (nT col .. nT col + nT - 1) are adjacent in memory. Every row of H is read once, by the
  // 16 lanes of a group with one vector load each, all loads issued before the first product.
  float opv[16][nT];
  const float ylane = lane < Mr ? yb[lane] : 0.f;  // y_r reaches the lanes of row r by a shuffle
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int r = 4 * s + grp;
    const bool rv = r < Mr;
    const float* hr = Hb + (size_t)r * D + nT * col;
    if (vec && rv) {
      if constexpr (nT == 4) {
        const floatx4 x = *reinterpret_cast<const floatx4*>(hr);
#pragma unroll
        for (int I = 0; I < 4; ++I) opv[s][I] = x[I];
      } else if constexpr (nT == 2) {
        const float2 x = *reinterpret_cast<const float2*>(hr);
        opv[s][0] = x.x;
        opv[s][1] = x.y;
      } else {
#pragma unroll
        for (int I = 0; I < nT; ++I) opv[s][I] = hr[I];
      }
    } else {
#pragma unroll
      for (int I = 0; I < nT; ++I) opv[s][I] = (rv && nT * col + I < D) ? hr[I] : 0.f;
    }
  }
  // 1. Gram tiles and c = H^T y: at k-step s, lane (col, grp) supplies H[4 s + grp][nT col + I] to tile row I (A
  // operand) and tile column I (B operand).
  floatx4 T[nU];
  float cp[nT];
#pragma unroll
  for (int u = 0; u < nU; ++u) T[u] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int I = 0; I < nT; ++I) cp[I] = 0.f;
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    if (4 * s >= Mr) break;
    const float yr = __shfl(ylane, 4 * s + grp);
#pragma unroll
    for (int I = 0; I < nT; ++I) cp[I] = __builtin_fmaf(opv[s][I], yr, cp[I]);
#pragma unroll
    for (int I = 0; I < nT; ++I)
#pragma unroll
      for (int J = I; J < nT; ++J)
        T[tile<nT>(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(opv[s][I], opv[s][J], T[tile<nT>(I, J)], 0, 0, 0);
  }
  // c[I] = c at position 16 I + col, the same in all four groups
  float c[nT], c0[nT];
#pragma unroll
  for (int I = 0; I < nT; ++I) {
    float v = cp[I];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    c[I] = v;
    c0[I] = v;
  }
  // G waits in LDS for the outputs as its upper tiles: tile (I, J), I <= J, entry (r, c) = G[nT r + I][nT c + J] at
  // Gt[tile(I, J) TS + 16 c + r]. Lane (grp, col) holds entries (4 grp + v, col): one 16-byte store per tile. 12.7 KiB
  // of LDS per wave in all, so that three waves fit a SIMD.
#pragma unroll
  for (int u = 0; u < nU; ++u) *reinterpret_cast<floatx4*>(Gt + u * TS + 16 * col + 4 * grp) = T[u];
  const float el2 = M == 4 ? 1.f : 5.f;
  const float n0 = N0v ? N0v[b] : N0s;
  const float sigma = (n0 / 2.f) / el2;
  // A = G + sigma I; identity on the padding. Diagonal tile (I, I) holds positions (16 I + 4 grp + v, 16 I + col).
#pragma unroll
  for (int I = 0; I < nT; ++I)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (4 * grp + v == col) T[tile<nT>(I, I)][v] = nT * col + I < D ? T[tile<nT>(I, I)][v] + sigma : 1.f;

  float eg[4];  // e_grp: this lane's right-hand side of the pivot-block solve
#pragma unroll
  for (int t = 0; t < 4; ++t) eg[t] = grp == t ? 1.f : 0.f;
  // 2. Block sweep over the pivot sets {k0 .. k0 + 3} of the tile order (padding pivots included: identity rows
  // sweep to themselves); the panel buffer alternates between two halves, so a block's stores never meet the
  // previous block's loads.
#pragma unroll
  for (int blk = 0; blk < (GIVEN ? 0 : DP / 4); ++blk) {
    const int k0 = 4 * blk, J0 = k0 / 16, g0 = (k0 % 16) / 4;
    float* P = pan[blk & 1];
    // rows k0 + s of tiles (J0, I >= J0) sit in group g0, register s; for I < J0 the same entries A[j][k0 + s] are
    // column 4 g0 + s of tile (I, J0)
    if (grp == g0) {
#pragma unroll
      for (int I = J0; I < nT; ++I) *reinterpret_cast<floatx4*>(P + (16 * I + col) * 4) = T[tile<nT>(J0, I)];
    }
    if ((col >> 2) == g0) {
#pragma unroll
      for (int I = 0; I < J0; ++I)
#pragma unroll
        for (int v = 0; v < 4; ++v) P[(16 * I + 4 * grp + v) * 4 + (col & 3)] = T[tile<nT>(I, J0)][v];
    }
    __syncthreads();
    floatx4 Vr[nT];
    float vt[nT];
#pragma unroll
    for (int I = 0; I < nT; ++I) {
      Vr[I] = *reinterpret_cast<const floatx4*>(P + (16 * I + col) * 4);  // A[k0 + t][16 I + col], t = 0..3
      vt[I] = P[(16 * I + col) * 4 + grp];                                 // A[k0 + grp][16 I + col]
    }
    floatx4 S4[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) S4[t] = *reinterpret_cast<const floatx4*>(P + (k0 + t) * 4);
    // the four c_S entries, from the lanes that hold them (col = 4 g0 + t)
    float cS[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) cS[t] = lane_value(c[J0], 4 * g0 + t);
    // A_SS = L D L^T (A_SS is the Schur complement of the pivots swept so far: symmetric positive definite), then two
    // solves: x = A_SS^-1 e_grp (column grp of the inverse, a lane's own right-hand side) and qv = -A_SS^-1 c_S.
    const float a00 = S4[0][0], a01 = S4[1][0], a02 = S4[2][0], a03 = S4[3][0];
    const float a11 = S4[1][1], a12 = S4[2][1], a13 = S4[3][1], a22 = S4[2][2], a23 = S4[3][2], a33 = S4[3][3];
    const float r0 = __builtin_amdgcn_rcpf(a00);
    const float l10 = a01 * r0, l20 = a02 * r0, l30 = a03 * r0;
    const float r1 = __builtin_amdgcn_rcpf(__builtin_fmaf(-l10, a01, a11));
    const float t21 = __builtin_fmaf(-l20, a01, a12), t31 = __builtin_fmaf(-l30, a01, a13);
    const float l21 = t21 * r1, l31 = t31 * r1;
    const float r2 = __builtin_amdgcn_rcpf(__builtin_fmaf(-l21, t21, __builtin_fmaf(-l20, a02, a22)));
    const float t32 = __builtin_fmaf(-l31, t21, __builtin_fmaf(-l30, a02, a23));
    const float l32 = t32 * r2;
    const float r3 =
        __builtin_amdgcn_rcpf(__builtin_fmaf(-l32, t32, __builtin_fmaf(-l31, t31, __builtin_fmaf(-l30, a03, a33))));
    auto solve = [&](float b0, float b1, float b2, float b3, float (&x)[4]) {
      const float y1 = __builtin_fmaf(-l10, b0, b1);
      const float y2 = __builtin_fmaf(-l21, y1, __builtin_fmaf(-l20, b0, b2));
      const float y3 = __builtin_fmaf(-l32, y2, __builtin_fmaf(-l31, y1, __builtin_fmaf(-l30, b0, b3)));
      x[3] = y3 * r3;
      x[2] = __builtin_fmaf(-l32, x[3], y2 * r2);
      x[1] = __builtin_fmaf(-l31, x[3], __builtin_fmaf(-l21, x[2], y1 * r1));
      x[0] = __builtin_fmaf(-l30, x[3], __builtin_fmaf(-l20, x[2], __builtin_fmaf(-l10, x[1], b0 * r0)));
    };
    // V = A_.S - E: minus one where the row is the pivot itself
#pragma unroll
    for (int t = 0; t < 4; ++t) Vr[J0][t] -= col == 4 * g0 + t ? 1.f : 0.f;
    vt[J0] -= col == 4 * g0 + grp ? 1.f : 0.f;
    float mg[4], qv[4];
    solve(eg[0], eg[1], eg[2], eg[3], mg);   // mg = A_SS^-1 e_grp; the update takes -mg
    solve(cS[0], cS[1], cS[2], cS[3], qv);   // qv = A_SS^-1 c_S; c takes -V qv
    float nu[nT];  // -(V A_SS^-1)[16 I + col][grp]: the A operand of the update
#pragma unroll
    for (int I = 0; I < nT; ++I) {
      nu[I] = __builtin_fmaf(-Vr[I][3], mg[3], __builtin_fmaf(-Vr[I][2], mg[2],
                             __builtin_fmaf(-Vr[I][1], mg[1], -Vr[I][0] * mg[0])));
      c[I] = __builtin_fmaf(-Vr[I][3], qv[3], __builtin_fmaf(-Vr[I][2], qv[2],
                            __builtin_fmaf(-Vr[I][1], qv[1], __builtin_fmaf(-Vr[I][0], qv[0], c[I]))));
    }
#pragma unroll
    for (int I = 0; I < nT; ++I)
#pragma unroll
      for (int J = I; J < nT; ++J)
        T[tile<nT>(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(nu[I], vt[J], T[tile<nT>(I, J)], 0, 0, 0);
  }

  // 3. Per coordinate i = lane, at position p = 16 (i % nT) + i / nT: z_i = c at p (register i % nT of lane i / nT);
  // (A^-1)_ii = 2 - T_pp, held by lane m + 16 (m >> 2), m = i / nT, register m & 3 of diagonal tile i % nT (G_ii
  // likewise, from LDS).
  const bool valid = lane < D;
  const int Ii = lane % nT, mi = lane / nT;
  const int src = (mi & 15) + 16 * ((mi & 15) >> 2);
  float z = 0.f, dgi = 0.f, cc0 = 0.f;
#pragma unroll
  for (int I = 0; I < nT; ++I) {
    const float zc = __shfl(c[I], mi & 15);
    const float c0c = __shfl(c0[I], mi & 15);
    const float td = __shfl(pick4(T[tile<nT>(I, I)], col & 3), src);
    if (Ii == I) {
      z = zc;
      cc0 = c0c;
      dgi = 2.f - td;
    }
  }
  float gii = 0.f;
  if (valid) gii = Gt[(Ii * nT - Ii * (Ii - 1) / 2) * TS + 17 * mi];
  const float mu = 1.f - sigma * dgi;
  const float vq = z / mu;
  const float lmax = M == 4 ? 1.f : 3.f;
  float l = fminf(fmaxf(2.f * ceilf(vq * 0.5f) - 1.f, -lmax), lmax);
  if (vq != vq) l = -lmax;
  if (GIVEN) l = valid ? (float)xg_in[(size_t)b * D + lane] : 0.f;
  if (!valid) l = 0.f;
  if (valid && !GIVEN) {
    if (o.z) o.z[(size_t)b * D + lane] = z;
    if (o.mu) o.mu[(size_t)b * D + lane] = mu;
    if (o.xg) o.xg[(size_t)b * D + lane] = (int8_t)l;
  }
  if (o.feat && !GIVEN) {
    float gs = gii;
    for (int off = 32; off; off >>= 1) gs += __shfl_xor(gs, off);
    const float gbar = gs / (float)D;
    uint16_t* f = o.feat + (size_t)b * o.feat_ld;
    if (valid) {
      f[lane] = bf16_of(cc0 / gbar);
      f[D + lane] = bf16_of(gii / gbar - 1.f);
      f[2 * D + lane] = bf16_of(l / lmax);
      f[3 * D + lane] = bf16_of(vq - l);
    }
    for (int x = 4 * D + lane; x < o.feat_ld; x += 64) f[x] = bf16_of(x == 4 * D ? sigma / gbar : 0.f);
  }

  // 4. h = 2 H^T (y - H x_g) (mimo.ising_delta_at) as 2 (c - G x_g): lane i = nT m + I takes the dot product of row
  // i with x_g (x_g in tile order, read as broadcasts). Row i is row m of tiles (I, J >= I), entries (m, c) at 16 c + m,
  // and column m of tiles (J < I, I), entries (c, m) at 16 m + c: a per-lane base and stride for each J. The row's
  // largest |G| rides along. scale = max(max|h|, max|G|), clamped at 1e-12.
  if (lane < DP) bc[16 * Ii + mi] = l;
  __syncthreads();
  float gx = 0.f, gm = 0.f;
  if (lane < DP) {
#pragma unroll
    for (int J = 0; J < nT; ++J) {
      const bool up = J >= Ii;
      const int tl = up ? Ii * nT - Ii * (Ii - 1) / 2 + (J - Ii) : J * nT - J * (J - 1) / 2 + (Ii - J);
      const float* g0 = Gt + tl * TS + (up ? mi : 16 * mi);
      const int st = up ? 16 : 1;
#pragma unroll
      for (int c4 = 0; c4 < 16; c4 += 4) {
        const floatx4 xv = *reinterpret_cast<const floatx4*>(bc + 16 * J + c4);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float g = g0[st * (c4 + e)];
          gx = __builtin_fmaf(g, xv[e], gx);
          gm = fmaxf(gm, fabsf(g));   // padding entries are zero
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
  const int P = 2 * K, N = P * D;
  if (valid) {
    const float hs = hval * inv;
    const int32_t h32 = (int32_t)rintf(hs * Q);
    if (o.hi) o.hi[(size_t)b * D + lane] = h32;
    if (o.hf) o.hf[(size_t)b * D + lane] = hs;
    for (int p = 0; p < P; ++p) {
      if (o.hN) o.hN[(size_t)b * N + p * D + lane] = h32;
      if (o.hNf) o.hNf[(size_t)b * N + p * D + lane] = hs;
    }
  }
  // 5. G / scale in row-major order. For each tile row I, lane (grp, col) reads G[nT (4 grp + k) + I][nT col + J],
  // k = 0..3, J < nT: entries (4 grp + k, col) of tile (I, J) for J >= I (one 16-byte load), entries (col, 4 grp + k)
  // of tile (J, I) for J < I (four loads at a constant stride). It writes rows nT (4 grp + k) + I, columns
  // nT col .. nT col + nT - 1, so each store instruction covers whole rows.
  if (o.Gi || o.Gf || o.Gh) {
    int32_t* Gib = o.Gi ? o.Gi + (size_t)b * D * D : nullptr;
    int16_t* Ghb = o.Gh ? o.Gh + (size_t)b * D * D : nullptr;
    float* Gfb = o.Gf ? o.Gf + (size_t)b * D * D : nullptr;
    const bool full = D == DP;
    const float invQ = inv * Q;   // rne(G invQ) = rne((G / scale) Q): Q is a power of two
    // with D = DP, row i = nT (4 grp + k) + I starts at lane base + (nT k + I) DP: the offsets are constants
    const size_t lb = (size_t)4 * nT * grp * DP + nT * col;
    int32_t* gil = Gib ? Gib + lb : nullptr;
    int16_t* ghl = Ghb ? Ghb + lb : nullptr;
    float* gfl = Gfb ? Gfb + lb : nullptr;
#pragma unroll
    for (int I = 0; I < nT; ++I) {
      floatx4 blk[nT];
#pragma unroll
      for (int J = 0; J < nT; ++J) {
        if (J >= I) {
          blk[J] = *reinterpret_cast<const floatx4*>(Gt + tile<nT>(I, J) * TS + 16 * col + 4 * grp);
        } else {
          const float* g0 = Gt + tile<nT>(J, I) * TS + 64 * grp + col;
          blk[J] = floatx4{g0[0], g0[16], g0[32], g0[48]};
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int i = nT * (4 * grp + k) + I, j0 = nT * col;
        const int at = (nT * k + I) * DP;
        // rne(g invQ) from the bits of g invQ + 1.5 2^23 (|g invQ| <= Q <= 2^14: the sum's low mantissa bits hold the
        // rounded integer, ties to even as rintf), so an int16 pair is one byte permute
        uint32_t gb[4];
#pragma unroll
        for (int J = 0; J < 4; ++J) {
          const float g = J < nT ? blk[J % nT][k] : 0.f;
          gb[J] = __builtin_bit_cast(uint32_t, g * invQ + 12582912.f) - 0x4B400000u;
        }
        if (full && nT == 4) {
          if (gil) *reinterpret_cast<int4*>(gil + at) = make_int4((int)gb[0], (int)gb[1], (int)gb[2], (int)gb[3]);
          if (ghl)
            *reinterpret_cast<uint2*>(ghl + at) =
                make_uint2(__builtin_amdgcn_perm(gb[1], gb[0], 0x05040100u), __builtin_amdgcn_perm(gb[3], gb[2], 0x05040100u));
          if (gfl) *reinterpret_cast<floatx4*>(gfl + at) = floatx4{blk[0][k] * inv, blk[1][k] * inv, blk[2][k] * inv,
                                                                 blk[3][k] * inv};
        } else if (full && nT == 2) {
          if (gil) *reinterpret_cast<int2*>(gil + at) = make_int2((int)gb[0], (int)gb[1]);
          if (ghl) *reinterpret_cast<uint32_t*>(ghl + at) = __builtin_amdgcn_perm(gb[1], gb[0], 0x05040100u);
          if (gfl) *reinterpret_cast<float2*>(gfl + at) = make_float2(blk[0][k] * inv, blk[1 % nT][k] * inv);
        } else if (i < D) {
#pragma unroll
          for (int J = 0; J < nT; ++J) {
            if (j0 + J >= D) continue;
            const size_t at = (size_t)i * D + j0 + J;
            if (Gib) Gib[at] = (int32_t)gb[J];
            if (Ghb) Ghb[at] = (int16_t)gb[J];
            if (Gfb) Gfb[at] = blk[J][k] * inv;
          }
        }
      }
    }
  }
}

template <int DP, bool GIVEN = false>
static void launch_mfma(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M,
                        int K, float Q, const FrontOut& o, hipStream_t st, const int8_t* xg_in = nullptr,
                        const int32_t* list = nullptr) {
  constexpr int nT = DP / 16;
  // vector loads of a lane's nT adjacent columns: whole rows of nT-aligned width and an aligned base
  const size_t align = nT == 4 ? 16 : nT == 2 ? 8 : 4;
  const int vec = (D == DP) && ((reinterpret_cast<uintptr_t>(H) % align) == 0);
  hipLaunchKernelGGL((front_end_mfma_kernel<DP, GIVEN>), dim3(B), dim3(64), 0, st, H, y, N0v, N0s, Mr, D, M, K, Q, vec,
                     o, xg_in, list);
}

bool mdi_build_mfma(const float* H, const float* y, const int8_t* xg, const int32_t* list, int n, int Mr, int D,
                    float Q, int32_t* Gi, int32_t* hi, hipStream_t st) {
  int qe = 0;
  if (D > 64 || Mr > 64 || !(Q > 0.f) || std::frexp(Q, &qe) != 0.5f) return false;
  FrontOut o{};
  o.Gi = Gi;
  o.hi = hi;
  if (D <= 16) launch_mfma<16, true>(H, y, nullptr, 1.f, n, Mr, D, 16, 1, Q, o, st, xg, list);
  else if (D <= 32) launch_mfma<32, true>(H, y, nullptr, 1.f, n, Mr, D, 16, 1, Q, o, st, xg, list);
  else if (D <= 48) launch_mfma<48, true>(H, y, nullptr, 1.f, n, Mr, D, 16, 1, Q, o, st, xg, list);
  else launch_mfma<64, true>(H, y, nullptr, 1.f, n, Mr, D, 16, 1, Q, o, st, xg, list);
  return true;
}

hipError_t launch_front_end(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M,
                            int K, float Q, const FrontOut& o, hipStream_t st) {
  // the expanded Chord inputs, the register-sweep form (form 1) and a fixed-point scale Q that is not a power of two
  // (this kernel rounds G (Q / scale), equal to (G / scale) Q only for Q = 2^k) take the register-sweep kernels
  int qe = 0;
  const bool q_pow2 = Q > 0.f && std::frexp(Q, &qe) == 0.5f;
  if (g_front_end_form == 2 && front_end_cpx(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st)) return hipGetLastError();
  if (g_front_end_form != 1 && D <= 64 && Mr <= 64 && !o.J && !o.Jf && q_pow2) {
    if (D <= 16) launch_mfma<16>(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
    else if (D <= 32) launch_mfma<32>(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
    else if (D <= 48) launch_mfma<48>(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
    else launch_mfma<64>(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
    return hipGetLastError();
  }
  return launch_front_end_sweep(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
}

}  // namespace arp
