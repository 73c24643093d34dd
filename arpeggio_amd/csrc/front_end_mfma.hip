// Front end for 16 and 32 users (D, Mr <= 64) with the whole MMSE inverse on the matrix cores.
//
// The same outputs as the register-sweep front end (front_end.hip) from a different schedule. One wave per frame:
//   1. G = H^T H in v_mfma_f32_16x16x4f32 accumulators (the upper 16x16 tiles of the coordinates in tile order,
//      position 16 I + m = coordinate nT m + I, so that a lane's operand columns are adjacent in memory), H read once
//      from global memory straight into the operands; c = H^T y by the VALU beside it. A copy of the tiles waits in
//      LDS (10 KiB at D = 64, in the accumulator layout) for the outputs.
//   2. A = G + sigma I (identity on the padding) is swept in place in the accumulators, four pivots at a time. For the
//      pivot set S = {k0 .. k0+3} the block sweep is
//          A_SS -> -A_SS^-1,  A_RS -> A_RS A_SS^-1,  A_SR -> A_SS^-1 A_SR,  A_RR -> A_RR - A_RS A_SS^-1 A_SR,
//      the composition of the four single-pivot sweeps. With V = A_.S - E_.S (E: ones at (k, k), k in S), all four
//      parts are one symmetric rank-4 update A <- A - (V A_SS^-1) V^T on every tile, except that the diagonal of A_SS
//      comes out 2 too high; no later block reads those entries, so the 2 comes off once, at the end. The update is
//      one MFMA per upper tile (A operand V A_SS^-1, B operand V^T); the panel (rows S, every column) goes through a
//      1 KiB LDS buffer, and the 4x4 inverse of A_SS is swept in registers by every lane. The right-hand side c rides
//      along as an extra column (c <- c - V A_SS^-1 c_S), so it ends as z = A^-1 c.
//   3. mu_i = 1 - sigma (A^-1)_ii, x_g = nearest level of z / mu (ties to the lower level).
//   4. h = 2 H^T (y - H x_g) (mimo.ising_delta_at) as 2 (c - G x_g) from the Gram tiles, with cross-lane sums;
//      scale = max(max|h|, max|G|), clamped at 1e-12; G_int = rne(G / scale * Q), h_int = rne(h / scale * Q),
//      G written row-major in runs of four entries gathered from the LDS copy.
// Reference semantics: mimo.mmse / _mu / quantize_levels (arpeggio/mimo/mimo.py:86-109, 176-182), mimo.ising_delta_at
// (mimo.py:506-523), chordfx.to_fixed (arpeggio/mimo/chordfx.py:60-76), chord_ref.compress
// (arpeggio/sim/x3model/chord_ref.py:47-59). The expanded Chord inputs (J, Jf) are the register-sweep kernel's.
#include "kernels.h"

namespace arp {

typedef float floatx4 __attribute__((ext_vector_type(4)));

static int g_front_end_form = 0;
void set_front_end_form(int form) { g_front_end_form = form; }

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

// Reduce-scatter over the 16 lanes of a group (lane bits 0..3): v holds NP partial sums per lane (NP a power of two,
// 4..16); afterwards v[0] is the sum over the 16 lanes of entry k(col), where step m (8, 4, 2, 1) keeps the upper half
// on lanes with bit m set while more than one entry is left.
template <int NP>
__device__ __forceinline__ float reduce_scatter16(float (&v)[NP], int col) {
  int n = NP;
#pragma unroll
  for (int m = 8; m >= 1; m >>= 1) {
    if (n > 1) {
      const bool up = (col & m) != 0;
      const int half = n / 2;
#pragma unroll
      for (int i = 0; i < NP / 2; ++i)
        if (i < half) {
          const float send = up ? v[i] : v[i + half];
          const float keep = up ? v[i + half] : v[i];
          v[i] = keep + __shfl_xor(send, m);
        }
      n = half;
    } else {
      v[0] += __shfl_xor(v[0], m);
    }
  }
  return v[0];
}

template <int DP>
__global__ __launch_bounds__(64) void front_end_mfma_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                            const float* __restrict__ N0v, float N0s, int Mr, int D,
                                                            int M, int K, float Q, int vec, FrontOut o) {
  constexpr int nT = DP / 16, nU = nT * (nT + 1) / 2;
  constexpr int NP = 4 * nT == 12 ? 16 : 4 * nT;        // row partial sums per lane in the G x reduction
  __shared__ __align__(16) floatx4 Gs[nU * 64];         // the Gram tiles as each lane holds them
  __shared__ __align__(16) float pan[2][DP * 4];        // panel rows S: pan[p * 4 + t] = A[k0 + t][p]
  __shared__ __align__(16) float bc[64];                // x_g in tile order
  const int b = blockIdx.x, lane = threadIdx.x, col = lane & 15, grp = lane >> 4;
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
  // G waits in LDS, in the accumulator layout, until the outputs
#pragma unroll
  for (int u = 0; u < nU; ++u) Gs[u * 64 + lane] = T[u];
  const float el2 = M == 4 ? 1.f : 5.f;
  const float n0 = N0v ? N0v[b] : N0s;
  const float sigma = (n0 / 2.f) / el2;
  // A = G + sigma I; identity on the padding. Diagonal tile (I, I) holds positions (16 I + 4 grp + v, 16 I + col).
#pragma unroll
  for (int I = 0; I < nT; ++I)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (4 * grp + v == col) T[tile<nT>(I, I)][v] = nT * col + I < D ? T[tile<nT>(I, I)][v] + sigma : 1.f;

  // 2. Block sweep over the pivot sets {k0 .. k0 + 3} of the tile order (padding pivots included: identity rows
  // sweep to themselves); the panel buffer alternates between two halves, so a block's stores never meet the
  // previous block's loads.
#pragma unroll
  for (int blk = 0; blk < DP / 4; ++blk) {
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
    // m = A_SS swept in place on its upper triangle (m[i][j], i <= j): afterwards m = -A_SS^-1
    float m[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = i; j < 4; ++j) m[i][j] = S4[j][i];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const float piv = __builtin_amdgcn_rcpf(m[p][p]);
      float rr[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j != p) rr[j] = (p < j ? m[p][j] : m[j][p]) * piv;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j)
          if (i != p && j != p) m[i][j] = __builtin_fmaf(-(i < p ? m[i][p] : m[p][i]), rr[j], m[i][j]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j != p) (p < j ? m[p][j] : m[j][p]) = rr[j];
      m[p][p] = -piv;
    }
    // V = A_.S - E: minus one where the row is the pivot itself
#pragma unroll
    for (int t = 0; t < 4; ++t) Vr[J0][t] -= col == 4 * g0 + t ? 1.f : 0.f;
    vt[J0] -= col == 4 * g0 + grp ? 1.f : 0.f;
    // mg = column grp of m (= -A_SS^-1), qv = m c_S (= -A_SS^-1 c_S)
    float mg[4], qv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float m0 = t < 1 ? m[t][0] : m[0][t], m1 = t < 2 ? m[t][1] : m[1][t];
      const float m2 = t < 3 ? m[t][2] : m[2][t], m3 = m[t][3];
      mg[t] = pick4(floatx4{m0, m1, m2, m3}, grp);
      qv[t] = __builtin_fmaf(m3, cS[3], __builtin_fmaf(m2, cS[2], __builtin_fmaf(m1, cS[1], m0 * cS[0])));
    }
    float nu[nT];  // -(V A_SS^-1)[16 I + col][grp]: the A operand of the update
#pragma unroll
    for (int I = 0; I < nT; ++I) {
      nu[I] = __builtin_fmaf(Vr[I][3], mg[3], __builtin_fmaf(Vr[I][2], mg[2],
                             __builtin_fmaf(Vr[I][1], mg[1], Vr[I][0] * mg[0])));
      c[I] = __builtin_fmaf(Vr[I][3], qv[3], __builtin_fmaf(Vr[I][2], qv[2],
                            __builtin_fmaf(Vr[I][1], qv[1], __builtin_fmaf(Vr[I][0], qv[0], c[I]))));
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
  const float gii = valid ? Gs[(Ii * nT - Ii * (Ii - 1) / 2) * 64 + src][mi & 3] : 0.f;
  const float mu = 1.f - sigma * dgi;
  const float vq = z / mu;
  const float lmax = M == 4 ? 1.f : 3.f;
  float l = fminf(fmaxf(2.f * ceilf(vq * 0.5f) - 1.f, -lmax), lmax);
  if (vq != vq) l = -lmax;
  if (!valid) l = 0.f;
  if (valid) {
    if (o.z) o.z[(size_t)b * D + lane] = z;
    if (o.mu) o.mu[(size_t)b * D + lane] = mu;
    if (o.xg) o.xg[(size_t)b * D + lane] = (int8_t)l;
  }
  if (o.feat) {
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

  // 4. h = 2 (c - G x_g) in tile order. Tile (I, J) entry (16 I + 4 grp + v, 16 J + col) adds G x at 16 J + col to
  // row 16 I + 4 grp + v (summed over the 16 lanes of the group by the reduce-scatter) and, for I < J, G x at
  // 16 I + 4 grp + v to row 16 J + col (summed over the four groups).
  if (lane < DP) bc[16 * Ii + mi] = l;
  __syncthreads();
  float ra[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) ra[k] = 0.f;
  float rb[nT];
#pragma unroll
  for (int J = 0; J < nT; ++J) rb[J] = 0.f;
  float gm = 0.f;
#pragma unroll
  for (int I = 0; I < nT; ++I) {
    const floatx4 xi = *reinterpret_cast<const floatx4*>(bc + 16 * I + 4 * grp);
#pragma unroll
    for (int J = I; J < nT; ++J) {
      const floatx4 g = Gs[tile<nT>(I, J) * 64 + lane];
      const float xj = bc[16 * J + col];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        ra[4 * I + v] = __builtin_fmaf(g[v], xj, ra[4 * I + v]);
        gm = fmaxf(gm, fabsf(g[v]));  // padding entries are zero
      }
      if (I < J)
        rb[J] = __builtin_fmaf(g[3], xi[3], __builtin_fmaf(g[2], xi[2], __builtin_fmaf(g[1], xi[1],
                                                                                       __builtin_fmaf(g[0], xi[0], rb[J]))));
    }
  }
#pragma unroll
  for (int J = 0; J < nT; ++J) {
    rb[J] += __shfl_xor(rb[J], 16);
    rb[J] += __shfl_xor(rb[J], 32);
  }
  reduce_scatter16<NP>(ra, col);
  // position 16 Ii + mi: part a in group mi >> 2 at the lane whose scatter index is 4 Ii + (mi & 3); part b in
  // register Ii of lane mi of any group
  const int ka = 4 * Ii + (mi & 3);
  const int ca = NP == 16 ? ka : NP == 8 ? 2 * ka : 4 * ka;
  const float gxa = __shfl(ra[0], 16 * ((mi & 15) >> 2) + (ca & 15));
  float gxb = 0.f;
#pragma unroll
  for (int J = 0; J < nT; ++J) {
    const float t = __shfl(rb[J], mi & 15);
    if (Ii == J) gxb = t;
  }
  const float hval = valid ? 2.f * (cc0 - (gxa + gxb)) : 0.f;
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
  // 5. G / scale in row-major order, gathered from the tiles in LDS: G_ij sits at tile (i % nT, j % nT), entry
  // (i / nT, j / nT), or at the mirror entry of tile (j % nT, i % nT). Each lane writes runs of four adjacent entries,
  // so a wave stores whole rows.
  if (o.Gi || o.Gf || o.Gh) {
    const float* Gf32 = reinterpret_cast<const float*>(Gs);
    int32_t* Gib = o.Gi ? o.Gi + (size_t)b * D * D : nullptr;
    int16_t* Ghb = o.Gh ? o.Gh + (size_t)b * D * D : nullptr;
    float* Gfb = o.Gf ? o.Gf + (size_t)b * D * D : nullptr;
    auto at = [&](int i, int j) -> float {
      const int ii = i % nT, mi2 = i / nT, jj = j % nT, mj2 = j / nT;
      const bool up = ii <= jj;
      const int I = up ? ii : jj, J = up ? jj : ii, mr = up ? mi2 : mj2, mc = up ? mj2 : mi2;
      return Gf32[((I * nT - I * (I - 1) / 2 + (J - I)) * 64 + mc + 16 * (mr >> 2)) * 4 + (mr & 3)];
    };
    if ((D & 3) == 0) {
      const int runs = D * D / 4;
      for (int q = lane; q < runs; q += 64) {
        const int i = q / (D / 4), j0 = 4 * (q % (D / 4));
        floatx4 gs;
#pragma unroll
        for (int e = 0; e < 4; ++e) gs[e] = at(i, j0 + e) * inv;
        if (Gib || Ghb) {
          int32_t g4[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) g4[e] = (int32_t)rintf(gs[e] * Q);
          if (Gib) *reinterpret_cast<int4*>(Gib + 4 * q) = make_int4(g4[0], g4[1], g4[2], g4[3]);
          if (Ghb) {
            const uint32_t lo = (uint32_t)(uint16_t)g4[0] | ((uint32_t)(uint16_t)g4[1] << 16);
            const uint32_t hi = (uint32_t)(uint16_t)g4[2] | ((uint32_t)(uint16_t)g4[3] << 16);
            *reinterpret_cast<uint2*>(Ghb + 4 * q) = make_uint2(lo, hi);
          }
        }
        if (Gfb) *reinterpret_cast<floatx4*>(Gfb + 4 * q) = gs;
      }
    } else {
      for (int x = lane; x < D * D; x += 64) {
        const float gs = at(x / D, x % D) * inv;
        if (Gib) Gib[x] = (int32_t)rintf(gs * Q);
        if (Ghb) Ghb[x] = (int16_t)rintf(gs * Q);
        if (Gfb) Gfb[x] = gs;
      }
    }
  }
}

template <int DP>
static void launch_mfma(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M,
                        int K, float Q, const FrontOut& o, hipStream_t st) {
  constexpr int nT = DP / 16;
  // vector loads of a lane's nT adjacent columns: whole rows of nT-aligned width and an aligned base
  const size_t align = nT == 4 ? 16 : nT == 2 ? 8 : 4;
  const int vec = (D == DP) && ((reinterpret_cast<uintptr_t>(H) % align) == 0);
  hipLaunchKernelGGL((front_end_mfma_kernel<DP>), dim3(B), dim3(64), 0, st, H, y, N0v, N0s, Mr, D, M, K, Q, vec, o);
}

hipError_t launch_front_end(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M,
                            int K, float Q, const FrontOut& o, hipStream_t st) {
  // the expanded Chord inputs and the register-sweep form (form 1) take the register-sweep kernels
  if (g_front_end_form == 0 && D <= 64 && Mr <= 64 && !o.J && !o.Jf) {
    if (D <= 16) launch_mfma<16>(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
    else if (D <= 32) launch_mfma<32>(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
    else if (D <= 48) launch_mfma<48>(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
    else launch_mfma<64>(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
    return hipGetLastError();
  }
  return launch_front_end_sweep(H, y, N0v, N0s, B, Mr, D, M, K, Q, o, st);
}

}  // namespace arp
