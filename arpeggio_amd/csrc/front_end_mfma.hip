// Front end for 16 and 32 users (D, Mr <= 64) with the whole MMSE inverse on the matrix cores.
//
// The same outputs as the register-sweep front end (front_end.hip) from a different schedule. One wave per frame:
//   1. G = H^T H in v_mfma_f32_16x16x4f32 accumulators (the upper 16x16 tiles), H staged in LDS (one coalesced pass
//      over HBM; the epilogue reads it again from there); c = H^T y by the VALU beside it. G stays in a second set of
//      registers for the outputs.
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
//   4. ybar = y - H x_g and h = 2 H^T ybar from the LDS copy of H, as mimo.ising_delta_at forms them;
//      scale = max(max|h|, max|G|), clamped at 1e-12; G_int = rne(G / scale * Q), h_int = rne(h / scale * Q).
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

template <int DP>
__global__ __launch_bounds__(64) void front_end_mfma_kernel(const float* __restrict__ H, const float* __restrict__ y,
                                                            const float* __restrict__ N0v, float N0s, int Mr, int D,
                                                            int M, int K, float Q, int vec4, FrontOut o) {
  constexpr int nT = DP / 16, nU = nT * (nT + 1) / 2, ld = DP + 1;
  extern __shared__ __align__(16) float sm[];
  const int Mp = (Mr + 3) & ~3;
  float* Hs = sm;                 // Mp x ld, rows padded against bank conflicts
  float* ys = Hs + Mp * ld;       // Mp
  float* bc = ys + Mp;            // 2 x 64: broadcast vectors of the epilogue
  float* pan = bc + 128;          // 2 x (DP x 4): panel rows S, pan[j * 4 + t] = A[k0 + t][j]
  const int b = blockIdx.x, lane = threadIdx.x, col = lane & 15, grp = lane >> 4;
  const float* Hb = H + (size_t)b * Mr * D;
  const float* yb = y + (size_t)b * Mr;
  // H (zero-padded to Mp x DP) into LDS, with 16-byte loads all in flight before the first store where rows allow
  if (vec4) {
    constexpr int CPR = DP / 4, RPP = 64 / CPR, NPASS = (64 + RPP - 1) / RPP;
    const int cr = lane % CPR, rr = lane / CPR;
    floatx4 v[NPASS];
#pragma unroll
    for (int t = 0; t < NPASS; ++t) {
      const int r = t * RPP + rr;
      v[t] = floatx4{0.f, 0.f, 0.f, 0.f};
      if (rr < RPP && r < Mr && 4 * cr < D) v[t] = *reinterpret_cast<const floatx4*>(Hb + (size_t)r * D + 4 * cr);
    }
#pragma unroll
    for (int t = 0; t < NPASS; ++t) {
      const int r = t * RPP + rr;
      if (rr < RPP && r < Mp)
#pragma unroll
        for (int e = 0; e < 4; ++e) Hs[r * ld + 4 * cr + e] = v[t][e];
    }
  } else {
    for (int x = lane; x < Mp * DP; x += 64) {
      const int r = x / DP, cc = x - r * DP;
      Hs[r * ld + cc] = (r < Mr && cc < D) ? Hb[(size_t)r * D + cc] : 0.f;
    }
  }
  if (lane < Mp) ys[lane] = lane < Mr ? yb[lane] : 0.f;
  __syncthreads();

  // 1. Gram tiles and c = H^T y. At k-step kk, lane (col, grp) supplies H[kk + grp][16 I + col] to tile row I (A
  // operand, rows of H^T) and to tile column I (B operand).
  floatx4 T[nU];
  float cp[nT];
#pragma unroll
  for (int u = 0; u < nU; ++u) T[u] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int I = 0; I < nT; ++I) cp[I] = 0.f;
#pragma unroll 2
  for (int kk = 0; kk < Mp; kk += 4) {
    const int r = kk + grp;
    float op[nT];
#pragma unroll
    for (int I = 0; I < nT; ++I) op[I] = Hs[r * ld + 16 * I + col];
    const float yr = ys[r];
#pragma unroll
    for (int I = 0; I < nT; ++I) cp[I] = __builtin_fmaf(op[I], yr, cp[I]);
#pragma unroll
    for (int I = 0; I < nT; ++I)
#pragma unroll
      for (int J = I; J < nT; ++J)
        T[tile<nT>(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(op[I], op[J], T[tile<nT>(I, J)], 0, 0, 0);
  }
  // c[I] = c_{16 I + col}, the same in all four groups
  float c[nT], c0[nT];
#pragma unroll
  for (int I = 0; I < nT; ++I) {
    float v = cp[I];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    c[I] = v;
    c0[I] = v;
  }
  floatx4 Gk[nU];
#pragma unroll
  for (int u = 0; u < nU; ++u) Gk[u] = T[u];
  const float el2 = M == 4 ? 1.f : 5.f;
  const float n0 = N0v ? N0v[b] : N0s;
  const float sigma = (n0 / 2.f) / el2;
  // A = G + sigma I; identity on the padding. Diagonal tile (I, I) holds (16 I + 4 grp + v, 16 I + col).
#pragma unroll
  for (int I = 0; I < nT; ++I)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (4 * grp + v == col) {
        const int i = 16 * I + col;
        T[tile<nT>(I, I)][v] = i < D ? T[tile<nT>(I, I)][v] + sigma : 1.f;
      }

  // 2. Block sweep over the pivot sets {k0 .. k0 + 3}; the panel buffer alternates between two halves, so a block's
  // stores never meet the previous block's loads.
#pragma unroll
  for (int blk = 0; blk < DP / 4; ++blk) {
    const int k0 = 4 * blk, J0 = k0 / 16, g0 = (k0 % 16) / 4;
    if (k0 >= D) break;  // padding pivots: identity, nothing to sweep
    float* P = pan + (blk & 1) * DP * 4;
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
    // m = A_SS swept in place: afterwards m = -A_SS^-1
    float m[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) m[i][j] = S4[i][j];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const float piv = __builtin_amdgcn_rcpf(m[p][p]);
      float rr[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) rr[j] = m[p][j] * piv;
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (i != p && j != p) m[i][j] = __builtin_fmaf(-m[i][p], rr[j], m[i][j]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j != p) {
          m[p][j] = rr[j];
          m[j][p] = rr[j];
        }
      m[p][p] = -piv;
    }
    // V = A_.S - E: minus one where the row is the pivot itself
#pragma unroll
    for (int t = 0; t < 4; ++t) Vr[J0][t] -= col == 4 * g0 + t ? 1.f : 0.f;
    vt[J0] -= col == 4 * g0 + grp ? 1.f : 0.f;
    // column grp of A_SS^-1 = -m, and w = A_SS^-1 c_S
    float mc[4], w[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      mc[t] = -pick4(floatx4{m[t][0], m[t][1], m[t][2], m[t][3]}, grp);
      w[t] = -(m[t][0] * cS[0] + m[t][1] * cS[1] + m[t][2] * cS[2] + m[t][3] * cS[3]);
    }
    float nu[nT];  // -(V A_SS^-1)[16 I + col][grp]: the A operand of the update
#pragma unroll
    for (int I = 0; I < nT; ++I) {
      nu[I] = -(Vr[I][0] * mc[0] + Vr[I][1] * mc[1] + Vr[I][2] * mc[2] + Vr[I][3] * mc[3]);
      c[I] -= Vr[I][0] * w[0] + Vr[I][1] * w[1] + Vr[I][2] * w[2] + Vr[I][3] * w[3];
    }
#pragma unroll
    for (int I = 0; I < nT; ++I)
#pragma unroll
      for (int J = I; J < nT; ++J)
        T[tile<nT>(I, J)] = __builtin_amdgcn_mfma_f32_16x16x4f32(nu[I], vt[J], T[tile<nT>(I, J)], 0, 0, 0);
  }

  // 3. Per coordinate i = lane: z_i = c[grp] of this lane; (A^-1)_ii = 2 - T_ii, held by lane col + 16 (col >> 2),
  // register col & 3 of the diagonal tile.
  const bool valid = lane < D;
  const int src = col + 16 * (col >> 2);
  float z = 0.f, dgi = 0.f, gii = 0.f, cc0 = 0.f;
#pragma unroll
  for (int I = 0; I < nT; ++I) {
    const float td = __shfl(pick4(T[tile<nT>(I, I)], col & 3), src);
    const float gd = __shfl(pick4(Gk[tile<nT>(I, I)], col & 3), src);
    if (grp == I) {
      z = c[I];
      dgi = 2.f - td;
      gii = gd;
      cc0 = c0[I];
    }
  }
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
    float gs = valid ? gii : 0.f;
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

  // 4. ybar_r = y_r - (H x_g)_r with lane r = row; h_j = 2 (H^T ybar)_j with lane j = column; H from LDS (row
  // stride ld: both the row reads and the column reads fall in distinct banks)
  bc[lane] = l;
  __syncthreads();
  float yr = 0.f;
  if (lane < Mp) {
    const float* Hr = Hs + lane * ld;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < DP; j += 4) {
      const floatx4 xv = *reinterpret_cast<const floatx4*>(bc + j);
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] = __builtin_fmaf(Hr[j + e], xv[e], acc[e]);
    }
    yr = ys[lane] - ((acc[0] + acc[1]) + (acc[2] + acc[3]));
  }
  bc[64 + lane] = yr;
  __syncthreads();
  float hsum = 0.f;
  {
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    const int jl = lane < DP ? lane : 0;
#pragma unroll 4
    for (int r = 0; r < Mp; r += 4) {
      const floatx4 yv = *reinterpret_cast<const floatx4*>(bc + 64 + r);
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] = __builtin_fmaf(Hs[(r + e) * ld + jl], yv[e], acc[e]);
    }
    hsum = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  }
  const float hval = valid ? 2.f * hsum : 0.f;
  float gm = 0.f;
#pragma unroll
  for (int I = 0; I < nT; ++I)
#pragma unroll
    for (int J = I; J < nT; ++J)
#pragma unroll
      for (int v = 0; v < 4; ++v) gm = fmaxf(gm, fabsf(Gk[tile<nT>(I, J)][v]));  // padding entries are zero
  float mx = fmaxf(gm, fabsf(hval));
  for (int off = 32; off; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  const float scale = fmaxf(mx, 1e-12f);
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
  }
  // G / scale from the kept Gram tiles: tile (I, J) entry (16 I + 4 grp + v, 16 J + col) and, for I < J, its mirror
  const size_t gb = (size_t)b * D * D;
#pragma unroll
  for (int I = 0; I < nT; ++I)
#pragma unroll
    for (int J = I; J < nT; ++J) {
      const int cj = 16 * J + col;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ri = 16 * I + 4 * grp + v;
        const float gs = Gk[tile<nT>(I, J)][v] / scale;
        if (ri < D && cj < D) {
          if (o.Gi) o.Gi[gb + (size_t)ri * D + cj] = (int32_t)rintf(gs * Q);
          if (o.Gf) o.Gf[gb + (size_t)ri * D + cj] = gs;
          if (I < J) {
            if (o.Gi) o.Gi[gb + (size_t)cj * D + ri] = (int32_t)rintf(gs * Q);
            if (o.Gf) o.Gf[gb + (size_t)cj * D + ri] = gs;
          }
        }
      }
    }
}

template <int DP>
static void launch_mfma(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M,
                        int K, float Q, const FrontOut& o, hipStream_t st) {
  const int vec4 = (D % 4 == 0) && ((reinterpret_cast<uintptr_t>(H) & 15) == 0);
  const int Mp = (Mr + 3) & ~3;
  const size_t lds = sizeof(float) * ((size_t)Mp * (DP + 1) + Mp + 128 + 8 * DP);
  hipLaunchKernelGGL((front_end_mfma_kernel<DP>), dim3(B), dim3(64), lds, st, H, y, N0v, N0s, Mr, D, M, K, Q, vec4, o);
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
