// FP32 annealers for MI355X: the Arpeggio (first-accept) and Chord (snapshot) anneals on float32 scaled couplings.
//
// The reference's float solvers (arpeggio/sim/arpx55.py:29-99 rule 'r', arpeggio/mimo/mimo.py:204-249) draw their
// uniforms from torch generators; these kernels draw them from the fixed-point path's Philox counters instead, so the
// acceptance draws match stream for stream: r = ((float)x * 2^-32 - 1/2) * 4T, flip iff z < r, which is exactly the
// fixed-point test x 4T > (z + 2T) 2^32 when evaluated without rounding. Every float expression is evaluated in a fixed
// order without fused multiply-add (the library is built with -ffp-contract=off), the same order as the CPU oracle's
// FP32 twins, so fields, energies and trajectories are reproducible bit for bit.
// Launch shapes and register layout follow the fixed-point kernels (anneal_arpeggio.hip, anneal_chord.hip).
#include "kernels.h"
#include "launch_bounds.h"

namespace arp {

__device__ __forceinline__ float uniform_r(uint32_t x, float T4) {
  return ((float)x * 2.3283064365386963e-10f - 0.5f) * T4;
}

template <int P, int CPL>
__global__ __launch_bounds__(ARP_ARPEGGIO_F32_MAX_THREADS) void arpeggio_f32_kernel(const float* __restrict__ Gt, const float* __restrict__ h,
                                                           const float* __restrict__ T, int L, float q, int R, int D,
                                                           const uint32_t* __restrict__ inst, uint32_t k0, uint32_t k1,
                                                           int start, int8_t* __restrict__ out_s,
                                                           float* __restrict__ out_de, long long* __restrict__ out_st,
                                                           float* __restrict__ out_w) {
  extern __shared__ float sGf[];
  const int b = blockIdx.x;
  const int nG = D * D;
  const float* g = Gt + (size_t)b * nG;
  for (int x = threadIdx.x; x < nG + ARP_WAVE * CPL; x += blockDim.x) sGf[x] = x < nG ? g[x] : 0.f;
  __syncthreads();
  const float* G = sGf;
  const int lane = lane_id(), wv = threadIdx.x >> 6, W = blockDim.x >> 6;
  const uint32_t in = inst[b];
  const float* hb = h + (size_t)b * D;
  for (int rr = wv; rr < R; rr += W) {
    const uint32_t r = (uint32_t)rr;
    const size_t t = (size_t)b * R + rr;
    bool valid[CPL];
    float w[CPL], gjj2[CPL], gjj4[CPL];
    uint32_t sb[CPL], bsb[CPL];
    uint64_t vmask[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const int j = c * ARP_WAVE + lane;
      valid[c] = j < D;
      const int nv = D - c * ARP_WAVE;
      vmask[c] = nv >= ARP_WAVE ? ~0ull : (nv <= 0 ? 0ull : ((1ull << nv) - 1ull));
      const float gd = valid[c] ? G[j * D + j] : 0.f;
      gjj2[c] = 2.f * gd;
      gjj4[c] = 4.f * gd;
      uint32_t bits = 0;
      if (valid[c]) {
        if (start == 0) {
#pragma unroll
          for (int p = 0; p < P; ++p)
            bits |= (philox_w0(0xFFFFFFFFu, (uint32_t)(p * D + j), r, in, k0, k1) >= 0x80000000u ? 1u : 0u) << p;
        } else {
          const uint32_t x = philox_w0(0xFFFFFFFFu, (uint32_t)j, r, in, k0, k1);
          bits = bal_pattern(P, (uint32_t)(((uint64_t)x * bal_count(P)) >> 32));
        }
      }
      sb[c] = bits;
      bsb[c] = bits;
    }
    // w_i = h_i - 2 acc_i, acc_i = sum_j G_ij m_j accumulated in index order j = 0..D-1
    float acc[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) acc[c] = 0.f;
#pragma unroll
    for (int c2 = 0; c2 < CPL; ++c2) {
      const int mj_lane = valid[c2] ? 2 * __builtin_popcount(sb[c2]) - P : 0;
      const int nj = D - c2 * ARP_WAVE < ARP_WAVE ? D - c2 * ARP_WAVE : ARP_WAVE;
      for (int l = 0; l < nj; ++l) {
        const float mj = (float)readlane(mj_lane, l);
        const float* col = G + (c2 * ARP_WAVE + l) * D;
#pragma unroll
        for (int c = 0; c < CPL; ++c) acc[c] = acc[c] + col[c * ARP_WAVE + lane] * mj;
      }
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) w[c] = valid[c] ? hb[c * ARP_WAVE + lane] - 2.f * acc[c] : 0.f;
    float de = 0.f, best = __builtin_inff();
    long long flips = 0, tests = 0;
    uint32_t rw[P][CPL][4];
    int n = 0;
    for (int k = 0; k < L; ++n) {
      const int nm = n & 3;
#pragma unroll
      for (int p = 0; p < P; ++p) {
        if (k >= L) break;
        if (nm == 0) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            uint32_t c0 = (uint32_t)(n >> 2), c1 = (uint32_t)(p * D + c * ARP_WAVE + lane), c2 = r, c3 = in;
            philox10(c0, c1, c2, c3, k0, k1);
            rw[p][c][0] = c0;
            rw[p][c][1] = c1;
            rw[p][c][2] = c2;
            rw[p][c][3] = c3;
          }
        }
        const float T4 = 4.f * T[k];
        float rv[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const uint32_t x = nm == 0 ? rw[p][c][0] : nm == 1 ? rw[p][c][1] : nm == 2 ? rw[p][c][2] : rw[p][c][3];
          rv[c] = uniform_r(x, T4);
        }
        long long tt = 1;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          uint64_t rem = vmask[c];
          while (rem) {
            const bool s_pos = (sb[c] >> p) & 1u;
            const float z = ((s_pos ? w[c] : -w[c]) + gjj2[c]) + q;
            const uint64_t m = ballot(z < rv[c]) & rem;
            if (m == 0) break;
            const int l = __builtin_ctzll(m);
            const int jf = c * ARP_WAVE + l;
            const float wj = __builtin_bit_cast(float, readlane(__builtin_bit_cast(int, w[c]), l));
            const float g4 = __builtin_bit_cast(float, readlane(__builtin_bit_cast(int, gjj4[c]), l));
            const bool so_pos = (readlane_u(sb[c], l) >> p) & 1u;
            de = de + ((so_pos ? 2.f : -2.f) * wj + g4);
            const float so4 = so_pos ? 4.f : -4.f;
            const float* col = G + jf * D;
#pragma unroll
            for (int cc = 0; cc < CPL; ++cc) w[cc] = w[cc] + so4 * col[cc * ARP_WAVE + lane];
            if (lane == l) sb[c] ^= 1u << p;
            ++flips;
            if (jf < D - 1) ++tt;
            rem &= (l == ARP_WAVE - 1) ? 0ull : (~0ull << (l + 1));
          }
        }
        tests += tt;
        if (de < best) {
          best = de;
#pragma unroll
          for (int c = 0; c < CPL; ++c) bsb[c] = sb[c];
        }
        ++k;
      }
    }
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const int j = c * ARP_WAVE + lane;
      if (valid[c]) {
#pragma unroll
        for (int p = 0; p < P; ++p) out_s[t * (size_t)(P * D) + p * D + j] = ((bsb[c] >> p) & 1u) ? 1 : -1;
        if (out_w) out_w[t * (size_t)D + j] = w[c];
      }
    }
    if (lane == 0) {
      out_de[t] = best;
      out_st[2 * t] = flips;
      out_st[2 * t + 1] = tests;
    }
  }
}

template <int P>
static hipError_t launch_f32_p(const float* Gt, const float* h, const float* T, int L, float q, int R, int D, int B,
                               const uint32_t* inst, uint32_t k0, uint32_t k1, int start, int8_t* out_s, float* out_de,
                               long long* out_st, float* out_w, int waves, hipStream_t st) {
  const int cpl = D <= ARP_WAVE ? 1 : 2;
  const size_t lds = sizeof(float) * ((size_t)D * D + ARP_WAVE * cpl);
  const void* fn = cpl == 1 ? (const void*)arpeggio_f32_kernel<P, 1> : (const void*)arpeggio_f32_kernel<P, 2>;
  if (lds > 65536) hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (cpl == 1)
    hipLaunchKernelGGL((arpeggio_f32_kernel<P, 1>), dim3(B), dim3(ARP_WAVE * waves), lds, st, Gt, h, T, L, q, R, D,
                       inst, k0, k1, start, out_s, out_de, out_st, out_w);
  else
    hipLaunchKernelGGL((arpeggio_f32_kernel<P, 2>), dim3(B), dim3(ARP_WAVE * waves), lds, st, Gt, h, T, L, q, R, D,
                       inst, k0, k1, start, out_s, out_de, out_st, out_w);
  return hipGetLastError();
}

hipError_t launch_arpeggio_f32(const float* Gt, const float* h, const float* T, int L, float q, int P, int R, int D,
                               int B, const uint32_t* inst, uint32_t k0, uint32_t k1, int start, int8_t* out_s,
                               float* out_de, long long* out_st, float* out_w, int waves, hipStream_t st) {
  switch (P) {
    case 2: return launch_f32_p<2>(Gt, h, T, L, q, R, D, B, inst, k0, k1, start, out_s, out_de, out_st, out_w, waves, st);
    case 4: return launch_f32_p<4>(Gt, h, T, L, q, R, D, B, inst, k0, k1, start, out_s, out_de, out_st, out_w, waves, st);
    default: return launch_f32_p<6>(Gt, h, T, L, q, R, D, B, inst, k0, k1, start, out_s, out_de, out_st, out_w, waves, st);
  }
}

// FP32 Chord: J (B, N, N) float symmetric (row j = column j), h, h0 (B, N), T, c (S) float, q float.
template <int MPL>
__global__ __launch_bounds__(ARP_CHORD_F32_MAX_THREADS) void chord_f32_kernel(const float* __restrict__ Jg, const float* __restrict__ hg,
                                                        const float* __restrict__ h0g, const float* __restrict__ Tf,
                                                        const float* __restrict__ Cf, float q, int S, int P, int R,
                                                        int N, uint32_t k0, uint32_t k1, int inst0,
                                                        float* __restrict__ out_e, int8_t* __restrict__ out_s,
                                                        float* __restrict__ out_u, long long* __restrict__ out_fl,
                                                        int lds_j) {
  extern __shared__ __align__(16) float sJ[];
  const int b = blockIdx.x;
  const size_t NN = (size_t)N * N;
  const float* Jb = Jg + (size_t)b * NN;
  if (lds_j) {
    for (size_t x = threadIdx.x; x < NN; x += blockDim.x) sJ[x] = Jb[x];
  }
  __syncthreads();
  const float* J = lds_j ? sJ : Jb;
  const int lane = lane_id(), wv = threadIdx.x >> 6, W = blockDim.x >> 6;
  const uint32_t inst = (uint32_t)(inst0 + b);
  const int per = N / P;
  float hv[MPL], h0v[MPL], jd4[MPL], hc[MPL];
  int cp[MPL];
  bool ok[MPL];
  uint64_t okm[MPL];
#pragma unroll
  for (int t = 0; t < MPL; ++t) {
    const int i = t * ARP_WAVE + lane;
    ok[t] = i < N;
    okm[t] = ballot(ok[t]);
    hv[t] = ok[t] ? hg[(size_t)b * N + i] : 0.f;
    h0v[t] = ok[t] ? h0g[(size_t)b * N + i] : 0.f;
    hc[t] = hv[t] - 2.f * h0v[t];
    jd4[t] = ok[t] ? 4.f * J[(size_t)i * N + i] : 0.f;
    cp[t] = ok[t] ? i / per : -1;
  }
  for (int r = wv; r < R; r += W) {
    uint64_t Sm[MPL];
#pragma unroll
    for (int t = 0; t < MPL; ++t) {
      const uint32_t x = philox_w0(0xFFFFFFFFu, (uint32_t)(t * ARP_WAVE + lane), (uint32_t)r, inst, k0, k1);
      Sm[t] = ballot(x >= 0x80000000u) & okm[t];
    }
    // u_i = h_i + 2 acc_i, acc_i = sum_j J_ij s_j in index order
    float acc[MPL];
#pragma unroll
    for (int t = 0; t < MPL; ++t) acc[t] = 0.f;
#pragma unroll
    for (int tj = 0; tj < MPL; ++tj) {
      const int nj = N - tj * ARP_WAVE < ARP_WAVE ? N - tj * ARP_WAVE : ARP_WAVE;
      for (int l = 0; l < nj; ++l) {
        const float sj = ((Sm[tj] >> l) & 1ull) ? 1.f : -1.f;
        const float* row = J + (size_t)(tj * ARP_WAVE + l) * N;
#pragma unroll
        for (int t = 0; t < MPL; ++t)
          if (ok[t]) acc[t] = acc[t] + row[t * ARP_WAVE + lane] * sj;
      }
    }
    float u[MPL];
#pragma unroll
    for (int t = 0; t < MPL; ++t) u[t] = hv[t] + 2.f * acc[t];
    // E0 = 1/2 sum_i s_i ((h_i - 2 h0_i) - u_i), summed in index order
    float e = 0.f;
#pragma unroll
    for (int t = 0; t < MPL; ++t) {
      const float si = ((Sm[t] >> lane) & 1ull) ? 1.f : -1.f;
      const float term = si * (hc[t] - u[t]);
      const int nl = N - t * ARP_WAVE < ARP_WAVE ? N - t * ARP_WAVE : ARP_WAVE;
      for (int l = 0; l < nl; ++l) e = e + __builtin_bit_cast(float, readlane(__builtin_bit_cast(int, term), l));
    }
    e = e * 0.5f;
    float best_e = __builtin_inff();
    uint64_t pf[MPL], best_m[MPL];
    uint32_t r0[MPL], r1[MPL], r2[MPL], r3[MPL];
#pragma unroll
    for (int t = 0; t < MPL; ++t) {
      pf[t] = 0;
      best_m[t] = Sm[t];
    }
    long long nfl = 0;
    int p = 0, n = 0;
    for (int k = 0; k < S; ++k) {
      const int nm = n & 3;
      const float T4 = 4.f * Tf[k], ck = Cf[k];
      uint64_t F[MPL];
      bool any = false;
#pragma unroll
      for (int t = 0; t < MPL; ++t) {
        const bool mine = cp[t] == p;
        F[t] = 0;
        if (ballot(mine) == 0) continue;
        if (nm == 0) {
          uint32_t c0 = (uint32_t)(n >> 2), c1 = (uint32_t)(t * ARP_WAVE + lane), c2 = (uint32_t)r, c3 = inst;
          philox10(c0, c1, c2, c3, k0, k1);
          r0[t] = c0;
          r1[t] = c1;
          r2[t] = c2;
          r3[t] = c3;
        }
        const uint32_t x = nm == 0 ? r0[t] : nm == 1 ? r1[t] : nm == 2 ? r2[t] : r3[t];
        const bool s_pos = (Sm[t] >> lane) & 1ull;
        const bool prev = (pf[t] >> lane) & 1ull;
        const float z = ((s_pos ? u[t] : -u[t]) + q) - (prev ? -ck : ck);
        F[t] = ballot(mine && z < uniform_r(x, T4));
        any |= F[t] != 0;
      }
#pragma unroll
      for (int t = 0; t < MPL; ++t) {
        pf[t] = F[t];
        uint64_t m = F[t];
        while (m) {
          const int l = __builtin_ctzll(m);
          m &= m - 1;
          const int j = t * ARP_WAVE + l;
          const bool sjp = (Sm[t] >> l) & 1ull;
          const float uj = __builtin_bit_cast(float, readlane(__builtin_bit_cast(int, u[t]), l));
          const float hj = __builtin_bit_cast(float, readlane(__builtin_bit_cast(int, hv[t]), l));
          const float h0j = __builtin_bit_cast(float, readlane(__builtin_bit_cast(int, h0v[t]), l));
          const float j4 = __builtin_bit_cast(float, readlane(__builtin_bit_cast(int, jd4[t]), l));
          e = e + ((sjp ? 2.f : -2.f) * ((uj - hj) + h0j) - j4);
          const float sj4 = sjp ? 4.f : -4.f;
          const float* row = J + (size_t)j * N;
#pragma unroll
          for (int tt = 0; tt < MPL; ++tt)
            if (ok[tt]) u[tt] = u[tt] - sj4 * row[tt * ARP_WAVE + lane];
          Sm[t] ^= 1ull << l;
          ++nfl;
        }
      }
      if (++p == P) {
        p = 0;
        ++n;
      }
      if ((any || k == 0) && e < best_e) {
        best_e = e;
#pragma unroll
        for (int t = 0; t < MPL; ++t) best_m[t] = Sm[t];
      }
    }
    const size_t o = (size_t)b * R + r;
    if (lane == 0) {
      out_e[o] = best_e;
      if (out_fl) out_fl[o] = nfl;
    }
#pragma unroll
    for (int t = 0; t < MPL; ++t)
      if (ok[t]) {
        out_s[o * N + t * ARP_WAVE + lane] = ((best_m[t] >> lane) & 1ull) ? 1 : -1;
        if (out_u) out_u[o * N + t * ARP_WAVE + lane] = u[t];
      }
  }
}

hipError_t launch_chord_f32(const float* J, const float* h, const float* h0, const float* T, const float* c, float q,
                            int S, int P, int R, int N, int B, uint32_t k0, uint32_t k1, int inst0, float* out_e,
                            int8_t* out_s, float* out_u, long long* out_fl, int waves, size_t max_lds,
                            hipStream_t st) {
  const size_t jb = sizeof(float) * (size_t)N * N;
  const int lds_j = jb <= max_lds ? 1 : 0;
  const size_t lds = lds_j ? jb : 0;
  const int mpl = (N + ARP_WAVE - 1) / ARP_WAVE;
  const void* fns[4] = {(const void*)chord_f32_kernel<1>, (const void*)chord_f32_kernel<2>,
                        (const void*)chord_f32_kernel<3>, (const void*)chord_f32_kernel<4>};
  if (lds > 65536) hipFuncSetAttribute(fns[mpl - 1], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  dim3 grid(B), block(ARP_WAVE * waves);
  switch (mpl) {
    case 1: hipLaunchKernelGGL(chord_f32_kernel<1>, grid, block, lds, st, J, h, h0, T, c, q, S, P, R, N, k0, k1, inst0, out_e, out_s, out_u, out_fl, lds_j); break;
    case 2: hipLaunchKernelGGL(chord_f32_kernel<2>, grid, block, lds, st, J, h, h0, T, c, q, S, P, R, N, k0, k1, inst0, out_e, out_s, out_u, out_fl, lds_j); break;
    case 3: hipLaunchKernelGGL(chord_f32_kernel<3>, grid, block, lds, st, J, h, h0, T, c, q, S, P, R, N, k0, k1, inst0, out_e, out_s, out_u, out_fl, lds_j); break;
    default: hipLaunchKernelGGL(chord_f32_kernel<4>, grid, block, lds, st, J, h, h0, T, c, q, S, P, R, N, k0, k1, inst0, out_e, out_s, out_u, out_fl, lds_j); break;
  }
  return hipGetLastError();
}

}  // namespace arp
