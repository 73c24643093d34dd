// Fixed-point Arpeggio (first-accept) annealer for MI355X (gfx950, wave64).
//
// Semantics: arpeggio/sim/local_code_a59/arpeggio_ref59.py:41-85 (mode 'arpeggio', plus mode 'par'), which the
// reference's one-thread-per-anneal CUDA kernel arpeggio/sim/arpeggio_fx.cu:35-100 also implements.
//
// Layout: one workgroup per frame (sub-problem). The frame's transposed Gram Gt (D x D, Gt[j*D + i] = G_ij) is
// staged once into LDS and shared by the workgroup's waves; wave w runs anneals r = w, w + W, ... One wave owns one
// anneal: lane l holds coordinates j = c*64 + l (c < CPL, D <= 64*CPL), so the D lanes of a pass are tested in
// parallel against their fields and the sequential index-order scan becomes
//     mask = ballot(accepted & j >= ptr);  j* = ffs(mask);  commit j*;  ptr = j* + 1;  re-test
// The commit reads one Gram column (D consecutive LDS words: bank-conflict free) and updates every lane's field.
// de, best, flips and tests are wave-uniform and live in scalar registers; spins are one bit per copy per lane.
// Each (lane, copy) Philox block gives the words of 4 consecutive sweeps, so a block is computed every 4th sweep of
// a copy instead of once per test.
#include "kernels.h"

namespace arp {

template <int P, int CPL>
__device__ __forceinline__ void arpeggio_anneal(const int32_t* __restrict__ G, const int32_t* __restrict__ hb,
                                                const int32_t* __restrict__ T, int L, int q, int D, uint32_t r,
                                                uint32_t inst, uint32_t k0, uint32_t k1, int start, int mode,
                                                int8_t* __restrict__ out_s, long long* __restrict__ out_de,
                                                long long* __restrict__ out_st, size_t t) {
  const int lane = lane_id();
  bool valid[CPL];
  int32_t w[CPL], dj2[CPL];
  uint32_t sb[CPL], bsb[CPL];
  uint64_t vmask[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const int j = c * ARP_WAVE + lane;
    valid[c] = j < D;
    const int nv = D - c * ARP_WAVE;
    vmask[c] = nv >= ARP_WAVE ? ~0ull : (nv <= 0 ? 0ull : ((1ull << nv) - 1ull));
    dj2[c] = valid[c] ? 2 * G[j * D + j] : 0;
    uint32_t bits = 0;
    if (valid[c]) {
      if (start == 0) {
#pragma unroll
        for (int p = 0; p < P; ++p)
          bits |= (philox_w0(0xFFFFFFFFu, (uint32_t)(p * D + j), r, inst, k0, k1) >= 0x80000000u ? 1u : 0u) << p;
      } else {
        const uint32_t x = philox_w0(0xFFFFFFFFu, (uint32_t)j, r, inst, k0, k1);
        bits = bal_pattern(P, (uint32_t)(((uint64_t)x * bal_count(P)) >> 32));
      }
    }
    sb[c] = bits;
    bsb[c] = bits;
    w[c] = valid[c] ? hb[j] : 0;
  }
  // w_i = h_i - 2 sum_j G_ij m_j, m_j = sum_p s_pj = 2 popcount - P
#pragma unroll
  for (int c2 = 0; c2 < CPL; ++c2) {
    const int mj_lane = valid[c2] ? 2 * __builtin_popcount(sb[c2]) - P : 0;
    const int nj = D - c2 * ARP_WAVE < ARP_WAVE ? D - c2 * ARP_WAVE : ARP_WAVE;
    for (int l = 0; l < nj; ++l) {
      const int mj = readlane(mj_lane, l);
      if (mj != 0) {
        const int32_t* col = G + (c2 * ARP_WAVE + l) * D;
#pragma unroll
        for (int c = 0; c < CPL; ++c) w[c] -= 2 * mj * col[c * ARP_WAVE + lane];
      }
    }
  }
  long long de = 0, best = 1LL << 62, flips = 0, tests = 0;
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
          uint32_t c0 = (uint32_t)(n >> 2), c1 = (uint32_t)(p * D + c * ARP_WAVE + lane), c2 = r, c3 = inst;
          philox10(c0, c1, c2, c3, k0, k1);
          rw[p][c][0] = c0;
          rw[p][c][1] = c1;
          rw[p][c][2] = c2;
          rw[p][c][3] = c3;
        }
      }
      const int32_t Tk = T[k];
      const uint32_t A = 4u * (uint32_t)Tk;
      const int32_t Bk = q + 2 * Tk;
      // flip iff s w + 2 G_jj + B_k < thr(x, 4 T_k)
      int32_t thr[CPL], zc[CPL];
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        const uint32_t x = nm == 0 ? rw[p][c][0] : nm == 1 ? rw[p][c][1] : nm == 2 ? rw[p][c][2] : rw[p][c][3];
        thr[c] = accept_threshold(x, A);
        zc[c] = dj2[c] + Bk;
      }
      long long tt = 1;
      if (mode == 0) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          uint64_t rem = vmask[c];
          while (rem) {
            const bool s_pos = (sb[c] >> p) & 1u;
            const int32_t z = (s_pos ? w[c] : -w[c]) + zc[c];
            const uint64_t m = ballot(z < thr[c]) & rem;
            if (m == 0) break;
            const int l = __builtin_ctzll(m);
            const int jf = c * ARP_WAVE + l;
            const int32_t wj = readlane(w[c], l);
            const int32_t djj = readlane(dj2[c], l);
            const uint32_t so_pos = (readlane_u(sb[c], l) >> p) & 1u;
            const int32_t so4 = so_pos ? 4 : -4;
            de += (long long)so4 * wj + 4LL * djj;
            const int32_t* col = G + jf * D;
#pragma unroll
            for (int cc = 0; cc < CPL; ++cc) w[cc] += so4 * col[cc * ARP_WAVE + lane];
            if (lane == l) sb[c] ^= 1u << p;
            ++flips;
            if (jf < D - 1) ++tt;
            rem &= (l == ARP_WAVE - 1) ? 0ull : (~0ull << (l + 1));
          }
        }
      } else {
        // 'par': decide every lane of the pass on one snapshot, then apply the flips in index order
        uint64_t fm[CPL];
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const bool s_pos = (sb[c] >> p) & 1u;
          const int32_t z = (s_pos ? w[c] : -w[c]) + zc[c];
          fm[c] = ballot(z < thr[c]) & vmask[c];
        }
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          uint64_t m = fm[c];
          while (m) {
            const int l = __builtin_ctzll(m);
            m &= m - 1;
            const int jf = c * ARP_WAVE + l;
            const int32_t wj = readlane(w[c], l);
            const int32_t djj = readlane(dj2[c], l);
            const uint32_t so_pos = (readlane_u(sb[c], l) >> p) & 1u;
            const int32_t so4 = so_pos ? 4 : -4;
            de += (long long)so4 * wj + 4LL * djj;
            const int32_t* col = G + jf * D;
#pragma unroll
            for (int cc = 0; cc < CPL; ++cc) w[cc] += so4 * col[cc * ARP_WAVE + lane];
            if (lane == l) sb[c] ^= 1u << p;
            ++flips;
          }
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
    }
  }
  if (lane == 0) {
    out_de[t] = best;
    out_st[2 * t] = flips;
    out_st[2 * t + 1] = tests;
  }
}

// grid.x = B frames, block = 64 W threads. LDS: D*D int32 (+ one wave of padding so that the lanes past D of the
// last column stay inside the allocation).
template <int P, int CPL>
__global__ __launch_bounds__(256) void arpeggio_fx_kernel(const int32_t* __restrict__ Gt, const int32_t* __restrict__ h,
                                                          const int32_t* __restrict__ T, int L, int q, int R, int D,
                                                          const uint32_t* __restrict__ inst, uint32_t k0, uint32_t k1,
                                                          int start, int mode, int8_t* __restrict__ out_s,
                                                          long long* __restrict__ out_de,
                                                          long long* __restrict__ out_st) {
  extern __shared__ int32_t sG[];
  const int b = blockIdx.x;
  const int nG = D * D;
  const int32_t* g = Gt + (size_t)b * nG;
  for (int x = threadIdx.x; x < nG + ARP_WAVE * CPL; x += blockDim.x) sG[x] = x < nG ? g[x] : 0;
  __syncthreads();
  const int wv = threadIdx.x >> 6, W = blockDim.x >> 6;
  const uint32_t in = inst[b];
  for (int r = wv; r < R; r += W)
    arpeggio_anneal<P, CPL>(sG, h + (size_t)b * D, T, L, q, D, (uint32_t)r, in, k0, k1, start, mode, out_s, out_de,
                            out_st, (size_t)b * R + r);
}

template <int P>
static hipError_t launch_p(const int32_t* Gt, const int32_t* h, const int32_t* T, int L, int q, int R, int D, int B,
                           const uint32_t* inst, uint32_t k0, uint32_t k1, int start, int mode, int8_t* out_s,
                           long long* out_de, long long* out_st, int waves, hipStream_t st) {
  const int cpl = D <= ARP_WAVE ? 1 : 2;
  const size_t lds = sizeof(int32_t) * ((size_t)D * D + ARP_WAVE * cpl);
  dim3 grid(B), block(ARP_WAVE * waves);
  if (lds > 65536) {
    hipFuncSetAttribute(cpl == 1 ? (const void*)arpeggio_fx_kernel<P, 1> : (const void*)arpeggio_fx_kernel<P, 2>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  }
  if (cpl == 1)
    hipLaunchKernelGGL((arpeggio_fx_kernel<P, 1>), grid, block, lds, st, Gt, h, T, L, q, R, D, inst, k0, k1, start,
                       mode, out_s, out_de, out_st);
  else
    hipLaunchKernelGGL((arpeggio_fx_kernel<P, 2>), grid, block, lds, st, Gt, h, T, L, q, R, D, inst, k0, k1, start,
                       mode, out_s, out_de, out_st);
  return hipGetLastError();
}

hipError_t launch_arpeggio_fx(const int32_t* Gt, const int32_t* h, const int32_t* T, int L, int q, int P, int R, int D,
                              int B, const uint32_t* inst, uint32_t k0, uint32_t k1, int start, int mode,
                              int8_t* out_s, long long* out_de, long long* out_st, int waves, hipStream_t st) {
  switch (P) {
    case 2: return launch_p<2>(Gt, h, T, L, q, R, D, B, inst, k0, k1, start, mode, out_s, out_de, out_st, waves, st);
    case 4: return launch_p<4>(Gt, h, T, L, q, R, D, B, inst, k0, k1, start, mode, out_s, out_de, out_st, waves, st);
    default: return launch_p<6>(Gt, h, T, L, q, R, D, B, inst, k0, k1, start, mode, out_s, out_de, out_st, waves, st);
  }
}

}  // namespace arp
