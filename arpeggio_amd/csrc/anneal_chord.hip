// Fixed-point Chord annealer for MI355X (gfx950, wave64).
//
// Semantics: arpeggio/mimo/chordfx.py:80-118 (chord_fx), the definition the reference's CUDA kernel
// arpeggio/mimo/chord_kernel.cu:40-175 matches bit for bit: copy coloring (at step k only copy k mod P may flip),
// Onsager term c_k against the previous step's flips, pinning q, one snapshot per step with every accepted flip
// committed, best-of-trajectory (first minimum), mode 'best' = lowest-energy anneal per frame (lowest r on ties).
//
// Layout: one workgroup per frame; the frame's symmetric coupling matrix J (N x N int16) sits in LDS when it fits
// (N = 256 takes 128 KiB of the CU's 160 KiB), otherwise rows are streamed from global memory through L2. One wave
// owns one anneal: lane l holds spins i = t*64 + l (slot t < MPL). Spins are wave-uniform 64-bit masks per slot (one
// ballot), so the copy test, the snapshot commit (masked ballot + find-first-set per committed flip), the energy and
// the best state are scalar work; fields u are per lane. A committed flip of spin j reads row j of J (= column j)
// once, N consecutive int16 words, and updates every lane's fields.
#include "kernels.h"
#include "launch_bounds.h"

namespace arp {

template <int MPL, bool LDSJ>
__global__ __launch_bounds__(ARP_CHORD_MAX_THREADS) void chord_fx_kernel(const int16_t* __restrict__ Jg, const int32_t* __restrict__ hg,
                                                       const int32_t* __restrict__ h0g, const int32_t* __restrict__ Tq,
                                                       const int32_t* __restrict__ Cq, int q, int S, int P, int R,
                                                       int N, uint32_t k0, uint32_t k1, int inst0, int mode,
                                                       int32_t* __restrict__ out_e, int8_t* __restrict__ out_s,
                                                       unsigned long long* __restrict__ stats) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int b = blockIdx.x;
  const size_t NN = (size_t)N * N;
  const int16_t* Jb = Jg + (size_t)b * NN;
  // LDS: [J (N*N int16, padded to 16 B)] [red: up to 16 waves x (e, r)]
  const size_t jbytes = LDSJ ? ((NN * 2 + 15) & ~(size_t)15) : 0;
  int16_t* J = LDSJ ? reinterpret_cast<int16_t*>(smem) : const_cast<int16_t*>(Jb);
  long long* red_e = reinterpret_cast<long long*>(smem + jbytes);
  int* red_r = reinterpret_cast<int*>(smem + jbytes + 16 * 8);
  if (LDSJ) {
    // N % 8 == 0 is guaranteed by the host (N is a multiple of 32), so N*N int16 is a whole number of 16 B words
    const int4* src = reinterpret_cast<const int4*>(Jb);
    int4* dst = reinterpret_cast<int4*>(J);
    for (size_t x = threadIdx.x; x < NN / 8; x += blockDim.x) dst[x] = src[x];
  }
  __syncthreads();
  const int lane = lane_id(), wv = threadIdx.x >> 6, W = blockDim.x >> 6;
  const uint32_t inst = (uint32_t)(inst0 + b);
  const int per = N / P;
  int32_t hv[MPL], hcv[MPL], jdiag8[MPL];
  int cp[MPL];
  bool ok[MPL];
  uint64_t okm[MPL];
#pragma unroll
  for (int t = 0; t < MPL; ++t) {
    const int i = t * ARP_WAVE + lane;
    ok[t] = i < N;
    okm[t] = ballot(ok[t]);
    hv[t] = ok[t] ? hg[(size_t)b * N + i] : 0;
    hcv[t] = ok[t] ? hv[t] - h0g[(size_t)b * N + i] : 0;  // h - h0: the energy step is 4 s (u - (h - h0)) - 8 J_jj
    jdiag8[t] = ok[t] ? 8 * (int32_t)J[(size_t)i * N + i] : 0;
    cp[t] = ok[t] ? i / per : -1;
  }
  long long wbest_e = 0x7FFFFFFFFFFFFFFFLL;
  int wbest_r = 0x7FFFFFFF;
  uint64_t wbest_m[MPL];
  unsigned long long n_flip = 0, n_eval = 0, n_rng = 0;
  for (int r = wv; r < R; r += W) {
    uint64_t Sm[MPL];  // bit l of slot t: spin t*64 + l is +1
#pragma unroll
    for (int t = 0; t < MPL; ++t) {
      const uint32_t x = philox_w0(0xFFFFFFFFu, (uint32_t)(t * ARP_WAVE + lane), (uint32_t)r, inst, k0, k1);
      Sm[t] = ballot(x >= 0x80000000u) & okm[t];
    }
    int32_t u[MPL];
#pragma unroll
    for (int t = 0; t < MPL; ++t) u[t] = hv[t];
    // u = h + 2 J s: row j of the symmetric J holds column j
#pragma unroll
    for (int tj = 0; tj < MPL; ++tj) {
      const int nj = N - tj * ARP_WAVE < ARP_WAVE ? N - tj * ARP_WAVE : ARP_WAVE;
      for (int l = 0; l < nj; ++l) {
        const int32_t s2 = ((Sm[tj] >> l) & 1ull) ? 2 : -2;
        const int16_t* row = J + (size_t)(tj * ARP_WAVE + l) * N;
#pragma unroll
        for (int t = 0; t < MPL; ++t)
          if (ok[t]) u[t] += s2 * (int32_t)row[t * ARP_WAVE + lane];
      }
    }
    // 2 E0 = sum_i s_i (h_i - 2 h0_i - u_i), one wave reduction per anneal
    long long e;
    {
      int32_t part = 0;
#pragma unroll
      for (int t = 0; t < MPL; ++t) {
        const int32_t s = ((Sm[t] >> lane) & 1ull) ? 1 : -1;
        if (ok[t]) part += s * (2 * hcv[t] - hv[t] - u[t]);
      }
      long long v = part;
#pragma unroll
      for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
      e = __builtin_amdgcn_readfirstlane((int)(v & 0xFFFFFFFF)) |
          ((long long)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32);
    }
    uint32_t r0[MPL], r1[MPL], r2[MPL], r3[MPL];
    uint64_t pf[MPL], best_m[MPL];
#pragma unroll
    for (int t = 0; t < MPL; ++t) {
      pf[t] = 0;
      best_m[t] = Sm[t];
    }
    long long best_e = 0x7FFFFFFFFFFFFFFFLL;
    int p = 0, n = 0;
    for (int k = 0; k < S; ++k) {
      const int nm = n & 3;
      const int32_t T = Tq[k], c = Cq[k];
      const uint32_t A = 4u * (uint32_t)T;
      uint64_t F[MPL];
      bool any = false;
#pragma unroll
      for (int t = 0; t < MPL; ++t) {
        const bool mine = cp[t] == p;
        const uint64_t act = ballot(mine);
        F[t] = 0;
        if (act == 0) continue;
        if (nm == 0) {
          uint32_t c0 = (uint32_t)(n >> 2), c1 = (uint32_t)(t * ARP_WAVE + lane), c2 = (uint32_t)r, c3 = inst;
          philox10(c0, c1, c2, c3, k0, k1);
          r0[t] = c0;
          r1[t] = c1;
          r2[t] = c2;
          r3[t] = c3;
          if (stats) n_rng += __builtin_popcountll(act);
        }
        const uint32_t x = nm == 0 ? r0[t] : nm == 1 ? r1[t] : nm == 2 ? r2[t] : r3[t];
        const bool s_pos = (Sm[t] >> lane) & 1ull;
        const bool prev = (pf[t] >> lane) & 1ull;
        const int32_t z = (s_pos ? u[t] : -u[t]) + q - (prev ? -c : c);
        F[t] = ballot(mine && accept_fx(x, A, z + 2 * T));
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
          const int32_t sj4 = ((Sm[t] >> l) & 1ull) ? 4 : -4;
          const int32_t uj = readlane(u[t], l);
          const int32_t hcj = readlane(hcv[t], l);
          const int32_t jd8 = readlane(jdiag8[t], l);
          e += (long long)sj4 * (uj - hcj) - jd8;
          const int16_t* row = J + (size_t)j * N;
#pragma unroll
          for (int tt = 0; tt < MPL; ++tt)
            if (ok[tt]) u[tt] -= sj4 * (int32_t)row[tt * ARP_WAVE + lane];
          Sm[t] ^= 1ull << l;
          if (stats) ++n_flip;
        }
      }
      if (++p == P) {
        p = 0;
        ++n;
      }
      if (any || k == 0) {
        if (stats) ++n_eval;
        if (e < best_e) {
          best_e = e;
#pragma unroll
          for (int t = 0; t < MPL; ++t) best_m[t] = Sm[t];
        }
      }
    }
    if (mode == 0) {
      if (lane == 0) out_e[(size_t)b * R + r] = (int32_t)best_e;
#pragma unroll
      for (int t = 0; t < MPL; ++t)
        if (ok[t]) out_s[((size_t)b * R + r) * N + t * ARP_WAVE + lane] = ((best_m[t] >> lane) & 1ull) ? 1 : -1;
    } else if (best_e < wbest_e) {
      wbest_e = best_e;
      wbest_r = r;
#pragma unroll
      for (int t = 0; t < MPL; ++t) wbest_m[t] = best_m[t];
    }
  }
  if (stats && lane == 0) {
    atomicAdd(&stats[0], n_flip);
    atomicAdd(&stats[1], n_eval);
    atomicAdd(&stats[2], n_rng);
    atomicAdd(&stats[3], (unsigned long long)S * (unsigned long long)((R - wv + W - 1) / W));
  }
  if (mode == 1) {
    if (lane == 0) {
      red_e[wv] = wbest_e;
      red_r[wv] = wbest_r;
    }
    __syncthreads();
    long long be = 0x7FFFFFFFFFFFFFFFLL;
    int br = 0x7FFFFFFF;
    for (int v = 0; v < W; ++v)
      if (red_e[v] < be || (red_e[v] == be && red_r[v] < br)) {
        be = red_e[v];
        br = red_r[v];
      }
    if (wbest_r == br) {
      if (lane == 0) out_e[b] = (int32_t)be;
#pragma unroll
      for (int t = 0; t < MPL; ++t)
        if (ok[t]) out_s[(size_t)b * N + t * ARP_WAVE + lane] = ((wbest_m[t] >> lane) & 1ull) ? 1 : -1;
    }
  }
}

template <int MPL>
static hipError_t launch_mpl(bool lds_j, size_t lds, dim3 grid, dim3 block, hipStream_t st, const int16_t* J,
                             const int32_t* h, const int32_t* h0, const int32_t* Tq, const int32_t* Cq, int q, int S,
                             int P, int R, int N, uint32_t k0, uint32_t k1, int inst0, int mode, int32_t* out_e,
                             int8_t* out_s, unsigned long long* stats) {
  if (lds_j) {
    hipFuncSetAttribute((const void*)chord_fx_kernel<MPL, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((chord_fx_kernel<MPL, true>), grid, block, lds, st, J, h, h0, Tq, Cq, q, S, P, R, N, k0, k1,
                       inst0, mode, out_e, out_s, stats);
  } else {
    hipLaunchKernelGGL((chord_fx_kernel<MPL, false>), grid, block, lds, st, J, h, h0, Tq, Cq, q, S, P, R, N, k0, k1,
                       inst0, mode, out_e, out_s, stats);
  }
  return hipGetLastError();
}

// max_lds: the device's dynamic LDS limit per workgroup; J goes to LDS when it fits.
hipError_t launch_chord_fx(const int16_t* J, const int32_t* h, const int32_t* h0, const int32_t* Tq,
                           const int32_t* Cq, int q, int S, int P, int R, int N, int B, uint32_t k0, uint32_t k1,
                           int inst0, int mode, int32_t* out_e, int8_t* out_s, unsigned long long* stats, int waves,
                           size_t max_lds, hipStream_t st) {
  const size_t jb = (((size_t)N * N * 2 + 15) & ~(size_t)15);
  const size_t red = 16 * 8 + 16 * 4;
  const bool lds_j = jb + red <= max_lds;
  const size_t lds = (lds_j ? jb : 0) + red;
  const int mpl = (N + ARP_WAVE - 1) / ARP_WAVE;
  dim3 grid(B), block(ARP_WAVE * waves);
  switch (mpl) {
    case 1: return launch_mpl<1>(lds_j, lds, grid, block, st, J, h, h0, Tq, Cq, q, S, P, R, N, k0, k1, inst0, mode, out_e, out_s, stats);
    case 2: return launch_mpl<2>(lds_j, lds, grid, block, st, J, h, h0, Tq, Cq, q, S, P, R, N, k0, k1, inst0, mode, out_e, out_s, stats);
    case 3: return launch_mpl<3>(lds_j, lds, grid, block, st, J, h, h0, Tq, Cq, q, S, P, R, N, k0, k1, inst0, mode, out_e, out_s, stats);
    default: return launch_mpl<4>(lds_j, lds, grid, block, st, J, h, h0, Tq, Cq, q, S, P, R, N, k0, k1, inst0, mode, out_e, out_s, stats);
  }
}

}  // namespace arp
