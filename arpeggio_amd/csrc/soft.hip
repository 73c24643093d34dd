// Soft output on MI355X: max-log bit LLRs from the MMSE output and from a candidate list.
//
// Reference semantics: mimo.mmse_soft_llr (arpeggio/mimo/mimo.py:111-129), mimo.select / dist (mimo.py:416-425) and
// mimo.list_llr (mimo.py:434-462), as softlib.detect_llr (arpeggio/mimo/softlib.py:19-65) uses them: LLR > 0 favours
// bit 1, dimension-major with the MSB first within a dimension (mimo.levels_to_bits, mimo.py:45-51).
//
// mmse_llr_kernel: one thread per (frame, coordinate). With the unbiased MMSE model z = mu l + e,
// Var(e) = E[l^2] mu (1 - mu), the squared distance to each level is d(l) = (z - mu l)^2 / max(Var, 1e-12), and
// LLR(bit k) = (min over levels with bit k = 0 - min over levels with bit k = 1) / 2.
//
// list_llr_kernel<DP>: one workgroup of max(64, DP) threads per frame; thread t keeps receive row t of H in registers
// and is coordinate t for the per-bit work. The candidates (R anneal outputs, then the MMSE levels) are formed into
// LDS as the decode kernels form them; each candidate's ||y - H l||^2 is a per-row dot product summed over the
// workgroup, and the first minimum is the decision. The list is completed, as mimo.list_llr does, by the D nb
// neighbours of the decision that differ from it in one Gray bit. Neighbour (j, k) moves coordinate j by delta, so
// its distance is d_best - 2 delta g_j + delta^2 n_j with g = H^T r (r the decision's residual) and n_j = ||H_.j||^2
// (H passes through LDS once, transposed, for the column sums). For bit (j, k) with the decision's value b:
//   m_b  = min(list candidates with bit b, every neighbour but (j, k))   (the smallest two neighbours suffice)
//   m_!b = min(list candidates with bit !b, neighbour (j, k))
// and LLR = clip((m_0 - m_1) / N0, +-clip).
#include "kernels.h"
#include "adaptive.h"

#include <cfloat>
#include <cstdint>

namespace arp {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// Gray bits of a level, MSB in bit nb-1 (16-QAM: -3 00, -1 01, +1 11, +3 10; QPSK: -1 0, +1 1)
__device__ __forceinline__ int level_gray(int l, int M) {
  if (M == 4) return l > 0 ? 1 : 0;
  return ((l > 0) ? 2 : 0) | ((l == 1 || l == -1) ? 1 : 0);
}

// the level whose Gray bits are g: the MSB is the sign, the LSB set for the inner levels
__device__ __forceinline__ int gray_level(int g, int M) {
  if (M == 4) return g ? 1 : -1;
  return ((g & 2) ? 1 : -1) * ((g & 1) ? 1 : 3);
}

__global__ __launch_bounds__(256) void mmse_llr_kernel(const float* __restrict__ z, const float* __restrict__ mu,
                                                       int B, int D, int M, float* __restrict__ llr) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (long)B * D) return;
  const float zz = z[x], m = mu[x];
  const float el2 = M == 4 ? 1.f : 5.f;
  const float var = fmaxf(el2 * m * (1.f - m), 1e-12f);
  const int nb = M == 4 ? 1 : 2, nl = M == 4 ? 2 : 4;
  float m0[2] = {FLT_MAX, FLT_MAX}, m1[2] = {FLT_MAX, FLT_MAX};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= nl) break;
    const float l = M == 4 ? (float)(2 * i - 1) : (float)(2 * i - 3);
    const float e = zz - m * l;
    const float d = e * e / var;
    const int g = level_gray((int)l, M);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (k >= nb) break;
      const int bit = (g >> (nb - 1 - k)) & 1;
      if (bit) m1[k] = fminf(m1[k], d);
      else m0[k] = fminf(m0[k], d);
    }
  }
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (k < nb) llr[x * nb + k] = (m0[k] - m1[k]) * 0.5f;
}

// Merge of (smallest, its index, second smallest): the running pair of the neighbour distances.
__device__ __forceinline__ void min2_merge(float& a1, int& ia, float& a2, float b1, int ib, float b2) {
  if (b1 < a1 || (b1 == a1 && ib < ia)) {
    a2 = fminf(a1, b2);
    a1 = b1;
    ia = ib;
  } else {
    a2 = fminf(a2, b1);
  }
}

template <int DP>
__global__ __launch_bounds__(DP > 64 ? DP : 64) void list_llr_kernel(
    const float* __restrict__ H, const float* __restrict__ y, const int8_t* __restrict__ xg,
    const int8_t* __restrict__ spins, int R, int P, int D, int Mr, int M, int delta, int vec4,
    const float* __restrict__ N0v, float N0s, float clip, int8_t* __restrict__ lhat, float* __restrict__ dist_out,
    float* __restrict__ llr) {
  constexpr int NTH = DP > 64 ? DP : 64, NW = NTH / 64;
  extern __shared__ float sm[];
  const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int C = R + 1;
  float* Ht = sm;                    // DP x (DP + 1): Ht[j (DP + 1) + r] = H[r][j]
  float* Lc = Ht + DP * (DP + 1);    // C x DP candidate levels
  float* dv = Lc + C * DP;           // C distances (padded to a multiple of 4)
  float* rs = dv + ((C + 3) & ~3);   // DP: the decision's residual
  float* red = rs + DP;              // 4 NW: cross-wave partials
  const float* Hb = H + (size_t)b * Mr * D;
  const bool row = t < Mr, coord = t < D;
  float hr[DP];
  if (vec4) {
#pragma unroll
    for (int q = 0; q < DP / 4; ++q) {
      floatx4 v = {0.f, 0.f, 0.f, 0.f};
      if (row && 4 * q < D) v = *reinterpret_cast<const floatx4*>(Hb + (size_t)t * D + 4 * q);
#pragma unroll
      for (int e = 0; e < 4; ++e) hr[4 * q + e] = v[e];
    }
  } else {
#pragma unroll
    for (int j = 0; j < DP; ++j) hr[j] = (row && j < D) ? Hb[(size_t)t * D + j] : 0.f;
  }
  const float yi = row ? y[(size_t)b * Mr + t] : 0.f;
  const float lmax = M == 4 ? 1.f : 3.f;
  const int N = delta ? P * D : (M == 4 ? D : 2 * D);
  const int x0 = coord ? (int)xg[(size_t)b * D + t] : 0;
  for (int c = 0; c < C; ++c) {
    float l = 0.f;
    if (coord) {
      if (c == R) {
        l = (float)x0;
      } else {
        const int8_t* s = spins + ((size_t)b * R + c) * N;
        if (delta) {
          int d = 0;
          for (int p = 0; p < P; ++p) d += s[p * D + t];
          l = fminf(fmaxf((float)(x0 + d), -lmax), lmax);
        } else {
          l = M == 4 ? (float)s[t] : (float)(2 * s[t] + s[D + t]);
        }
      }
    }
    if (t < DP) Lc[c * DP + t] = l;
  }
  // H transposed into LDS for the column sums (thread t writes row t: bank (j + t) mod 64, no conflicts)
  if (t < DP)
#pragma unroll
    for (int j = 0; j < DP; ++j) Ht[j * (DP + 1) + t] = hr[j];
  __syncthreads();
  // candidate distances; the decision is the first minimum (MMSE last), its residual row kept per thread
  int best = 0;
  float bd = 0.f, ebest = 0.f;
  for (int c = 0; c < C; ++c) {
    const float* lc = Lc + c * DP;
    float sacc = 0.f;
#pragma unroll
    for (int j = 0; j < DP; j += 4) {
      const floatx4 lv = *reinterpret_cast<const floatx4*>(lc + j);
#pragma unroll
      for (int e = 0; e < 4; ++e) sacc = __builtin_fmaf(hr[j + e], lv[e], sacc);
    }
    const float e = row ? yi - sacc : 0.f;
    float part = e * e;
    for (int o = 32; o; o >>= 1) part += __shfl_xor(part, o);
    if (NW > 1) {
      if (lane == 0) red[wv] = part;
      __syncthreads();
      part = 0.f;
      for (int w = 0; w < NW; ++w) part += red[w];
      __syncthreads();
    }
    if (t == 0) dv[c] = part;
    if (c == 0 || part < bd) {
      bd = part;
      best = c;
      ebest = e;
    }
  }
  if (t < DP) rs[t] = ebest;
  __syncthreads();
  if (dist_out)
    for (int c = t; c < C; c += NTH) dist_out[(size_t)b * C + c] = dv[c];
  // g_j = H_.j^T r and n_j = ||H_.j||^2 for coordinate j = t
  float g = 0.f, n = 0.f;
  if (t < DP) {
    const float* hc = Ht + t * (DP + 1);
    for (int r = 0; r < DP; r += 4) {
      const floatx4 rv = *reinterpret_cast<const floatx4*>(rs + r);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float h = hc[r + e];
        g = __builtin_fmaf(h, rv[e], g);
        n = __builtin_fmaf(h, h, n);
      }
    }
  }
  const int nb = M == 4 ? 1 : 2;
  const int lb = coord ? (int)Lc[best * DP + t] : 0;
  if (lhat && coord) lhat[(size_t)b * D + t] = (int8_t)lb;
  const int gb = level_gray(lb, M);
  // neighbour distances, and the smallest two over the frame (index j nb + k)
  float dn[2] = {FLT_MAX, FLT_MAX};
  float a1 = FLT_MAX, a2 = FLT_MAX;
  int ia = 0x7FFFFFFF;
  if (coord) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (k >= nb) break;
      const int lf = gray_level(gb ^ (1 << (nb - 1 - k)), M);
      const float dl = (float)(lf - lb);
      dn[k] = bd - 2.f * dl * g + dl * dl * n;
      min2_merge(a1, ia, a2, dn[k], t * nb + k, FLT_MAX);
    }
  }
  for (int o = 32; o; o >>= 1) {
    const float b1 = __shfl_xor(a1, o), b2 = __shfl_xor(a2, o);
    const int ib = __shfl_xor(ia, o);
    min2_merge(a1, ia, a2, b1, ib, b2);
  }
  if (NW > 1) {
    __syncthreads();
    if (lane == 0) {
      red[3 * wv] = a1;
      red[3 * wv + 1] = __builtin_bit_cast(float, ia);
      red[3 * wv + 2] = a2;
    }
    __syncthreads();
    a1 = FLT_MAX;
    a2 = FLT_MAX;
    ia = 0x7FFFFFFF;
    for (int w = 0; w < NW; ++w) min2_merge(a1, ia, a2, red[3 * w], __builtin_bit_cast(int, red[3 * w + 1]),
                                            red[3 * w + 2]);
  }
  if (!coord) return;
  // per-bit minima over the list
  float m0[2] = {FLT_MAX, FLT_MAX}, m1[2] = {FLT_MAX, FLT_MAX};
  for (int c = 0; c < C; ++c) {
    const int gc = level_gray((int)Lc[c * DP + t], M);
    const float dc = dv[c];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (k >= nb) break;
      if ((gc >> (nb - 1 - k)) & 1) m1[k] = fminf(m1[k], dc);
      else m0[k] = fminf(m0[k], dc);
    }
  }
  const float n0 = N0v ? N0v[b] : N0s;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (k >= nb) break;
    const int bit = (gb >> (nb - 1 - k)) & 1;
    const float other = (ia == t * nb + k) ? a2 : a1;   // the smallest neighbour but (j, k)
    float mb = bit ? m1[k] : m0[k], mn = bit ? m0[k] : m1[k];
    mb = fminf(mb, other);
    mn = fminf(mn, dn[k]);
    const float v0 = bit ? mn : mb, v1 = bit ? mb : mn;
    llr[((size_t)b * D + t) * nb + k] = fminf(fmaxf((v0 - v1) / n0, -clip), clip);
  }
}

size_t list_llr_lds_bytes(int DP, int R) {
  const int NW = DP > 64 ? DP / 64 : 1;
  return sizeof(float) * ((size_t)DP * (DP + 1) + (size_t)(R + 1) * DP + ((R + 4) & ~3) + DP + 4 * NW);
}

hipError_t launch_mmse_llr(const float* z, const float* mu, int B, int D, int M, float* llr, hipStream_t st) {
  const long n = (long)B * D;
  hipLaunchKernelGGL(mmse_llr_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, z, mu, B, D, M, llr);
  return hipGetLastError();
}

hipError_t launch_list_llr(const float* H, const float* y, const int8_t* xg, const int8_t* spins, int B, int R, int P,
                           int D, int Mr, int M, int delta, const float* N0v, float N0s, float clip, int8_t* lhat,
                           float* dist_out, float* llr, hipStream_t st) {
  const int DP = D <= 32 && Mr <= 32 ? 32 : D <= 64 && Mr <= 64 ? 64 : 128;
  const size_t lds = list_llr_lds_bytes(DP, R);
  const int vec4 = (D % 4 == 0) && ((reinterpret_cast<uintptr_t>(H) & 15) == 0);
#define ARP_LIST_LLR(DPv)                                                                                            \
  do {                                                                                                               \
    if (lds > 65536)                                                                                                 \
      (void)hipFuncSetAttribute((const void*)list_llr_kernel<DPv>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);  \
    hipLaunchKernelGGL((list_llr_kernel<DPv>), dim3(B), dim3(DPv > 64 ? DPv : 64), lds, st, H, y, xg, spins, R, P, D, \
                       Mr, M, delta, vec4, N0v, N0s, clip, lhat, dist_out, llr);                                           \
  } while (0)
  if (DP == 32) ARP_LIST_LLR(32);
  else if (DP == 64) ARP_LIST_LLR(64);
  else ARP_LIST_LLR(128);
#undef ARP_LIST_LLR
  return hipGetLastError();
}

}  // namespace arp

extern "C" int arp_soft_mmse_llr(const float* z, const float* mu, int B, int D, int M, float* llr, void* stream) {
  if (D < 1 || D > 128 || B < 0) return arp::set_error(ARP_E_RANGE, "soft_mmse_llr: D or B out of range");
  if (M != 4 && M != 16) return arp::set_error(ARP_E_ARG, "soft_mmse_llr: M must be 4 or 16");
  if (B == 0) return arp::set_error(0, "");
  if (!z || !mu || !llr) return arp::set_error(ARP_E_ARG, "soft_mmse_llr: null pointer");
  return arp::hip_status(arp::launch_mmse_llr(z, mu, B, D, M, llr, (hipStream_t)stream), "soft_mmse_llr launch");
}

extern "C" int arp_soft_list_llr(const float* H, const float* y, const int8_t* xg, const int8_t* spins, int B, int R,
                                 int P, int D, int Mr, int M, int delta, const float* N0, float N0_scalar, float clip,
                                 int8_t* lhat, float* dist, float* llr, void* stream) {
  if (D < 1 || D > 128 || Mr < 1 || Mr > 128) return arp::set_error(ARP_E_RANGE, "soft_list_llr: D or Mr outside 1..128");
  if (R < 0 || R > 63) return arp::set_error(ARP_E_RANGE, "soft_list_llr: R outside 0..63");
  if (M != 4 && M != 16) return arp::set_error(ARP_E_ARG, "soft_list_llr: M must be 4 or 16");
  if (delta && (P < 1 || P > 6)) return arp::set_error(ARP_E_ARG, "soft_list_llr: P outside 1..6");
  if (!(clip > 0.f)) return arp::set_error(ARP_E_ARG, "soft_list_llr: clip must be positive");
  if (B < 0) return arp::set_error(ARP_E_ARG, "soft_list_llr: B < 0");
  if (B == 0) return arp::set_error(0, "");
  if (!H || !y || !xg || (R > 0 && !spins) || !llr) return arp::set_error(ARP_E_ARG, "soft_list_llr: null pointer");
  if (!N0 && !(N0_scalar > 0.f)) return arp::set_error(ARP_E_ARG, "soft_list_llr: N0 must be positive");
  return arp::hip_status(arp::launch_list_llr(H, y, xg, spins, B, R, P, D, Mr, M, delta, N0, N0_scalar, clip, lhat,
                                              dist, llr, (hipStream_t)stream),
                         "soft_list_llr launch");
}
