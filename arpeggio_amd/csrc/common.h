// Shared device helpers for the MI355X (gfx950, CDNA4) detector kernels: wave64 ballots, lane reads, and the
// counter-based Philox4x32-10 generator whose counter layout the annealers share with the reference
// (arpeggio/mimo/chordfx.py:28-56, arpeggio/mimo/chord_kernel.cu:14-24, arpeggio/sim/arpeggio_fx.cu:10-22).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ARP_WAVE 64

namespace arp {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of two 32x32->64 multiplies with the Random123
// multipliers and Weyl key increments. Each product is written as one 64-bit multiply so that it issues as a single
// v_mad_u64_u32 (both halves) rather than a v_mul_hi_u32 / v_mul_lo_u32 pair. The key bumps are wave-uniform and
// land on the scalar unit.
__device__ __forceinline__ void philox10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                                         uint32_t k1) {
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    // each three-input XOR is one v_bitop3_b32 (truth table 0x96 = a ^ b ^ c)
    const uint32_t x0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
    const uint32_t x2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
    c0 = x0;
    c1 = lo1;
    c2 = x2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Round keys held in vector registers (every lane the same value) for use inside long loops: the scalar register
// file stays free for the loop's masks and counters, and the key XORs read a vector operand at the same cost.
struct PhiloxKeys {
  uint32_t k0[10], k1[10];
};

__device__ __forceinline__ PhiloxKeys philox_keys_v(uint32_t k0, uint32_t k1) {
  PhiloxKeys K;
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    uint32_t a = k0 + (uint32_t)i * 0x9E3779B9u, b = k1 + (uint32_t)i * 0xBB67AE85u;
    asm volatile("v_mov_b32 %0, %1" : "=v"(K.k0[i]) : "s"(a));
    asm volatile("v_mov_b32 %0, %1" : "=v"(K.k1[i]) : "s"(b));
  }
  return K;
}

__device__ __forceinline__ void philox10_kv(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                                            const PhiloxKeys& K) {
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t x0 = __builtin_amdgcn_bitop3_b32(hi1, c1, K.k0[i], 0x96);
    const uint32_t x2 = __builtin_amdgcn_bitop3_b32(hi0, c3, K.k1[i], 0x96);
    c0 = x0;
    c1 = lo1;
    c2 = x2;
    c3 = lo0;
  }
}

// sets bit l of a wave-uniform 64-bit mask with one scalar instruction
__device__ __forceinline__ void mask_set(uint64_t& m, int l) { asm("s_bitset1_b64 %0, %1" : "+s"(m) : "s"(l) : "scc"); }

__device__ __forceinline__ uint32_t philox_w0(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1) {
  philox10(c0, c1, c2, c3, k0, k1);
  return c0;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

__device__ __forceinline__ uint32_t readlane_u(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (ARP_WAVE - 1)); }

// Exact form of the reference's 64-bit decision  x * A > z * 2^32  (arpeggio_fx.cu:76-79, chord_kernel.cu:114)
// for 0 <= A < 2^31 and any int32 z: with x*A = hi*2^32 + lo, it holds iff hi > z, or hi == z and lo > 0
// (for z < 0 both sides agree because x*A >= 0). One 32x32->64 product instead of a 64-bit signed compare.
__device__ __forceinline__ bool accept_fx(uint32_t x, uint32_t A, int32_t z) {
  const uint64_t pr = (uint64_t)x * A;
  const int32_t hi = (int32_t)(pr >> 32);
  return hi > z || (hi == z && (uint32_t)pr != 0u);
}

// The same test as a threshold that depends on x and A only: x*A > z*2^32  <=>  z < hi + (lo != 0). One threshold per
// (lane, pass) turns every re-test of the first-accept scan into a single compare.
__device__ __forceinline__ int32_t accept_threshold(uint32_t x, uint32_t A) {
  const uint64_t pr = (uint64_t)x * A;
  return (int32_t)(pr >> 32) + ((uint32_t)pr != 0u ? 1 : 0);
}

// The i-th P-bit word with P/2 ones in increasing numeric order (arpeggio_ref59.py:31-38), for P = 2, 4, 6.
__device__ __forceinline__ uint32_t bal_pattern(int P, uint32_t i) {
  // P = 2: 1, 2.  P = 4: 3, 5, 6, 9, 10, 12.  P = 6: the 20 six-bit words of weight 3, ascending.
  if (P == 2) return i ? 2u : 1u;
  if (P == 4) {
    const uint32_t t4 = (3u) | (5u << 4) | (6u << 8) | (9u << 12) | (10u << 16) | (12u << 20);
    return (t4 >> (4 * i)) & 15u;
  }
  // pack 20 words of 6 bits into two 64-bit constants (10 words each)
  const uint64_t lo = (7ull) | (11ull << 6) | (13ull << 12) | (14ull << 18) | (19ull << 24) | (21ull << 30) |
                      (22ull << 36) | (25ull << 42) | (26ull << 48) | (28ull << 54);
  const uint64_t hi = (35ull) | (37ull << 6) | (38ull << 12) | (41ull << 18) | (42ull << 24) | (44ull << 30) |
                      (49ull << 36) | (50ull << 42) | (52ull << 48) | (56ull << 54);
  return (uint32_t)(((i < 10 ? lo : hi) >> (6 * (i < 10 ? i : i - 10))) & 63ull);
}

__device__ __forceinline__ uint32_t bal_count(int P) { return P == 2 ? 2u : (P == 4 ? 6u : 20u); }

}  // namespace arp
