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
// de, best, flips and the spin masks are wave-uniform and live in scalar registers.
// Each (lane, copy) Philox block gives the words of 4 consecutive sweeps, so a block is computed every 4th sweep of
// a copy instead of once per test.
#include "kernels.h"
#include "launch_bounds.h"
#include "adaptive.h"

namespace arp {

// a < b for wave-uniform 64-bit values, on the scalar unit: the sign of the high word of a - b (no overflow within
// the documented ranges)
__device__ __forceinline__ bool s64_lt(long long a, long long b) {
  uint32_t lo, hi;
  asm("s_sub_u32 %0, %2, %4\n\ts_subb_u32 %1, %3, %5"
      : "=&s"(lo), "=s"(hi)
      : "s"((uint32_t)a), "s"((uint32_t)((unsigned long long)a >> 32)), "s"((uint32_t)b),
        "s"((uint32_t)((unsigned long long)b >> 32))
      : "scc");
  return (int32_t)hi < 0;
}

// (a ^ s) + b in one v_xad_u32 (s wave-uniform); the compiler splits the xor from the add when a subtraction follows
__device__ __forceinline__ int32_t xad(int32_t a, int32_t s, int32_t b) {
  int32_t r;
  asm("v_xad_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(s), "v"(b));
  return r;
}

// One anneal per wave. G4 is the frame's Gram column block in LDS pre-scaled by 4 (G4[j*D + i] = 4 G_ij), so a flip of
// lane j with old spin s adds s*G4[j][i] to every w_i and 4 (s w_j + 2 G_jj) to de.
// Spins: wave masks (bit l of Sm[p][c]: spin of copy p at coordinate c*64 + l is +1) while the start is built, then
// per-lane signs (0 / -1) for the passes. Test: lane j of copy p flips iff q_j = s w_j + 2 G_jj <= tq_j, tq_j the
// high word of one multiply-add per pass; q_j = (w_j ^ sg_j) + (2 G_jj - sg_j) is one xor-add from the sign mask sg_j
// (0 / -1), and the committed lane's q_j is also
// its energy step (de += 4 q_j), so a flip reads two lane values (q and the sign). A pass's flips gather in a scalar
// mask that updates the signs, the flip count and the last-lane count once at the end of the pass. The scalar unit
// and the vector unit each issue about one wave instruction per 4 cycles per SIMD, so the loop splits its work
// between them (per committed flip 7 vector and about 11 scalar instructions).
// tests (the reference's comparator passes) = L + flips - flips on the last lane (arpeggio_ref59.py:60-71).
// w_i = h_i - 2 sum_j G_ij m_j = h_i - sum_j G4_ij (m_j / 2) for the state in Sm (m_j = sum_p s_pj, even: P is even)
template <int P, int CPL>
__device__ __forceinline__ void init_fields(const int32_t* __restrict__ G4, const int32_t* __restrict__ hb, int D,
                                            const bool* valid, uint64_t (&Sm)[P][CPL], int32_t* w) {
  const int lane = lane_id();
#pragma unroll
  for (int c = 0; c < CPL; ++c) w[c] = valid[c] ? hb[c * ARP_WAVE + lane] : 0;
#pragma unroll
  for (int c2 = 0; c2 < CPL; ++c2) {
    if (P == 2) {
      // m_j / 2 = +1 where both copies are +1, -1 where both are -1, 0 otherwise: visit only those columns
      const int nv = D - c2 * ARP_WAVE;
      const uint64_t vm = nv >= ARP_WAVE ? ~0ull : ((1ull << nv) - 1ull);
      uint64_t plus = Sm[0][c2] & Sm[P - 1][c2], minus = ~(Sm[0][c2] | Sm[P - 1][c2]) & vm;
      while (plus) {
        const int l = __builtin_ctzll(plus);
        plus &= plus - 1;
        const int32_t* col = G4 + (c2 * ARP_WAVE + l) * D;
#pragma unroll
        for (int c = 0; c < CPL; ++c) w[c] -= col[c * ARP_WAVE + lane];
      }
      while (minus) {
        const int l = __builtin_ctzll(minus);
        minus &= minus - 1;
        const int32_t* col = G4 + (c2 * ARP_WAVE + l) * D;
#pragma unroll
        for (int c = 0; c < CPL; ++c) w[c] += col[c * ARP_WAVE + lane];
      }
    } else {
      const int nj = D - c2 * ARP_WAVE < ARP_WAVE ? D - c2 * ARP_WAVE : ARP_WAVE;
      for (int l = 0; l < nj; ++l) {
        int half_m = 0;
#pragma unroll
        for (int p = 0; p < P; ++p) half_m += (int)((Sm[p][c2] >> l) & 1ull);
        half_m -= P / 2;
        if (half_m != 0) {
          const int32_t* col = G4 + (c2 * ARP_WAVE + l) * D;
#pragma unroll
          for (int c = 0; c < CPL; ++c) w[c] -= half_m * col[c * ARP_WAVE + lane];
        }
      }
    }
  }
}

// Start energy up to a constant: 2 E0 + 2 P tr(G) = -(h + w).m (2 E0 = -2 h.m + 2 m^T G m - 2 P tr G and
// m^T G m = (h - w).m / 2), summed over the wave in 64 bits.
template <int P, int CPL>
__device__ __forceinline__ long long start_energy(const int32_t* __restrict__ hb, const bool* valid,
                                                  uint64_t (&Sm)[P][CPL], const int32_t* w) {
  const int lane = lane_id();
  long long e = 0;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    int m = 0;
#pragma unroll
    for (int p = 0; p < P; ++p) m += (int)((Sm[p][c] >> lane) & 1ull);
    m = 2 * m - P;
    if (valid[c]) e -= (long long)(hb[c * ARP_WAVE + lane] + w[c]) * m;
  }
#pragma unroll
  for (int o = 32; o; o >>= 1) e += __shfl_xor(e, o);
  return e;
}

// DS: D as a compile-time constant (0: D at run time); with it the column address l*D folds into the LDS address
// shift and the lane masks into constants.
// fmix32 (the MurmurHash3 finaliser) and xorshift32 (13, 17, 5): the production engine's stream primitives
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  return h ^ (h >> 16);
}

__device__ __forceinline__ uint32_t xs32(uint32_t x) {
  x ^= x << 13;
  x ^= x >> 17;
  return x ^ (x << 5);
}

// RNG 1: the production engine's xorshift32 lane streams (one state per coordinate, seeded from the key, instance
// and anneal by fmix32; start spin of copy p = bit 31 - p of the first state, each pass draws u = x >> 16 from the
// next, RB = 16 fraction bits), as arpeggio/sim/x3model/chord_ref.py:17-45 and 99-107 define them. Cq (per pass, may be
// NULL): an Onsager term subtracted from the decision bound, B_k = q - c_k + 2 T_k (chord_ref.py:102), read for RNG 1
// and RNG 2 (Philox streams with the term); RNG 0, the plain annealers, has none.
template <int P, int CPL, int MODE, bool FULL, int DS = 0, int RNG = 0>
__device__ __forceinline__ void arpeggio_anneal(const int32_t* __restrict__ G4, const int32_t* __restrict__ hb,
                                                const int32_t* __restrict__ T, int L, int q, int D, uint32_t r,
                                                uint32_t inst, uint32_t k0, uint32_t k1, int start,
                                                int8_t* __restrict__ out_s, long long* __restrict__ out_de,
                                                long long* __restrict__ out_st, size_t t,
                                                const uint32_t* __restrict__ thr, int gate,
                                                int8_t* __restrict__ used, const int32_t* __restrict__ Cq = nullptr) {
  if (DS) D = DS;
  const int lane = lane_id();
  bool valid[CPL];
  int32_t w[CPL], dj2[CPL];
  uint64_t vmask[CPL];
  uint64_t Sm[P][CPL], Rm[P][CPL];  // start spins as wave masks; Rm: the random start beside a learned one
  uint32_t xsr[CPL];                // RNG 1: this lane's xorshift state per column block
  const uint32_t xs_base = RNG == 1 ? fmix32(fmix32(fmix32(k0 ^ inst) ^ k1) ^ r) : 0u;
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const int j = c * ARP_WAVE + lane;
    valid[c] = j < D;
    const int nv = D - c * ARP_WAVE;
    vmask[c] = nv >= ARP_WAVE ? ~0ull : (nv <= 0 ? 0ull : ((1ull << nv) - 1ull));
    dj2[c] = valid[c] ? G4[j * D + j] >> 1 : 0;
    if (RNG == 1) {
      const uint32_t st = fmix32(xs_base + (uint32_t)(j + 1) * 0x9E3779B9u);
      xsr[c] = xs32(st ? st : 0x9E3779B9u);
#pragma unroll
      for (int p = 0; p < P; ++p) Sm[p][c] = ballot(valid[c] && ((xsr[c] >> (31 - p)) & 1u));
    } else if (thr) {
      // learned start: spin +1 iff x + thr_pj >= 2^32, x the random start's word; Rm holds the random start
#pragma unroll
      for (int p = 0; p < P; ++p) {
        uint32_t x = 0u, th = 0u;
        if (valid[c]) {  // lanes past D read nothing: the threshold block holds P*D words per frame
          x = philox_w0(0xFFFFFFFFu, (uint32_t)(p * D + j), r, inst, k0, k1);
          th = thr[p * D + j];
        }
        Sm[p][c] = ballot(valid[c] && (uint64_t)x + th >= 0x100000000ull);
        Rm[p][c] = ballot(valid[c] && x >= 0x80000000u);
      }
    } else if (start == 0) {
#pragma unroll
      for (int p = 0; p < P; ++p)
        Sm[p][c] = ballot(valid[c] &&
                          philox_w0(0xFFFFFFFFu, (uint32_t)(p * D + j), r, inst, k0, k1) >= 0x80000000u);
    } else {
      const uint32_t x = valid[c] ? philox_w0(0xFFFFFFFFu, (uint32_t)j, r, inst, k0, k1) : 0u;
      const uint32_t bits = bal_pattern(P, (uint32_t)(((uint64_t)x * bal_count(P)) >> 32));
#pragma unroll
      for (int p = 0; p < P; ++p) Sm[p][c] = ballot(valid[c] && ((bits >> p) & 1u));
    }
  }
  init_fields<P, CPL>(G4, hb, D, valid, Sm, w);
  if (thr) {
    // energy gate: keep the learned start unless the random start (Rm) has lower energy
    bool take = true;
    if (gate) {
      int32_t wr[CPL];
      init_fields<P, CPL>(G4, hb, D, valid, Rm, wr);
      const long long el = start_energy<P, CPL>(hb, valid, Sm, w);
      const long long er = start_energy<P, CPL>(hb, valid, Rm, wr);
      take = el <= er;
      if (!take) {
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          w[c] = wr[c];
#pragma unroll
          for (int p = 0; p < P; ++p) Sm[p][c] = Rm[p][c];
        }
      }
    }
    if (used && lane == 0) used[t] = take ? 1 : 0;
  }
  // From here the spins live as per-lane signs: sgp[p][c] = 0 / -1 for spin +1 / -1 of copy p at coordinate
  // c*64 + lane (Bs: the best state's).
  int32_t sgp[P][CPL], Bs[P][CPL];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      sgp[p][c] = __builtin_amdgcn_inverse_ballot_w64(Sm[p][c]) ? 0 : -1;
      Bs[p][c] = sgp[p][c];
    }
  long long de4 = 0, best4 = 1LL << 60;   // de / 4
  int32_t fc[CPL];                        // this lane's flip count
#pragma unroll
  for (int c = 0; c < CPL; ++c) fc[c] = 0;
  const PhiloxKeys K = philox_keys_v(k0, k1);
  uint32_t rw[P][CPL][4];
  for (int kb = 0; kb < L; kb += 4 * P) {
    const int n4 = kb / (4 * P);
    int32_t Tb[4 * P], Cb[4 * P];
#pragma unroll
    for (int i = 0; i < 4 * P; ++i) Tb[i] = (FULL || kb + i < L) ? T[kb + i] : 0;  // FULL: L % 4P == 0
#pragma unroll
    for (int i = 0; i < 4 * P; ++i) Cb[i] = (RNG != 0 && Cq && (FULL || kb + i < L)) ? Cq[kb + i] : 0;
#pragma unroll
    for (int nm = 0; nm < 4; ++nm) {
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int k = kb + nm * P + p;
        if (!FULL && k >= L) break;
        if (RNG != 1 && nm == 0) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            uint32_t c0 = (uint32_t)n4, c1 = (uint32_t)(p * D + c * ARP_WAVE + lane), c2 = r, c3 = inst;
            philox10_kv(c0, c1, c2, c3, K);
            rw[p][c][0] = c0;
            rw[p][c][1] = c1;
            rw[p][c][2] = c2;
            rw[p][c][3] = c3;
          }
        }
        const int32_t Tk = Tb[nm * P + p];
        const uint32_t A = 4u * (uint32_t)Tk;
        // Decision x A > (q_j + B_k) 2^32 (B_k = q + 2 T_k, q_j = s w_j + 2 G_jj) as q_j <= tq with
        // tq = floor((x A - 1) / 2^32) - B_k: the high word of one 32x32+64 multiply-add with addend -1 (x A = 0
        // gives -1), then one subtraction; the documented ranges keep tq in int32.
        const int32_t Bk = q + 2 * Tk - (RNG != 0 ? Cb[nm * P + p] : 0);
        int32_t tq[CPL];
        if (RNG == 1) {
          // u A > (q_j + B_k) 2^16 with the 16-bit draw u of the lane's next xorshift state
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            xsr[c] = xs32(xsr[c]);
            tq[c] = (int32_t)(((uint64_t)(xsr[c] >> 16) * A + ~0ull) >> 16) - Bk;
          }
        } else {
#pragma unroll
          for (int c = 0; c < CPL; ++c) tq[c] = (int32_t)(((uint64_t)rw[p][c][nm] * A + ~0ull) >> 32) - Bk;
        }
        // Within a pass a lane flips at most once and is then behind the scan, so the lanes still to be tested keep
        // their start-of-pass signs sg and q = (w ^ sg) - sg + dj2.
        if (MODE == 0) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            // fm collects the pass's flips; the signs take them at the end of the pass. The committed lane's old
            // spin is its start-of-pass sign, read from that lane.
            const int32_t sg = sgp[p][c];
            // q = (w ^ sg) - sg + dj2 = (w ^ sg) + e2: one xor-add per re-test. With one column block the field
            // update is an xor-add too; with two or more, the plain form measures faster (the loads of the column
            // blocks then issue together).
            const int32_t e2 = dj2[c] - sg;
            uint64_t rem = vmask[c], fm = 0;  // rem: lanes the scan has not passed yet
            int32_t qv = CPL == 1 ? (w[c] ^ sg) + e2 : ((w[c] ^ sg) - sg) + dj2[c];
            uint64_t m = ballot(qv <= tq[c]) & rem;
            while (m) {
              const int l = __builtin_ctzll(m);
              de4 += readlane(qv, l);
              const int32_t sl = readlane(sg, l);   // 0: old spin +1 (w += 4 G col), -1: old spin -1 (w -= ...)
              const int32_t* col = G4 + (c * ARP_WAVE + l) * D;
              if constexpr (CPL == 1) {
                w[0] = xad(col[lane], sl, w[0]) - sl;
                qv = (w[0] ^ sg) + e2;
              } else {
#pragma unroll
                for (int cc = 0; cc < CPL; ++cc) w[cc] += (col[cc * ARP_WAVE + lane] ^ sl) - sl;
                qv = ((w[c] ^ sg) - sg) + dj2[c];
              }
              mask_set(fm, l);
              rem = vmask[c] & ((~1ull) << l);  // the scan passes lanes in index order: the lanes above l remain
              m = ballot(qv <= tq[c]) & rem;
            }
            if (fm) {  // passes without a flip skip the sign update (the empty asm keeps the branch)
              asm volatile("");
              const bool fl = __builtin_amdgcn_inverse_ballot_w64(fm);
              sgp[p][c] ^= fl ? -1 : 0;
              fc[c] += fl ? 1 : 0;
            }
          }
        } else {
          // 'par': every lane of the pass decided on one snapshot, the flips then applied in index order
          uint64_t fm[CPL];
#pragma unroll
          for (int c = 0; c < CPL; ++c)
            fm[c] = ballot(((w[c] ^ sgp[p][c]) - sgp[p][c]) + dj2[c] <= tq[c]) & vmask[c];
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            const int32_t sg = sgp[p][c];
            uint64_t m = fm[c];
            while (m) {
              const int l = __builtin_ctzll(m);
              m &= m - 1;
              de4 += readlane(((w[c] ^ sg) - sg) + dj2[c], l);
              const int32_t sl = readlane(sg, l);
              const int32_t* col = G4 + (c * ARP_WAVE + l) * D;
#pragma unroll
              for (int cc = 0; cc < CPL; ++cc) w[cc] += (col[cc * ARP_WAVE + lane] ^ sl) - sl;
            }
            const bool fl = __builtin_amdgcn_inverse_ballot_w64(fm[c]);
            sgp[p][c] ^= fl ? -1 : 0;
            fc[c] += fl ? 1 : 0;
          }
        }
        // best-of-trajectory, strict (the first minimum wins). The test runs on the scalar unit (sign of the 64-bit
        // difference) and the copy behind a wave-uniform branch: the empty asm keeps the compiler from turning the
        // copy into vector selects on every pass.
        if (s64_lt(de4, best4)) {
          asm volatile("");
          best4 = de4;
#pragma unroll
          for (int pp = 0; pp < P; ++pp)
#pragma unroll
            for (int c = 0; c < CPL; ++c) Bs[pp][c] = sgp[pp][c];
        }
      }
    }
  }
  // the anneal's spin row from a wave-uniform base and a lane offset computed here (the volatile mbcnt pair keeps the
  // compiler from hoisting a per-lane 64-bit address out of the pass loop, where it would hold two registers)
  int8_t* const srow = out_s + t * (size_t)(P * D);
  int lane_now;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lane_now));
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    const int j = c * ARP_WAVE + lane_now;
    if (valid[c]) {
#pragma unroll
      for (int p = 0; p < P; ++p) srow[p * D + j] = Bs[p][c] ? -1 : 1;
    }
  }
  // flips = the wave's sum of the lane counts; tests = L + flips - flips on the last lane (D - 1)
  int fl = 0;
#pragma unroll
  for (int c = 0; c < CPL; ++c) fl += fc[c];
#pragma unroll
  for (int o = 32; o; o >>= 1) fl += __shfl_xor(fl, o);
  int last = 0;
#pragma unroll
  for (int c = 0; c < CPL; ++c)
    if (c == (D - 1) / ARP_WAVE) last = readlane(fc[c], (D - 1) % ARP_WAVE);
  if (lane == 0) {
    out_de[t] = 4 * best4;
    out_st[2 * t] = fl;
    out_st[2 * t + 1] = MODE == 0 ? (long long)L + fl - last : (long long)L;
  }
}

// Two anneals per wave for D <= 32 (P = 2, first accept, Philox streams): lanes 0-31 run anneal r of the frame, lanes
// 32-63 anneal r + 1 (when `two`), each half lane for lane the one-anneal form above with coordinate j = lane & 31 and
// its own Philox stream word r + half; results equal arpeggio_anneal's for both anneals. The scan commits one flip in
// each half per step: the two halves' next accepted lanes come from the two 32-bit words of one ballot, each half's
// lanes read their own Gram column, and a half with no flip left reads the zero padding past the Gram (column D), so
// its fields stay. A pass takes max(flips of the two) steps, and the per-pass work (Philox block, thresholds, the
// sign update, the best-state test) is shared by the two anneals.
template <bool FULL, int DS = 0>
__device__ __forceinline__ void arpeggio_anneal_pair(const int32_t* __restrict__ G4, const int32_t* __restrict__ hb,
                                                     const int32_t* __restrict__ T, int L, int q, int D, uint32_t r,
                                                     bool two, uint32_t inst, uint32_t k0, uint32_t k1, int start,
                                                     int8_t* __restrict__ out_s, long long* __restrict__ out_de,
                                                     long long* __restrict__ out_st, size_t t,
                                                     const uint32_t* __restrict__ thr, int gate,
                                                     int8_t* __restrict__ used) {
  constexpr int P = 2;
  constexpr uint64_t LO = 0xFFFFFFFFull, HI = ~0xFFFFFFFFull;
  if (DS) D = DS;
  const int lane = lane_id(), hf = lane >> 5, j = lane & 31;
  const uint32_t rr = r + (uint32_t)hf;
  const bool valid = j < D && (hf == 0 || two);
  const uint64_t vm = ballot(valid);
  const int32_t dj2 = valid ? G4[j * D + j] >> 1 : 0;
  uint64_t Sm[P], Rm[P];
  if (thr) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      uint32_t x = 0u, th = 0u;
      if (valid) {
        x = philox_w0(0xFFFFFFFFu, (uint32_t)(p * D + j), rr, inst, k0, k1);
        th = thr[p * D + j];
      }
      Sm[p] = ballot(valid && (uint64_t)x + th >= 0x100000000ull);
      Rm[p] = ballot(valid && x >= 0x80000000u);
    }
  } else if (start == 0) {
#pragma unroll
    for (int p = 0; p < P; ++p)
      Sm[p] = ballot(valid && philox_w0(0xFFFFFFFFu, (uint32_t)(p * D + j), rr, inst, k0, k1) >= 0x80000000u);
  } else {
    const uint32_t x = valid ? philox_w0(0xFFFFFFFFu, (uint32_t)j, rr, inst, k0, k1) : 0u;
    const uint32_t bits = bal_pattern(P, (uint32_t)(((uint64_t)x * bal_count(P)) >> 32));
#pragma unroll
    for (int p = 0; p < P; ++p) Sm[p] = ballot(valid && ((bits >> p) & 1u));
  }
  // fields w_j = h_j - sum_i G4_ij (m_i / 2): per half the columns with m_i / 2 = +1 (both spins +1) or -1 (both -1)
  auto fields = [&](const uint64_t (&S)[P]) {
    int32_t w = valid ? hb[j] : 0;
    const uint64_t plus = S[0] & S[1];
    uint64_t nz = plus | (~(S[0] | S[1]) & vm);
    while (nz) {
      const uint32_t lo = (uint32_t)nz, hi = (uint32_t)(nz >> 32);
      const int l0 = lo ? __builtin_ctz(lo) : D, l1 = hi ? __builtin_ctz(hi) : D;
      const int32_t s0 = lo && ((plus >> l0) & 1ull) ? -1 : 0, s1 = hi && ((plus >> (32 + l1)) & 1ull) ? -1 : 0;
      if (lo) nz &= ~(1ull << l0);
      if (hi) nz &= ~(1ull << (32 + l1));
      const int lc = hf ? l1 : l0;
      const int32_t sc = hf ? s1 : s0;  // -1: w -= col, 0: w += col
      w += (G4[lc * D + j] ^ sc) - sc;
    }
    return w;
  };
  // start energy up to a constant, per half: -(h + w).m summed over the half's lanes (every lane holds its half's)
  auto energy = [&](const uint64_t (&S)[P], int32_t w) {
    const int m = 2 * ((int)((S[0] >> lane) & 1ull) + (int)((S[1] >> lane) & 1ull)) - P;
    long long e = valid ? -(long long)(hb[j] + w) * m : 0;
#pragma unroll
    for (int o = 16; o; o >>= 1) e += __shfl_xor(e, o);
    return e;
  };
  int32_t w = fields(Sm);
  if (thr) {
    bool take = true;
    if (gate) {
      const int32_t wr = fields(Rm);
      take = energy(Sm, w) <= energy(Rm, wr);
      const uint64_t keep = ballot(take);
#pragma unroll
      for (int p = 0; p < P; ++p) Sm[p] = (Sm[p] & keep) | (Rm[p] & ~keep);
      w = take ? w : wr;
    }
    if (used && j == 0 && (hf == 0 || two)) used[t + hf] = take ? 1 : 0;
  }
  int32_t sgp[P], Bs[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    sgp[p] = __builtin_amdgcn_inverse_ballot_w64(Sm[p]) ? 0 : -1;
    Bs[p] = sgp[p];
  }
  long long de0 = 0, de1 = 0, best0 = 1LL << 60, best1 = 1LL << 60;  // per half, de / 4
  int32_t fc = 0;
  const PhiloxKeys K = philox_keys_v(k0, k1);
  uint32_t rw[P][4];
  for (int kb = 0; kb < L; kb += 4 * P) {
    const int n4 = kb / (4 * P);
    int32_t Tb[4 * P];
#pragma unroll
    for (int i = 0; i < 4 * P; ++i) Tb[i] = (FULL || kb + i < L) ? T[kb + i] : 0;
#pragma unroll
    for (int nm = 0; nm < 4; ++nm) {
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int k = kb + nm * P + p;
        if (!FULL && k >= L) break;
        if (nm == 0) {
          uint32_t c0 = (uint32_t)n4, c1 = (uint32_t)(p * D + j), c2 = rr, c3 = inst;
          philox10_kv(c0, c1, c2, c3, K);
          rw[p][0] = c0;
          rw[p][1] = c1;
          rw[p][2] = c2;
          rw[p][3] = c3;
        }
        const int32_t Tk = Tb[nm * P + p];
        const uint32_t A = 4u * (uint32_t)Tk;
        const int32_t Bk = q + 2 * Tk;
        const int32_t tq = (int32_t)(((uint64_t)rw[p][nm] * A + ~0ull) >> 32) - Bk;
        const int32_t sg = sgp[p], e2 = dj2 - sg;
        uint64_t rem = vm, fm = 0;
        int32_t qv = (w ^ sg) + e2;
        uint64_t m = ballot(qv <= tq) & rem;
        while (m) {
          const uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
          int l0 = D, l1 = D;
          int32_t sl0 = 0, sl1 = 0;
          uint64_t rn = 0;
          if (lo) {
            l0 = __builtin_ctz(lo);
            de0 += readlane(qv, l0);
            sl0 = readlane(sg, l0);
            mask_set(fm, l0);
            rn = ((~1ull) << l0) & LO;
          }
          if (hi) {
            l1 = __builtin_ctz(hi);
            de1 += readlane(qv, 32 + l1);
            sl1 = readlane(sg, 32 + l1);
            mask_set(fm, 32 + l1);
            rn |= (~1ull) << (32 + l1);
          }
          rem &= rn;
          const int lc = hf ? l1 : l0;
          const int32_t sl = hf ? sl1 : sl0;
          w = ((G4[lc * D + j] ^ sl) + w) - sl;
          qv = (w ^ sg) + e2;
          m = ballot(qv <= tq) & rem;
        }
        if (fm) {
          asm volatile("");
          const bool fl = __builtin_amdgcn_inverse_ballot_w64(fm);
          sgp[p] ^= fl ? -1 : 0;
          fc += fl ? 1 : 0;
        }
        const bool lt0 = s64_lt(de0, best0), lt1 = s64_lt(de1, best1);
        if (lt0 || lt1) {
          asm volatile("");
          if (lt0) best0 = de0;
          if (lt1) best1 = de1;
          const bool u = __builtin_amdgcn_inverse_ballot_w64((lt0 ? LO : 0ull) | (lt1 ? HI : 0ull));
#pragma unroll
          for (int pp = 0; pp < P; ++pp) Bs[pp] = u ? sgp[pp] : Bs[pp];
        }
      }
    }
  }
  const size_t th = t + (size_t)hf;
  if (valid) {
#pragma unroll
    for (int p = 0; p < P; ++p) out_s[th * (size_t)(P * D) + p * D + j] = Bs[p] ? -1 : 1;
  }
  int fl = fc;
#pragma unroll
  for (int o = 16; o; o >>= 1) fl += __shfl_xor(fl, o);
  const int last0 = readlane(fc, D - 1), last1 = readlane(fc, 32 + D - 1);
  if (j == 0 && (hf == 0 || two)) {
    out_de[th] = 4 * (hf ? best1 : best0);
    out_st[2 * th] = fl;
    out_st[2 * th + 1] = (long long)L + fl - (hf ? last1 : last0);
  }
}

// grid.x = B frames (or list slots), block = 64 W threads. LDS: 4 Gt, D*D int32 (+ one wave of padding so that the
// lanes past D of the last column stay inside the allocation). With a frame list, workgroup i anneals frame list[i]
// if i < *count and exits otherwise (a uniform exit before any barrier); its outputs go to slot i. Anneal rr of a
// workgroup uses Philox stream r0 + rr, so a later launch with r0 = R continues a frame's anneals with new streams.
// Stages frame b's Gram into LDS as 4 Gt (from the int16 Gh when given, else the int32 Gt), with CPL waves of zero
// padding past it.
template <int CPL>
__device__ __forceinline__ void stage_gram(int32_t* __restrict__ sG, const int32_t* __restrict__ Gt,
                                           const int16_t* __restrict__ Gh, int b, int D) {
  const int nG = D * D;
  if (Gh) {
    const int16_t* g = Gh + (size_t)b * nG;
    if ((nG & 3) == 0) {  // four entries per 8-byte load (frames start on 8-byte boundaries)
      for (int x4 = threadIdx.x; x4 < nG / 4; x4 += blockDim.x) {
        const uint2 w = reinterpret_cast<const uint2*>(g)[x4];
        reinterpret_cast<int4*>(sG)[x4] = make_int4(4 * (int32_t)(int16_t)(w.x & 0xFFFFu), 4 * ((int32_t)w.x >> 16),
                                                    4 * (int32_t)(int16_t)(w.y & 0xFFFFu), 4 * ((int32_t)w.y >> 16));
      }
      for (int x = nG + threadIdx.x; x < nG + ARP_WAVE * CPL; x += blockDim.x) sG[x] = 0;
    } else {
      for (int x = threadIdx.x; x < nG + ARP_WAVE * CPL; x += blockDim.x) sG[x] = x < nG ? 4 * (int32_t)g[x] : 0;
    }
  } else {
    const int32_t* g = Gt + (size_t)b * nG;
    for (int x = threadIdx.x; x < nG + ARP_WAVE * CPL; x += blockDim.x) sG[x] = x < nG ? 4 * g[x] : 0;
  }
}

// D = 64: eight waves per SIMD (at most 64 registers), against seven at 66.
template <int P, int CPL, int MODE, bool FULL, int DS, int RNG>
__global__ __launch_bounds__(ARP_ARPEGGIO_MAX_THREADS) __attribute__((amdgpu_waves_per_eu(DS == 64 ? 8 : 1))) void arpeggio_fx_kernel(const int32_t* __restrict__ Gt,
                                                                               const int32_t* __restrict__ h,
                                                                               const int32_t* __restrict__ T, int L,
                                                                               int q, int R, int D,
                                                                               const uint32_t* __restrict__ inst,
                                                                               uint32_t k0, uint32_t k1, int start,
                                                                               int8_t* __restrict__ out_s,
                                                                               long long* __restrict__ out_de,
                                                                               long long* __restrict__ out_st,
                                                                               AnnealExtra ex) {
  extern __shared__ int32_t sG[];
  if (DS) D = DS;
  const int slot = blockIdx.x;
  int b = slot;
  if (ex.list) {
    if (slot >= *ex.count) return;
    b = ex.list[slot];
  }
  stage_gram<CPL>(sG, Gt, ex.Gh, b, D);
  __syncthreads();
  // the wave index as a wave-uniform value: anneal index and output row in scalar registers
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), W = blockDim.x >> 6;
  const uint32_t in = inst[b];
  const uint32_t* thr = ex.thr ? ex.thr + (size_t)b * P * D : nullptr;
  for (int r = wv; r < R; r += W)
    arpeggio_anneal<P, CPL, MODE, FULL, DS, RNG>(sG, h + (size_t)b * D, T, L, q, D, (uint32_t)(ex.r0 + r), in, k0, k1,
                                                 start, out_s, out_de, out_st, (size_t)slot * R + r, thr, ex.gate,
                                                 ex.used, ex.C);
}

// The same launch with two anneals per wave (D <= 32, P = 2, first accept): wave w runs the pairs (2i, 2i + 1),
// i = w, w + W, ...; an odd R leaves the last pair's upper half idle.
template <bool FULL, int DS>
__global__ __launch_bounds__(ARP_ARPEGGIO_MAX_THREADS) void arpeggio_pair_kernel(const int32_t* __restrict__ Gt,
                                                                                 const int32_t* __restrict__ h,
                                                                                 const int32_t* __restrict__ T, int L,
                                                                                 int q, int R, int D,
                                                                                 const uint32_t* __restrict__ inst,
                                                                                 uint32_t k0, uint32_t k1, int start,
                                                                                 int8_t* __restrict__ out_s,
                                                                                 long long* __restrict__ out_de,
                                                                                 long long* __restrict__ out_st,
                                                                                 AnnealExtra ex) {
  extern __shared__ int32_t sG[];
  if (DS) D = DS;
  const int slot = blockIdx.x;
  int b = slot;
  if (ex.list) {
    if (slot >= *ex.count) return;
    b = ex.list[slot];
  }
  stage_gram<1>(sG, Gt, ex.Gh, b, D);
  __syncthreads();
  const int wv = threadIdx.x >> 6, W = blockDim.x >> 6;
  const uint32_t in = inst[b];
  const uint32_t* thr = ex.thr ? ex.thr + (size_t)b * 2 * D : nullptr;
  for (int r = 2 * wv; r < R; r += 2 * W)
    arpeggio_anneal_pair<FULL, DS>(sG, h + (size_t)b * D, T, L, q, D, (uint32_t)(ex.r0 + r), r + 1 < R, in, k0, k1,
                                   start, out_s, out_de, out_st, (size_t)slot * R + r, thr, ex.gate, ex.used);
}

template <int P, int CPL, int MODE, bool FULL, int DS, int RNG = 0>
static void launch_pcmf(dim3 grid, dim3 block, size_t lds, hipStream_t st, const int32_t* Gt, const int32_t* h,
                        const int32_t* T, int L, int q, int R, int D, const uint32_t* inst, uint32_t k0, uint32_t k1,
                        int start, int8_t* out_s, long long* out_de, long long* out_st, const AnnealExtra& ex) {
  if (lds > 65536)
    hipFuncSetAttribute((const void*)arpeggio_fx_kernel<P, CPL, MODE, FULL, DS, RNG>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((arpeggio_fx_kernel<P, CPL, MODE, FULL, DS, RNG>), grid, block, lds, st, Gt, h, T, L, q, R, D,
                     inst, k0, k1, start, out_s, out_de, out_st, ex);
}

// FULL: L is a whole number of 4 P-pass Philox blocks, so the pass loop needs no bound test
template <int P, int CPL, int MODE>
static void launch_pcm(dim3 grid, dim3 block, size_t lds, hipStream_t st, const int32_t* Gt, const int32_t* h,
                       const int32_t* T, int L, int q, int R, int D, const uint32_t* inst, uint32_t k0, uint32_t k1,
                       int start, int8_t* out_s, long long* out_de, long long* out_st, const AnnealExtra& ex) {
  // 16, 32 and 64 users (D = 32, 64, 128; two copies, first accept): D as a compile-time constant
  constexpr bool STATIC = P == 2 && MODE == 0;
  constexpr int DSA = STATIC ? (CPL == 1 ? 64 : 128) : 0, DSB = STATIC && CPL == 1 ? 32 : 0;
  const bool full = L % (4 * P) == 0;
#define ARP_PCMF(F, DSV) \
  launch_pcmf<P, CPL, MODE, F, DSV>(grid, block, lds, st, Gt, h, T, L, q, R, D, inst, k0, k1, start, out_s, out_de, \
                                    out_st, ex)
  if (STATIC && (D == DSA || D == DSB)) {
    if (D == DSA && full) ARP_PCMF(true, DSA);
    else if (D == DSA) ARP_PCMF(false, DSA);
    else if (full) ARP_PCMF(true, DSB);
    else ARP_PCMF(false, DSB);
    return;
  }
  if (full) ARP_PCMF(true, 0);
  else ARP_PCMF(false, 0);
#undef ARP_PCMF
}

template <int P>
static hipError_t launch_p(const int32_t* Gt, const int32_t* h, const int32_t* T, int L, int q, int R, int D, int B,
                           const uint32_t* inst, uint32_t k0, uint32_t k1, int start, int mode, int8_t* out_s,
                           long long* out_de, long long* out_st, int waves, const AnnealExtra& ex, hipStream_t st) {
  const int cpl = D <= ARP_WAVE ? 1 : 2;
  const size_t lds = sizeof(int32_t) * ((size_t)D * D + ARP_WAVE * cpl);
  dim3 grid(B), block(ARP_WAVE * waves);
  if (ex.rng || ex.C) {
    // the production engine's lane-shared Chord: the parallel pass with the Onsager term, xorshift lane streams
    // (RNG 1) or Philox (RNG 2)
    const bool full = L % (4 * P) == 0;
#define ARP_LANE(CPLV, F, RV) \
  launch_pcmf<P, CPLV, 1, F, 0, RV>(grid, block, lds, st, Gt, h, T, L, q, R, D, inst, k0, k1, start, out_s, out_de, \
                                    out_st, ex)
    if (ex.rng) {
      if (cpl == 1 && full) ARP_LANE(1, true, 1);
      else if (cpl == 1) ARP_LANE(1, false, 1);
      else if (full) ARP_LANE(2, true, 1);
      else ARP_LANE(2, false, 1);
    } else {
      if (cpl == 1 && full) ARP_LANE(1, true, 2);
      else if (cpl == 1) ARP_LANE(1, false, 2);
      else if (full) ARP_LANE(2, true, 2);
      else ARP_LANE(2, false, 2);
    }
#undef ARP_LANE
    return hipGetLastError();
  }
  if (P == 2 && mode == 0 && D <= 32 && anneals_per_wave(D) == 2) {
    // two anneals per wave: half the waves per workgroup for the same anneals
    const int pw = (R + 1) / 2 < waves ? (R + 1) / 2 : waves;
    const dim3 pblock(ARP_WAVE * pw);
    const bool full = L % 8 == 0;
#define ARP_PAIR(F, DSV)                                                                                             \
  hipLaunchKernelGGL((arpeggio_pair_kernel<F, DSV>), grid, pblock, lds, st, Gt, h, T, L, q, R, D, inst, k0, k1, start, \
                     out_s, out_de, out_st, ex)
    if (D == 32 && full) ARP_PAIR(true, 32);
    else if (D == 32) ARP_PAIR(false, 32);
    else if (full) ARP_PAIR(true, 0);
    else ARP_PAIR(false, 0);
#undef ARP_PAIR
    return hipGetLastError();
  }
  if (cpl == 1) {
    if (mode == 0) launch_pcm<P, 1, 0>(grid, block, lds, st, Gt, h, T, L, q, R, D, inst, k0, k1, start, out_s, out_de, out_st, ex);
    else launch_pcm<P, 1, 1>(grid, block, lds, st, Gt, h, T, L, q, R, D, inst, k0, k1, start, out_s, out_de, out_st, ex);
  } else {
    if (mode == 0) launch_pcm<P, 2, 0>(grid, block, lds, st, Gt, h, T, L, q, R, D, inst, k0, k1, start, out_s, out_de, out_st, ex);
    else launch_pcm<P, 2, 1>(grid, block, lds, st, Gt, h, T, L, q, R, D, inst, k0, k1, start, out_s, out_de, out_st, ex);
  }
  return hipGetLastError();
}

// Persistent mixed-size launch (P = 2): `wg` workgroups of `waves` waves pull tiles from a device counter in the
// global tile order (the caller puts the largest frames first); a tile's frames are staged side by side in LDS, each
// as 4 Gt plus one wave of padding per 64 coordinates, and the waves run the tile's nf x R anneals. Every anneal is
// the same function with the same arguments as in arpeggio_fx_kernel, so results equal the per-class launches.
// The counter is cleared by arpeggio_batch_reset_kernel on the same stream before each launch.
__global__ void arpeggio_batch_reset_kernel(unsigned* counter) { *counter = 0u; }

__global__ __launch_bounds__(ARP_BATCHED_MAX_THREADS) void arpeggio_batched_kernel(ArpBatchArgs a, unsigned* __restrict__ counter) {
  extern __shared__ int32_t sG[];
  __shared__ int s_tile;
  const int W = blockDim.x >> 6, wv = threadIdx.x >> 6;
  const int R = a.R;
  while (true) {
    if (threadIdx.x == 0) s_tile = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int t = s_tile;
    __syncthreads();  // every thread holds t before the next tile's fetch may overwrite s_tile
    if (t >= a.n_tiles) break;
    int ci = 0;
    for (int i = 1; i < a.n_classes; ++i)
      if (t >= a.c[i].first_tile) ci = i;
    const ArpBatchClass& C = a.c[ci];
    const int D = C.D, nG = D * D, cpl = D <= ARP_WAVE ? 1 : 2;
    const int stride = nG + ARP_WAVE * cpl;
    const int f0 = (t - C.first_tile) * C.tile;
    const int nf = C.B - f0 < C.tile ? C.B - f0 : C.tile;
    const int32_t* g = C.Gt + (size_t)f0 * nG;
    for (int x = threadIdx.x; x < nf * stride; x += blockDim.x) {
      const int f = x / stride, e = x - f * stride;
      sG[x] = e < nG ? 4 * g[(size_t)f * nG + e] : 0;
    }
    __syncthreads();
    const bool full = C.L % 8 == 0;
    const bool pair = C.apw == 2 && D <= 32;  // two anneals per wave: units are pairs (r, r + 1)
    const int RU = pair ? (R + 1) / 2 : R;
    for (int u = wv; u < nf * RU; u += W) {
      const int f = u / RU, r = (u - f * RU) * (pair ? 2 : 1), b = f0 + f;
      const int32_t* gs = sG + f * stride;
      const int32_t* hb = C.h + (size_t)b * D;
      const size_t at = (size_t)b * R + r;
      const uint32_t* thr = C.thr ? C.thr + (size_t)b * 2 * D : nullptr;
      if (pair) {
        if (full && D == 32)
          arpeggio_anneal_pair<true, 32>(gs, hb, C.T, C.L, C.q, D, (uint32_t)r, r + 1 < R, C.inst[b], a.k0, a.k1,
                                         C.start, C.out_s, C.out_de, C.out_st, at, thr, C.gate, C.used);
        else if (full)
          arpeggio_anneal_pair<true>(gs, hb, C.T, C.L, C.q, D, (uint32_t)r, r + 1 < R, C.inst[b], a.k0, a.k1, C.start,
                                     C.out_s, C.out_de, C.out_st, at, thr, C.gate, C.used);
        else
          arpeggio_anneal_pair<false>(gs, hb, C.T, C.L, C.q, D, (uint32_t)r, r + 1 < R, C.inst[b], a.k0, a.k1,
                                      C.start, C.out_s, C.out_de, C.out_st, at, thr, C.gate, C.used);
      } else if (cpl == 1) {
        if (full && D == 64)
          arpeggio_anneal<2, 1, 0, true, 64>(gs, hb, C.T, C.L, C.q, D, (uint32_t)r, C.inst[b], a.k0, a.k1, C.start,
                                             C.out_s, C.out_de, C.out_st, at, thr, C.gate, C.used);
        else if (full && D == 32)
          arpeggio_anneal<2, 1, 0, true, 32>(gs, hb, C.T, C.L, C.q, D, (uint32_t)r, C.inst[b], a.k0, a.k1, C.start,
                                             C.out_s, C.out_de, C.out_st, at, thr, C.gate, C.used);
        else if (full)
          arpeggio_anneal<2, 1, 0, true>(gs, hb, C.T, C.L, C.q, D, (uint32_t)r, C.inst[b], a.k0, a.k1, C.start,
                                         C.out_s, C.out_de, C.out_st, at, thr, C.gate, C.used);
        else
          arpeggio_anneal<2, 1, 0, false>(gs, hb, C.T, C.L, C.q, D, (uint32_t)r, C.inst[b], a.k0, a.k1, C.start,
                                          C.out_s, C.out_de, C.out_st, at, thr, C.gate, C.used);
      } else {
        if (full && D == 128)
          arpeggio_anneal<2, 2, 0, true, 128>(gs, hb, C.T, C.L, C.q, D, (uint32_t)r, C.inst[b], a.k0, a.k1, C.start,
                                              C.out_s, C.out_de, C.out_st, at, thr, C.gate, C.used);
        else if (full)
          arpeggio_anneal<2, 2, 0, true>(gs, hb, C.T, C.L, C.q, D, (uint32_t)r, C.inst[b], a.k0, a.k1, C.start,
                                         C.out_s, C.out_de, C.out_st, at, thr, C.gate, C.used);
        else
          arpeggio_anneal<2, 2, 0, false>(gs, hb, C.T, C.L, C.q, D, (uint32_t)r, C.inst[b], a.k0, a.k1, C.start,
                                          C.out_s, C.out_de, C.out_st, at, thr, C.gate, C.used);
      }
    }
    __syncthreads();  // the tile's LDS is free again
  }
}

hipError_t launch_arpeggio_batched(const ArpBatchArgs& a, unsigned* counter, int wg, int waves, size_t lds,
                                   hipStream_t st) {
  if (lds > 65536) {
    const hipError_t e =
        hipFuncSetAttribute((const void*)arpeggio_batched_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(arpeggio_batch_reset_kernel, dim3(1), dim3(1), 0, st, counter);
  hipLaunchKernelGGL(arpeggio_batched_kernel, dim3(wg), dim3(ARP_WAVE * waves), lds, st, a, counter);
  return hipGetLastError();
}

hipError_t launch_arpeggio_ex(const int32_t* Gt, const int32_t* h, const int32_t* T, int L, int q, int P, int R, int D,
                              int B, const uint32_t* inst, uint32_t k0, uint32_t k1, int start, int mode,
                              int8_t* out_s, long long* out_de, long long* out_st, int waves, const AnnealExtra& ex,
                              hipStream_t st) {
  switch (P) {
    case 2: return launch_p<2>(Gt, h, T, L, q, R, D, B, inst, k0, k1, start, mode, out_s, out_de, out_st, waves, ex, st);
    case 4: return launch_p<4>(Gt, h, T, L, q, R, D, B, inst, k0, k1, start, mode, out_s, out_de, out_st, waves, ex, st);
    default: return launch_p<6>(Gt, h, T, L, q, R, D, B, inst, k0, k1, start, mode, out_s, out_de, out_st, waves, ex, st);
  }
}

hipError_t launch_arpeggio_fx(const int32_t* Gt, const int32_t* h, const int32_t* T, int L, int q, int P, int R, int D,
                              int B, const uint32_t* inst, uint32_t k0, uint32_t k1, int start, int mode,
                              int8_t* out_s, long long* out_de, long long* out_st, int waves, hipStream_t st,
                              const uint32_t* thr, int gate, int8_t* used) {
  AnnealExtra ex{thr, used, nullptr, nullptr, gate, 0};
  return launch_arpeggio_ex(Gt, h, T, L, q, P, R, D, B, inst, k0, k1, start, mode, out_s, out_de, out_st, waves, ex,
                            st);
}

}  // namespace arp
