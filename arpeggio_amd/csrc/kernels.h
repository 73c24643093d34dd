// Launch wrappers shared by the kernel translation units and the C ABI (capi.hip).
#pragma once
#include "common.h"

namespace arp {

struct FrontOut {
  int8_t* xg;      // (B, D) MMSE hard levels
  int32_t* Gi;     // (B, D, D) fixed-point Gram (symmetric; also its own transpose)
  int32_t* hi;     // (B, D)
  float* scale;    // (B,)
  float* Gf;       // (B, D, D) G / scale
  float* hf;       // (B, D)   h / scale
  int16_t* J;      // (B, N, N) expanded dDI couplings, N = 2 K D
  int32_t* hN;     // (B, N)
  float* z;        // (B, D) LMMSE output
  float* mu;       // (B, D)
  float* Jf;       // (B, N, N) expanded dDI couplings, float (G / scale)
  float* hNf;      // (B, N)
};

hipError_t launch_arpeggio_fx(const int32_t* Gt, const int32_t* h, const int32_t* T, int L, int q, int P, int R, int D,
                              int B, const uint32_t* inst, uint32_t k0, uint32_t k1, int start, int mode,
                              int8_t* out_s, long long* out_de, long long* out_st, int waves, hipStream_t st);
hipError_t launch_chord_fx(const int16_t* J, const int32_t* h, const int32_t* h0, const int32_t* Tq,
                           const int32_t* Cq, int q, int S, int P, int R, int N, int B, uint32_t k0, uint32_t k1,
                           int inst0, int mode, int32_t* out_e, int8_t* out_s, unsigned long long* stats, int waves,
                           size_t max_lds, hipStream_t st);
hipError_t launch_front_end(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M,
                            int K, float Q, const FrontOut& o, hipStream_t st);
hipError_t launch_decode_select(const float* H, const float* y, const int8_t* xg, const int8_t* spins, int B, int R,
                                int P, int D, int Mr, int M, int delta, const int8_t* ltrue, int8_t* lhat,
                                int32_t* biterr, float* dist_out, hipStream_t st);

hipError_t launch_arpeggio_f32(const float* Gt, const float* h, const float* T, int L, float q, int P, int R, int D,
                               int B, const uint32_t* inst, uint32_t k0, uint32_t k1, int start, int8_t* out_s,
                               float* out_de, long long* out_st, float* out_w, int waves, hipStream_t st);
hipError_t launch_chord_f32(const float* J, const float* h, const float* h0, const float* T, const float* c, float q,
                            int S, int P, int R, int N, int B, uint32_t k0, uint32_t k1, int inst0, float* out_e,
                            int8_t* out_s, float* out_u, long long* out_fl, int waves, size_t max_lds,
                            hipStream_t st);

}  // namespace arp
