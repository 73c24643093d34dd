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
  uint16_t* feat;  // (B, feat_ld) bf16 learned-start features: H^T y / g, G_jj / g - 1, x_g / lmax, z/mu - x_g,
                   // (N0/2) / g, zero padding; g = mean_j G_jj
  int feat_ld;
  int16_t* Gh;     // (B, D, D) fixed-point Gram as int16 (|G_int| <= Q <= 2^14 fits): Gi at half the bytes
};

// One class of a mixed batch for the persistent Arpeggio launch: frames of one size D with their own buffers (the
// layouts of launch_arpeggio_fx) and schedule; `tile` frames form one unit of work, and the class's tiles are
// numbered from first_tile in the launch's global tile order.
struct ArpBatchClass {
  const int32_t* Gt;     // (B, D, D)
  const int32_t* h;      // (B, D)
  const uint32_t* inst;  // (B,)
  const int32_t* T;      // (L,)
  int8_t* out_s;         // (B, R, P D)
  long long* out_de;     // (B, R)
  long long* out_st;     // (B, R, 2)
  int D, B, L, q, start, tile, first_tile, gate;
  const uint32_t* thr;   // (B, 2 D) learned-start thresholds or nullptr
  int8_t* used;          // (B, R) kept flags of the energy gate or nullptr
  int apw;               // anneals per wave (2: pairs, for D <= 32)
};
#define ARP_MAX_CLASSES 8
struct ArpBatchArgs {
  ArpBatchClass c[ARP_MAX_CLASSES];
  int n_classes, n_tiles, R, pad_;
  uint32_t k0, k1;
};

// Anneals per wave64 of the first-accept launches (two copies) for frames of D coordinates: 2 packs two anneals into
// a wave (D <= 32 only), 1 runs one; set per host thread by arp_set_anneals_per_wave.
int anneals_per_wave(int D);

// number of learned-start features for D coordinates (before padding)
__host__ __device__ inline int feat_count(int D) { return 4 * D + 1; }

hipError_t launch_arpeggio_fx(const int32_t* Gt, const int32_t* h, const int32_t* T, int L, int q, int P, int R, int D,
                              int B, const uint32_t* inst, uint32_t k0, uint32_t k1, int start, int mode,
                              int8_t* out_s, long long* out_de, long long* out_st, int waves, hipStream_t st,
                              const uint32_t* thr = nullptr, int gate = 0, int8_t* used = nullptr);
hipError_t launch_arpeggio_batched(const ArpBatchArgs& a, unsigned* counter, int wg, int waves, size_t lds,
                                   hipStream_t st);
hipError_t launch_chord_fx(const int16_t* J, const int32_t* h, const int32_t* h0, const int32_t* Tq,
                           const int32_t* Cq, int q, int S, int P, int R, int N, int B, uint32_t k0, uint32_t k1,
                           int inst0, int mode, int32_t* out_e, int8_t* out_s, unsigned long long* stats, int waves,
                           size_t max_lds, hipStream_t st);
// Front end: the matrix-core block sweep (front_end_mfma.hip) for D, Mr <= 64 without the expanded Chord inputs, the
// register sweep (front_end.hip) otherwise; set_front_end_form(1) forces the register sweep everywhere.
hipError_t launch_front_end(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M,
                            int K, float Q, const FrontOut& o, hipStream_t st);
hipError_t launch_front_end_sweep(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D,
                                  int M, int K, float Q, const FrontOut& o, hipStream_t st);
void set_front_end_form(int form);
int front_end_form();  // 2: the frames' H are real lifts of complex channels (front end and decodes read half of H)
// The MDI sub-problem build at 64 users (D = Mr = 128) on the complex-lift front end with the decisions given; false
// (nothing launched) for a Q that is not a power of two or a misaligned H.
bool mdi_build_cpx128(const float* H, const float* y, const int8_t* xg, const int32_t* list, int n, float Q,
                      int32_t* Gi, int32_t* hi, hipStream_t st);
// The complex-lift front end (front_end_cpx.hip) for H = [[Re, -Im], [Im, Re]], D = Mr in {32, 64}; false (nothing
// launched) for other shapes, the expanded Chord inputs, a Q that is not a power of two, or a misaligned H.
bool front_end_cpx(const float* H, const float* y, const float* N0v, float N0s, int B, int Mr, int D, int M, int K,
                   float Q, const FrontOut& o, hipStream_t st);
// The MDI sub-problem build (Gi, hi around a given decision xg for the frames of a list) on the matrix-core front end;
// false (nothing launched) for D or Mr > 64 or a Q that is not a power of two.
bool mdi_build_mfma(const float* H, const float* y, const int8_t* xg, const int32_t* list, int n, int Mr, int D,
                    float Q, int32_t* Gi, int32_t* hi, hipStream_t st);
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

hipError_t launch_gemm_bf16(const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, int M, int N,
                            int K, const float* bias, int act, float* Cf, float* Cpre, uint16_t* Cb, long ldc,
                            hipStream_t st);
hipError_t launch_gelu_bwd(const float* dH, const float* Apre, long n, uint16_t* dAb, float* dAf, hipStream_t st);
hipError_t launch_bce(const float* z, long ld, const int8_t* y, int M, int N, uint16_t* dzb, float* dzf,
                      float* loss_acc, hipStream_t st);
hipError_t launch_colsum(const float* X, long ld, int M, int N, float* out, hipStream_t st);
hipError_t launch_adam(float* W, const float* G, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                       int step, uint16_t* Wb, hipStream_t st);
hipError_t launch_cast_bf16(const float* x, long n, uint16_t* y, hipStream_t st);
hipError_t launch_mlp_start_fused(const uint16_t* X, int M, int F, const uint16_t* W1t, const float* b1,
                                  const uint16_t* W2t, const float* b2, const uint16_t* W3t, const float* b3, int N,
                                  float inv_temp, uint32_t* thr, hipStream_t st);
hipError_t launch_start_thr(const float* z, long ld, int M, int N, float inv_temp, uint32_t* thr, hipStream_t st);

}  // namespace arp
