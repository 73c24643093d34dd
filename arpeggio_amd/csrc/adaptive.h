// Optional arguments of the first-accept and decode kernels for frame-list (compacted) launches, and the launch
// wrappers of the early-stop path (certificate + compaction). Included by the kernels that take them and by the C ABI.
#pragma once
#include "../../include/arpeggio_hip.h"
#include "common.h"

namespace arp {

// Optional arguments of arpeggio_fx_kernel: learned-start thresholds (B, P D) with the energy gate and the kept flags,
// a frame list (slot i anneals frame list[i] for i < *count), the first stream index r0, an Onsager term C per pass
// (subtracted from the decision bound), and rng 1 for the production xorshift lane streams (parallel pass only).
struct AnnealExtra {
  const uint32_t* thr;
  int8_t* used;
  const int32_t* list;
  const int32_t* count;
  int gate;
  int r0;
  const int32_t* C;
  int rng;
  const int16_t* Gh;  // the Gram as int16 in place of Gt (the front end's Gh output)
};

// Optional arguments of the decode kernels: a frame list (as AnnealExtra; spins are read from slot i, everything else
// from frame list[i]), a prior candidate per frame (levels and distance) that a new candidate must beat strictly, and
// the chosen candidate's distance out.
struct DecodeExtra {
  const int32_t* list;
  const int32_t* count;
  const int8_t* prior_l;
  const float* prior_d;
  float* best_d;
};

hipError_t launch_arpeggio_ex(const int32_t* Gt, const int32_t* h, const int32_t* T, int L, int q, int P, int R, int D,
                              int B, const uint32_t* inst, uint32_t k0, uint32_t k1, int start, int mode,
                              int8_t* out_s, long long* out_de, long long* out_st, int waves, const AnnealExtra& ex,
                              hipStream_t st);
hipError_t launch_decode_ex(const float* H, const float* y, const int8_t* xg, const int8_t* spins, int B, int R, int P,
                            int D, int Mr, int M, int delta, const int8_t* ltrue, int8_t* lhat, int32_t* biterr,
                            float* dist_out, const DecodeExtra& ex, hipStream_t st);
// The C ABI's error state (capi.hip): set_error(0, "") clears it and returns ARP_OK; hip_status maps a HIP error.
int set_error(int code, const char* msg);
int hip_status(hipError_t e, const char* what);

// The whole chain of arp_detect_adaptive on a caller's workspace (detect.hip).
int detect_chain(const float* H, const float* y, const float* N0, float N0_scalar, int B, int Mr, int D, int M,
                 const int32_t* T1, int L1, int R1, const int32_t* T2, int L2, int R2, int q, float thr_scale,
                 const uint32_t* inst, uint32_t k0, uint32_t k1, const arp_mlp_weights* mlp, int gate,
                 const int8_t* ltrue, int8_t* lhat, int32_t* biterr, void* work, size_t work_bytes, void* stream);

hipError_t launch_certify(const float* best_d, const float* N0v, float N0s, float thr_scale, int B, int32_t* list,
                          int32_t* count, hipStream_t st);

}  // namespace arp
