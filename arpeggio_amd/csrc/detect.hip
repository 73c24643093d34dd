// Whole detection chain in one C call: front end -> [warm-start MLP] -> stage-1 first-accept anneal -> decode ->
// [certificate -> stage-2 anneal of the open frames -> decode]. Every buffer between the stages is carved from one
// workspace, so a C caller runs the chain of the Python pipelines (arpeggio_amd.detector.DetectPipeline with R2 = 0,
// AdaptivePipeline otherwise) with a single call and no allocation of its own beyond that workspace.
#include <stdio.h>

#include "../../include/arpeggio_hip.h"
#include "adaptive.h"
#include "kernels.h"
#include "launch_bounds.h"

namespace arp {

struct DetectWork {
  int8_t* xg;
  int16_t* Gh;
  int32_t* hi;
  float* scale;
  uint16_t* feat;
  uint32_t* thr;
  int8_t* spins1;
  long long* de1;
  long long* st1;
  int8_t* used1;
  float* d1;
  float* d2;
  int32_t* list;
  int32_t* count;
  int8_t* spins2;
  long long* de2;
  long long* st2;
  int8_t* used2;
  size_t bytes;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Carves the workspace in a fixed order (each piece 256-byte aligned); base NULL only measures.
static DetectWork carve(uint8_t* base, int B, int D, int R1, int R2, int feat_ld, bool learned) {
  DetectWork w{};
  size_t off = 0;
  auto take = [&](size_t bytes) -> uint8_t* {
    uint8_t* p = base ? base + off : nullptr;
    off += align256(bytes);
    return p;
  };
  const size_t b = (size_t)B, N = 2 * (size_t)D;
  w.xg = (int8_t*)take(b * D);
  w.Gh = (int16_t*)take(b * D * D * 2);
  w.hi = (int32_t*)take(b * D * 4);
  w.scale = (float*)take(b * 4);
  if (learned) {
    w.feat = (uint16_t*)take(b * feat_ld * 2);
    w.thr = (uint32_t*)take(b * N * 4);
    w.used1 = (int8_t*)take(b * R1);
    if (R2 > 0) w.used2 = (int8_t*)take(b * R2);
  }
  w.spins1 = (int8_t*)take(b * R1 * N);
  w.de1 = (long long*)take(b * R1 * 8);
  w.st1 = (long long*)take(b * R1 * 16);
  w.d1 = (float*)take(b * 4);
  if (R2 > 0) {
    w.d2 = (float*)take(b * 4);
    w.list = (int32_t*)take(b * 4);
    w.count = (int32_t*)take(4);
    w.spins2 = (int8_t*)take(b * R2 * N);
    w.de2 = (long long*)take(b * R2 * 8);
    w.st2 = (long long*)take(b * R2 * 16);
  }
  w.bytes = off;
  return w;
}

static int feat_ld_of(int D) { return (feat_count(D) + 31) / 32 * 32; }

extern "C" size_t arp_detect_workspace_bytes(int B, int D, int R1, int R2, int learned) {
  if (B < 0 || D < 1 || D > 128 || R1 < 1 || R2 < 0) return 0;
  return carve(nullptr, B, D, R1, R2, feat_ld_of(D), learned != 0).bytes;
}

int detect_chain(const float* H, const float* y, const float* N0, float N0_scalar, int B, int Mr, int D, int M,
                 const int32_t* T1, int L1, int R1, const int32_t* T2, int L2, int R2, int q, float thr_scale,
                 const uint32_t* inst, uint32_t k0, uint32_t k1, const arp_mlp_weights* mlp, int gate,
                 const int8_t* ltrue, int8_t* lhat, int32_t* biterr, void* work, size_t work_bytes, void* stream) {
  char msg[256];
  if (D < 1 || D > 128 || Mr < 1 || Mr > 128) {
    snprintf(msg, sizeof msg, "detect: D = %d, Mr = %d outside 1..128", D, Mr);
    return set_error(ARP_E_RANGE, msg);
  }
  if (M != 4 && M != 16) return set_error(ARP_E_ARG, "detect: M must be 4 or 16");
  if (B < 0 || R1 < 1 || R1 > 256 || R2 < 0 || R2 > 256 || L1 < 1 || (R2 > 0 && L2 < 1)) {
    snprintf(msg, sizeof msg, "detect: B = %d, R1 = %d, L1 = %d, R2 = %d, L2 = %d", B, R1, L1, R2, L2);
    return set_error(ARP_E_ARG, msg);
  }
  if (B == 0) return set_error(0, "");
  if (!H || !y || !T1 || (R2 > 0 && !T2) || !inst || !lhat) return set_error(ARP_E_ARG, "detect: null pointer");
  if (biterr && !ltrue) return set_error(ARP_E_ARG, "detect: biterr needs ltrue");
  const bool learned = mlp != nullptr;
  if (learned && (mlp->H != 256 || mlp->F != feat_ld_of(D) || 2 * D % 64 != 0 || 2 * D > 256 || !mlp->W1t ||
                  !mlp->b1 || !mlp->W2t || !mlp->b2 || !mlp->W3t || !mlp->b3))
    return set_error(ARP_E_ARG, "detect: warm-start weights must be F = feature row, H = 256, N = 2D in "
                                "{64, 128, 192, 256}");
  const int fld = feat_ld_of(D);
  const size_t need = carve(nullptr, B, D, R1, R2, fld, learned).bytes;
  if (!work || work_bytes < need) {
    snprintf(msg, sizeof msg, "detect: workspace of %zu bytes, the shape needs %zu", work_bytes, need);
    return set_error(ARP_E_RANGE, msg);
  }
  const int waves1 = R1 < (D > 64 ? 8 : 4) ? R1 : (D > 64 ? 8 : 4);
  const int waves2 = R2 < (D > 64 ? 8 : 4) ? R2 : (D > 64 ? 8 : 4);
  if (ARP_WAVE * waves1 > ARP_ARPEGGIO_MAX_THREADS || ARP_WAVE * waves2 > ARP_ARPEGGIO_MAX_THREADS)
    return set_error(ARP_E_ARG, "detect: anneals per workgroup exceed the kernel's bound");
  const DetectWork w = carve((uint8_t*)work, B, D, R1, R2, fld, learned);
  hipStream_t st = (hipStream_t)stream;
  FrontOut o{w.xg, nullptr, w.hi, w.scale, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
             nullptr, nullptr, learned ? w.feat : nullptr, learned ? fld : 0, w.Gh};
  hipError_t e = launch_front_end(H, y, N0, N0_scalar, B, Mr, D, M, 1, 4096.f, o, st);
  if (e != hipSuccess) return hip_status(e, "detect: front end");
  if (learned) {
    e = launch_mlp_start_fused(w.feat, B, fld, mlp->W1t, mlp->b1, mlp->W2t, mlp->b2, mlp->W3t, mlp->b3, 2 * D,
                               mlp->inv_temp, w.thr, st);
    if (e != hipSuccess) return hip_status(e, "detect: warm-start MLP");
  }
  AnnealExtra ex1{w.thr, w.used1, nullptr, nullptr, learned && gate ? 1 : 0, 0, nullptr, 0, w.Gh};
  e = launch_arpeggio_ex(nullptr, w.hi, T1, L1, q, 2, R1, D, B, inst, k0, k1, 0, 0, w.spins1, w.de1, w.st1, waves1, ex1,
                         st);
  if (e != hipSuccess) return hip_status(e, "detect: stage-1 anneal");
  e = launch_decode_ex(H, y, w.xg, w.spins1, B, R1, 2, D, Mr, M, 1, ltrue, lhat, biterr, nullptr,
                       DecodeExtra{nullptr, nullptr, nullptr, nullptr, w.d1}, st);
  if (e != hipSuccess) return hip_status(e, "detect: stage-1 decode");
  if (R2 == 0) return set_error(0, "");
  e = launch_certify(w.d1, N0, N0_scalar, thr_scale, B, w.list, w.count, st);
  if (e != hipSuccess) return hip_status(e, "detect: certificate");
  AnnealExtra ex2{w.thr, w.used2, w.list, w.count, learned && gate ? 1 : 0, R1, nullptr, 0, w.Gh};
  e = launch_arpeggio_ex(nullptr, w.hi, T2, L2, q, 2, R2, D, B, inst, k0, k1, 0, 0, w.spins2, w.de2, w.st2, waves2, ex2,
                         st);
  if (e != hipSuccess) return hip_status(e, "detect: stage-2 anneal");
  e = launch_decode_ex(H, y, w.xg, w.spins2, B, R2, 2, D, Mr, M, 1, ltrue, lhat, biterr, nullptr,
                       DecodeExtra{w.list, w.count, lhat, w.d1, w.d2}, st);
  return hip_status(e, "detect: stage-2 decode");
}

}  // namespace arp

extern "C" int arp_detect_adaptive(const float* H, const float* y, const float* N0, float N0_scalar, int B, int Mr,
                                   int D, int M, const int32_t* T1, int L1, int R1, const int32_t* T2, int L2, int R2,
                                   int q, float thr_scale, const uint32_t* inst, uint32_t k0, uint32_t k1,
                                   const arp_mlp_weights* mlp, int gate, const int8_t* ltrue, int8_t* lhat,
                                   int32_t* biterr, void* work, size_t work_bytes, void* stream) {
  return arp::detect_chain(H, y, N0, N0_scalar, B, Mr, D, M, T1, L1, R1, T2, L2, R2, q, thr_scale, inst, k0, k1, mlp,
                           gate, ltrue, lhat, biterr, work, work_bytes, stream);
}

