// Allocation-free variant of the whole-chain entry point: arp_detect_adaptive's chain with its workspace in a
// file-scope device array (`scratch`) instead of caller memory, for callers that cannot allocate (code under HIP graph
// capture that must not call hipMalloc, tight serving loops, harnesses that forbid allocation).
//
// The array's device address on each device is taken once, with that device current, by hipGetSymbolAddress on the
// entry point's first call there; host code uses that address, never the array's name. Every call on a device shares
// that device's array and leaves its contents behind; the chain writes every piece before reading it (the certificate
// clears the open-frame count itself), so no call depends on an earlier one. Two calls that overlap on one device
// (concurrent streams, or host threads) race on the array.
#include <stdio.h>

#include <mutex>

#include "adaptive.h"

namespace arp {

// 32 MiB per device: about 1,500 frames of 32 users (D = 64) with R1 + R2 = 20 anneals per frame and the learned
// start; arp_detect_workspace_bytes gives a shape's exact need.
#define ARP_SCRATCH_BYTES (32u << 20)
static __device__ __align__(256) uint8_t scratch[ARP_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

static int scratch_on_current_device(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_status(e, "detect_noalloc: hipGetDevice");
  if (dev < 0 || dev >= 64) return set_error(ARP_E_RANGE, "detect_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_status(e, "detect_noalloc: hipGetSymbolAddress(scratch)");
    if (!p) return set_error(ARP_E_HIP, "detect_noalloc: hipGetSymbolAddress(scratch) gave a null address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return set_error(0, "");
}

}  // namespace arp

extern "C" int arp_detect_adaptive_noalloc(const float* H, const float* y, const float* N0, float N0_scalar, int B,
                                           int Mr, int D, int M, const int32_t* T1, int L1, int R1,
                                           const int32_t* T2, int L2, int R2, int q, float thr_scale,
                                           const uint32_t* inst, uint32_t k0, uint32_t k1,
                                           const arp_mlp_weights* mlp, int gate, const int8_t* ltrue, int8_t* lhat,
                                           int32_t* biterr, void* stream) {
  if (B > 0) {
    const size_t need = arp_detect_workspace_bytes(B, D, R1, R2, mlp != nullptr);
    if (need > ARP_SCRATCH_BYTES) {
      char msg[200];
      snprintf(msg, sizeof msg,
               "detect_noalloc: the shape needs %zu bytes of scratch, the library holds %u bytes per device", need,
               (unsigned)ARP_SCRATCH_BYTES);
      return arp::set_error(ARP_E_RANGE, msg);
    }
  }
  void* ws = nullptr;
  if (B > 0) {
    const int rc = arp::scratch_on_current_device(&ws);
    if (rc != ARP_OK) return rc;
  }
  return arp::detect_chain(H, y, N0, N0_scalar, B, Mr, D, M, T1, L1, R1, T2, L2, R2, q, thr_scale, inst, k0, k1, mlp,
                           gate, ltrue, lhat, biterr, ws, ARP_SCRATCH_BYTES, stream);
}
