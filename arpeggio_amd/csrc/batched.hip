// The mixed-size batch as whole-chain C entry points: per class front end, one persistent anneal over every class,
// per class decode (arp_detect_batched), and the launch-shape autotuner of the persistent anneal (arp_tune_batched).
//
// No reference counterpart: the reference detects one fixed shape per call with a batch-scalar N0
// (arpeggio/mimo/mimo.py:61-82, arpeggio/mimo/fxdetect.py:24-43). Every anneal here equals arp_anneal_arpeggio's on the
// same class buffers, stream for stream (Philox instance word = the frame's inst).
#include "kernels.h"
#include "adaptive.h"

#include <cstdint>
#include <cstring>

namespace {

constexpr size_t kAlign = 256;
constexpr int kLdsPerCu = 160 * 1024;

size_t up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Per-class buffers carved from the workspace, in class order.
struct ClassBufs {
  int8_t* xg;
  int32_t* Gi;
  int32_t* hi;
  int8_t* spins;
  long long* de;
  long long* st;
};

size_t class_bytes(const arp_batch_frames& c, int R) {
  const size_t D = 2 * (size_t)c.nt, B = c.B;
  return up(B * D) + up(B * D * D * 4) + up(B * D * 4) + up(B * R * 2 * D) + up(B * R * 8) + up(B * R * 16);
}

char* carve(char* p, const arp_batch_frames& c, int R, ClassBufs& b) {
  const size_t D = 2 * (size_t)c.nt, B = c.B;
  b.xg = reinterpret_cast<int8_t*>(p);
  p += up(B * D);
  b.Gi = reinterpret_cast<int32_t*>(p);
  p += up(B * D * D * 4);
  b.hi = reinterpret_cast<int32_t*>(p);
  p += up(B * D * 4);
  b.spins = reinterpret_cast<int8_t*>(p);
  p += up(B * R * 2 * D);
  b.de = reinterpret_cast<long long*>(p);
  p += up(B * R * 8);
  b.st = reinterpret_cast<long long*>(p);
  p += up(B * R * 16);
  return p;
}

int check_classes(const arp_batch_frames* classes, int n, int R, const int32_t* T, int L, float Q) {
  if (!classes || n < 1 || n > ARP_MAX_CLASSES)
    return arp::set_error(ARP_E_ARG, "detect_batched: 1..8 classes expected");
  if (R < 1 || L < 1 || L % 2 || !T) return arp::set_error(ARP_E_ARG, "detect_batched: R, L (a multiple of 2) or T");
  if (!(Q > 0.f) || Q > 16384.f) return arp::set_error(ARP_E_ARG, "detect_batched: Q outside (0, 2^14]");
  for (int i = 0; i < n; ++i) {
    const arp_batch_frames& c = classes[i];
    if (c.nt < 1 || c.nt > 64) return arp::set_error(ARP_E_RANGE, "detect_batched: a class has nt outside 1..64");
    if (c.M != 4 && c.M != 16) return arp::set_error(ARP_E_ARG, "detect_batched: a class has M other than 4, 16");
    if (c.B < 0) return arp::set_error(ARP_E_ARG, "detect_batched: a class has B < 0");
    if (c.B > 0 && (!c.H || !c.y || !c.inst || !c.lhat))
      return arp::set_error(ARP_E_ARG, "detect_batched: a class has a null pointer");
    if (c.biterr && !c.ltrue) return arp::set_error(ARP_E_ARG, "detect_batched: biterr needs ltrue");
    if (!c.N0 && !(c.N0_scalar > 0.f)) return arp::set_error(ARP_E_ARG, "detect_batched: N0 must be positive");
  }
  return ARP_OK;
}

// the per-class front ends and the class table of the persistent launch (tiles from the LDS budget)
int front_and_table(const arp_batch_frames* classes, int n, const int32_t* T, int L, int q, int R, float Q,
                    int lds_budget, int apw, void* work, size_t work_bytes, hipStream_t st, arp_batch_class* tab,
                    ClassBufs* bufs, uint32_t** counter) {
  char* p = static_cast<char*>(work);
  for (int i = 0; i < n; ++i) p = carve(p, classes[i], R, bufs[i]);
  *counter = reinterpret_cast<uint32_t*>(p);
  for (int i = 0; i < n; ++i) {
    const arp_batch_frames& c = classes[i];
    const int D = 2 * c.nt;
    if (c.B > 0) {
      arp::FrontOut o{};
      o.xg = bufs[i].xg;
      o.Gi = bufs[i].Gi;
      o.hi = bufs[i].hi;
      const hipError_t e = arp::launch_front_end(c.H, c.y, c.N0, c.N0_scalar, c.B, D, D, c.M, 1, Q, o, st);
      if (e != hipSuccess) return arp::hip_status(e, "detect_batched front end");
    }
    const size_t stride = (size_t)D * D + ARP_WAVE * (D <= ARP_WAVE ? 1 : 2);
    const int tile = (int)(lds_budget / (4 * stride)) > 1 ? (int)(lds_budget / (4 * stride)) : 1;
    tab[i] = arp_batch_class{bufs[i].Gi, bufs[i].hi, c.inst, T, bufs[i].spins, (int64_t*)bufs[i].de,
                             (int64_t*)bufs[i].st, D, c.B, L, q, 0, tile, nullptr, nullptr, 0, apw};
  }
  (void)work_bytes;
  return ARP_OK;
}

}  // namespace

extern "C" size_t arp_detect_batched_workspace_bytes(const arp_batch_frames* classes, int n_classes, int R) {
  if (!classes || n_classes < 1 || n_classes > ARP_MAX_CLASSES || R < 1) return 0;
  size_t s = 0;
  for (int i = 0; i < n_classes; ++i) s += class_bytes(classes[i], R);
  return s + kAlign;  // the launch counter
}

extern "C" int arp_detect_batched(const arp_batch_frames* classes, int n_classes, const int32_t* T, int L, int q,
                                  int R, float Q, uint32_t k0, uint32_t k1, int workgroups, int waves, int lds_budget,
                                  void* work, size_t work_bytes, void* stream) {
  int rc = check_classes(classes, n_classes, R, T, L, Q);
  if (rc) return rc;
  if (!work || work_bytes < arp_detect_batched_workspace_bytes(classes, n_classes, R))
    return arp::set_error(ARP_E_ARG, "detect_batched: workspace smaller than arp_detect_batched_workspace_bytes");
  if (lds_budget < 1 || lds_budget > kLdsPerCu) return arp::set_error(ARP_E_ARG, "detect_batched: lds_budget");
  const hipStream_t st = (hipStream_t)stream;
  arp_batch_class tab[ARP_MAX_CLASSES];
  ClassBufs bufs[ARP_MAX_CLASSES];
  uint32_t* counter = nullptr;
  rc = front_and_table(classes, n_classes, T, L, q, R, Q, lds_budget, 0, work, work_bytes, st, tab, bufs, &counter);
  if (rc) return rc;
  rc = arp_anneal_arpeggio_batched(tab, n_classes, R, k0, k1, counter, workgroups, waves, stream);
  if (rc) return rc;
  for (int i = 0; i < n_classes; ++i) {
    const arp_batch_frames& c = classes[i];
    if (c.B == 0) continue;
    const int D = 2 * c.nt;
    const hipError_t e = arp::launch_decode_select(c.H, c.y, bufs[i].xg, bufs[i].spins, c.B, R, 2, D, D, c.M, 1,
                                                   c.ltrue, c.lhat, c.biterr, nullptr, st);
    if (e != hipSuccess) return arp::hip_status(e, "detect_batched decode");
  }
  return arp::set_error(0, "");
}

extern "C" int arp_tune_batched(const arp_batch_frames* classes, int n_classes, const int32_t* T, int L, int q, int R,
                                float Q, uint32_t k0, uint32_t k1, int reps, void* work, size_t work_bytes,
                                int* best_shape, float* best_ms, void* stream) {
  int rc = check_classes(classes, n_classes, R, T, L, Q);
  if (rc) return rc;
  if (!work || work_bytes < arp_detect_batched_workspace_bytes(classes, n_classes, R))
    return arp::set_error(ARP_E_ARG, "tune_batched: workspace smaller than arp_detect_batched_workspace_bytes");
  if (!best_shape || reps < 1) return arp::set_error(ARP_E_ARG, "tune_batched: best_shape or reps");
  const hipStream_t st = (hipStream_t)stream;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                              hipSuccess)
    return arp::set_error(ARP_E_HIP, "tune_batched: device query failed");
  size_t need = 0;
  bool small = false;  // a class of D <= 32: anneals per wave 1 and 2 are both candidates
  for (int i = 0; i < n_classes; ++i) {
    const size_t D = 2 * (size_t)classes[i].nt;
    small = small || D <= 32;
    const size_t s = 4 * (D * D + ARP_WAVE * (D <= ARP_WAVE ? 1 : 2));
    need = need > s ? need : s;
  }
  const int budgets[4] = {(int)need, (int)(2 * need < (size_t)kLdsPerCu ? 2 * need : kLdsPerCu), kLdsPerCu / 2,
                          kLdsPerCu};
  arp_batch_class tab[ARP_MAX_CLASSES];
  ClassBufs bufs[ARP_MAX_CLASSES];
  uint32_t* counter = nullptr;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return arp::set_error(ARP_E_HIP, "tune_batched: event creation failed");
  float best = -1.f;
  int bw = 0, bv = 0, bb = 0, ba = 1;
  int tried[4] = {-1, -1, -1, -1};
  for (int k = 0; k < 4 && !rc; ++k) {
    const int budget = budgets[k];
    bool seen = budget < (int)need;
    for (int j = 0; j < k; ++j) seen = seen || tried[j] == budget;
    tried[k] = budget;
    if (seen) continue;
    rc = front_and_table(classes, n_classes, T, L, q, R, Q, budget, 1, work, work_bytes, st, tab, bufs, &counter);
    const int per_cu = kLdsPerCu / budget > 1 ? kLdsPerCu / budget : 1;
    for (int cand = 0; cand < (small ? 6 : 3); ++cand) {
      if (rc) break;
      const int waves = 4 << (cand % 3), apw = 1 + cand / 3;
      for (int i = 0; i < n_classes; ++i) tab[i].apw = apw;
      const int wg = cus * per_cu;
      rc = arp_anneal_arpeggio_batched(tab, n_classes, R, k0, k1, counter, wg, waves, stream);  // warm-up
      (void)hipEventRecord(e0, st);
      for (int r = 0; r < reps && !rc; ++r)
        rc = arp_anneal_arpeggio_batched(tab, n_classes, R, k0, k1, counter, wg, waves, stream);
      (void)hipEventRecord(e1, st);
      if (rc) break;
      if (hipEventSynchronize(e1) != hipSuccess) {
        rc = arp::set_error(ARP_E_HIP, "tune_batched: event synchronize failed");
        break;
      }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      ms /= (float)reps;
      if (best < 0.f || ms < best) {
        best = ms;
        bw = wg;
        bv = waves;
        bb = budget;
        ba = apw;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  best_shape[0] = bw;
  best_shape[1] = bv;
  best_shape[2] = bb;
  best_shape[3] = ba;
  if (best_ms) *best_ms = best;
  return arp::set_error(0, "");
}
