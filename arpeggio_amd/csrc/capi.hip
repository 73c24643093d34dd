// C ABI entry points of libarpeggio_hip.so: argument checks, launch-shape choice, error reporting.
#include <stdarg.h>
#include <stdio.h>

#include <string>

#include "../../include/arpeggio_hip.h"
#include "kernels.h"
#include "launch_bounds.h"
#include "adaptive.h"
#include "mdi.h"

namespace arp {

__global__ void philox_kat_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  uint32_t c0 = in[6 * t], c1 = in[6 * t + 1], c2 = in[6 * t + 2], c3 = in[6 * t + 3];
  philox10(c0, c1, c2, c3, in[6 * t + 4], in[6 * t + 5]);
  out[4 * t] = c0;
  out[4 * t + 1] = c1;
  out[4 * t + 2] = c2;
  out[4 * t + 3] = c3;
}
}  // namespace arp

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

static int ok() {
  g_err.clear();
  return ARP_OK;
}

static int hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) return fail(ARP_E_HIP, "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  return ok();
}

static size_t device_max_lds() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) return 65536;
  return (size_t)v;
}

// Anneals (waves) per workgroup of the Arpeggio kernels: a frame of D <= 64 stages 17 KiB of Gram, so four waves
// per workgroup already fill a CU; a frame of D <= 128 stages 65 KiB (two workgroups per CU), so eight waves. The
// kernels are compiled for workgroups of up to ARP_ARPEGGIO_MAX_THREADS threads.
static thread_local int g_waves_override = 0;

static int arpeggio_waves(int R, int D) {
  const int w = g_waves_override > 0 ? g_waves_override : (D > 64 ? 8 : 4);
  return R < w ? R : w;
}

// Anneals per wave64 of the first-accept launches for D <= 32 (two copies): 0 = the library's choice (kAutoApw), 1 or
// 2. Frames of more than 32 coordinates always run one anneal per wave.
static thread_local int g_apw = 0;
static constexpr int kAutoApw = 1;

namespace arp {
int anneals_per_wave(int D) { return D <= 32 ? (g_apw ? g_apw : kAutoApw) : 1; }
}  // namespace arp

// A workgroup of `waves` wave64s must fit the kernel's compiled bound.
static bool block_fits(int waves, int max_threads) { return waves >= 1 && ARP_WAVE * waves <= max_threads; }

namespace arp {
int set_error(int code, const char* msg) { return code == ARP_OK ? ok() : fail(code, "%s", msg); }
int hip_status(hipError_t e, const char* what) { return hip_check(e, what); }
}  // namespace arp

extern "C" const char* arp_last_error(void) { return g_err.c_str(); }

extern "C" int arp_abi_version(void) { return 5; }

extern "C" int arp_set_anneal_waves(int waves) {
  if (waves < 0 || waves > 16) return fail(ARP_E_ARG, "set_anneal_waves: %d outside 0..16", waves);
  g_waves_override = waves;
  return ok();
}

extern "C" int arp_set_anneals_per_wave(int apw) {
  if (apw < 0 || apw > 2) return fail(ARP_E_ARG, "set_anneals_per_wave: %d outside 0..2", apw);
  g_apw = apw;
  return ok();
}

extern "C" int arp_set_front_end_form(int form) {
  if (form < 0 || form > 2) return fail(ARP_E_ARG, "set_front_end_form: %d, expected 0, 1 or 2", form);
  arp::set_front_end_form(form);
  return ok();
}

extern "C" int arp_philox_kat(const uint32_t* in, uint32_t* out, int n, void* stream) {
  if (n < 0) return fail(ARP_E_ARG, "philox_kat: n = %d < 0", n);
  if (n == 0) return ok();
  if (!in || !out) return fail(ARP_E_ARG, "philox_kat: null pointer");
  hipLaunchKernelGGL(arp::philox_kat_kernel, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)stream, in, out, n);
  return hip_check(hipGetLastError(), "philox_kat launch");
}

extern "C" int arp_anneal_arpeggio(const int32_t* Gt, const int32_t* h, const int32_t* T, int L, int q, int P, int R,
                                   int D, int B, const uint32_t* inst, uint32_t k0, uint32_t k1, int start, int mode,
                                   int8_t* out_s, int64_t* out_de, int64_t* out_st, void* stream) {
  if (P != 2 && P != 4 && P != 6) return fail(ARP_E_ARG, "arpeggio: P = %d, expected 2, 4 or 6", P);
  if (D < 1 || D > 128) return fail(ARP_E_RANGE, "arpeggio: D = %d outside 1..128", D);
  if (L < 1) return fail(ARP_E_ARG, "arpeggio: L = %d, need at least one pass", L);
  if (R < 1) return fail(ARP_E_ARG, "arpeggio: R = %d < 1", R);
  if (B < 0) return fail(ARP_E_ARG, "arpeggio: B = %d < 0", B);
  if (start != 0 && start != 1) return fail(ARP_E_ARG, "arpeggio: start = %d, expected 0 (rand) or 1 (bal)", start);
  if (mode != 0 && mode != 1) return fail(ARP_E_ARG, "arpeggio: mode = %d, expected 0 or 1", mode);
  if (B == 0) return ok();
  if (!Gt || !h || !T || !inst || !out_s || !out_de || !out_st) return fail(ARP_E_ARG, "arpeggio: null pointer");
  const int waves = arpeggio_waves(R, D);
  if (!block_fits(waves, ARP_ARPEGGIO_MAX_THREADS))
    return fail(ARP_E_ARG, "arpeggio: %d anneals per workgroup (%d threads) exceed the kernel's bound of %d threads",
                waves, ARP_WAVE * waves, ARP_ARPEGGIO_MAX_THREADS);
  return hip_check(arp::launch_arpeggio_fx(Gt, h, T, L, q, P, R, D, B, inst, k0, k1, start, mode, out_s,
                                           (long long*)out_de, (long long*)out_st, waves, (hipStream_t)stream),
                   "arpeggio launch");
}

extern "C" int arp_anneal_chord(const int16_t* J, const int32_t* h, const int32_t* h0, const int32_t* Tq,
                                const int32_t* Cq, int q, int S, int P, int R, int N, int B, uint32_t k0, uint32_t k1,
                                int inst0, int mode, int32_t* out_e, int8_t* out_s, uint64_t* stats, void* stream) {
  if (N < 32 || N > 256 || N % 32) return fail(ARP_E_RANGE, "chord: N = %d, expected a multiple of 32 up to 256", N);
  if (P < 1 || N % P) return fail(ARP_E_ARG, "chord: P = %d does not divide N = %d", P, N);
  if (R < 1) return fail(ARP_E_ARG, "chord: R = %d < 1", R);
  if (S < 0 || B < 0) return fail(ARP_E_ARG, "chord: S = %d, B = %d", S, B);
  if (mode != 0 && mode != 1) return fail(ARP_E_ARG, "chord: mode = %d, expected 0 or 1", mode);
  if (B == 0) return ok();
  if (!J || !h || !h0 || (S > 0 && (!Tq || !Cq)) || !out_e || !out_s) return fail(ARP_E_ARG, "chord: null pointer");
  // a frame with N > 128 stages up to 128 KiB of J, one workgroup per CU: sixteen anneals per workgroup fill it
  const int wmax = g_waves_override > 0 ? g_waves_override : (N > 128 ? 16 : 4);
  const int waves = R < wmax ? R : wmax;
  if (!block_fits(waves, ARP_CHORD_MAX_THREADS))
    return fail(ARP_E_ARG, "chord: %d anneals per workgroup (%d threads) exceed the kernel's bound of %d threads", waves,
                ARP_WAVE * waves, ARP_CHORD_MAX_THREADS);
  return hip_check(arp::launch_chord_fx(J, h, h0, Tq, Cq, q, S, P, R, N, B, k0, k1, inst0, mode, out_e, out_s,
                                        (unsigned long long*)stats, waves, device_max_lds(), (hipStream_t)stream),
                   "chord launch");
}

extern "C" int arp_front_end(const float* H, const float* y, const float* N0, float N0_scalar, int B, int Mr, int D,
                             int M, int K, float Q, int8_t* xg, int32_t* Gi, int32_t* hi, float* scale, float* Gf,
                             float* hf, int16_t* J, int32_t* hN, float* z, float* mu, float* Jf, float* hNf,
                             uint16_t* feat, int feat_ld, int16_t* Gh, void* stream) {
  if (D < 1 || D > 128) return fail(ARP_E_RANGE, "front_end: D = %d outside 1..128", D);
  if (Mr < 1 || Mr > 128) return fail(ARP_E_RANGE, "front_end: Mr = %d outside 1..128", Mr);
  if (M != 4 && M != 16) return fail(ARP_E_ARG, "front_end: M = %d, expected 4 or 16", M);
  if (K < 1 || K > 3) return fail(ARP_E_ARG, "front_end: K = %d outside 1..3", K);
  if ((J || Jf) && 2 * K * D > 256) return fail(ARP_E_RANGE, "front_end: expanded N = %d > 256", 2 * K * D);
  if (B < 0) return fail(ARP_E_ARG, "front_end: B = %d < 0", B);
  if (B == 0) return ok();
  if (!H || !y) return fail(ARP_E_ARG, "front_end: null input");
  if (feat && feat_ld < arp::feat_count(D))
    return fail(ARP_E_ARG, "front_end: feat_ld = %d < %d features", feat_ld, arp::feat_count(D));
  arp::FrontOut o{xg, Gi, hi, scale, Gf, hf, J, hN, z, mu, Jf, hNf, feat, feat_ld, Gh};
  return hip_check(arp::launch_front_end(H, y, N0, N0_scalar, B, Mr, D, M, K, Q, o, (hipStream_t)stream),
                   "front_end launch");
}

extern "C" int arp_decode_select(const float* H, const float* y, const int8_t* xg, const int8_t* spins, int B, int R,
                                 int P, int D, int Mr, int M, int delta, const int8_t* ltrue, int8_t* lhat,
                                 int32_t* biterr, float* dist, void* stream) {
  if (D < 1 || D > 128 || Mr < 1 || Mr > 256) return fail(ARP_E_RANGE, "decode: D = %d, Mr = %d", D, Mr);
  if (R < 0 || R > 256) return fail(ARP_E_RANGE, "decode: R = %d outside 0..256", R);
  if (M != 4 && M != 16) return fail(ARP_E_ARG, "decode: M = %d", M);
  if (delta && (P < 1 || P > 6)) return fail(ARP_E_ARG, "decode: P = %d", P);
  if (B < 0) return fail(ARP_E_ARG, "decode: B = %d < 0", B);
  if (B == 0) return ok();
  if (!H || !y || !xg || (R > 0 && !spins)) return fail(ARP_E_ARG, "decode: null input");
  if (biterr && !ltrue) return fail(ARP_E_ARG, "decode: biterr needs ltrue");
  return hip_check(arp::launch_decode_select(H, y, xg, spins, B, R, P, D, Mr, M, delta, ltrue, lhat, biterr, dist,
                                             (hipStream_t)stream),
                   "decode launch");
}

extern "C" int arp_anneal_arpeggio_f32(const float* Gt, const float* h, const float* T, int L, float q, int P, int R,
                                       int D, int B, const uint32_t* inst, uint32_t k0, uint32_t k1, int start,
                                       int8_t* out_s, float* out_de, int64_t* out_st, float* out_w, void* stream) {
  if (P != 2 && P != 4 && P != 6) return fail(ARP_E_ARG, "arpeggio_f32: P = %d, expected 2, 4 or 6", P);
  if (D < 1 || D > 128) return fail(ARP_E_RANGE, "arpeggio_f32: D = %d outside 1..128", D);
  if (L < 1 || R < 1 || B < 0) return fail(ARP_E_ARG, "arpeggio_f32: L = %d, R = %d, B = %d", L, R, B);
  if (start != 0 && start != 1) return fail(ARP_E_ARG, "arpeggio_f32: start = %d", start);
  if (B == 0) return ok();
  if (!Gt || !h || !T || !inst || !out_s || !out_de || !out_st) return fail(ARP_E_ARG, "arpeggio_f32: null pointer");
  const int waves = arpeggio_waves(R, D);
  if (!block_fits(waves, ARP_ARPEGGIO_F32_MAX_THREADS))
    return fail(ARP_E_ARG, "arpeggio_f32: %d anneals per workgroup (%d threads) exceed the kernel's bound of %d threads",
                waves, ARP_WAVE * waves, ARP_ARPEGGIO_F32_MAX_THREADS);
  return hip_check(arp::launch_arpeggio_f32(Gt, h, T, L, q, P, R, D, B, inst, k0, k1, start, out_s, out_de,
                                            (long long*)out_st, out_w, waves, (hipStream_t)stream),
                   "arpeggio_f32 launch");
}

extern "C" int arp_anneal_chord_f32(const float* J, const float* h, const float* h0, const float* T, const float* c,
                                    float q, int S, int P, int R, int N, int B, uint32_t k0, uint32_t k1, int inst0,
                                    float* out_e, int8_t* out_s, float* out_u, int64_t* out_fl, void* stream) {
  if (N < 1 || N > 256) return fail(ARP_E_RANGE, "chord_f32: N = %d outside 1..256", N);
  if (P < 1 || N % P) return fail(ARP_E_ARG, "chord_f32: P = %d does not divide N = %d", P, N);
  if (R < 1 || S < 0 || B < 0) return fail(ARP_E_ARG, "chord_f32: R = %d, S = %d, B = %d", R, S, B);
  if (B == 0) return ok();
  if (!J || !h || !h0 || (S > 0 && (!T || !c)) || !out_e || !out_s) return fail(ARP_E_ARG, "chord_f32: null pointer");
  const int wmax = g_waves_override > 0 ? g_waves_override : 4;
  const int waves = R < wmax ? R : wmax;
  if (!block_fits(waves, ARP_CHORD_F32_MAX_THREADS))
    return fail(ARP_E_ARG, "chord_f32: %d anneals per workgroup (%d threads) exceed the kernel's bound of %d threads",
                waves, ARP_WAVE * waves, ARP_CHORD_F32_MAX_THREADS);
  return hip_check(arp::launch_chord_f32(J, h, h0, T, c, q, S, P, R, N, B, k0, k1, inst0, out_e, out_s, out_u,
                                         (long long*)out_fl, waves, device_max_lds(), (hipStream_t)stream),
                   "chord_f32 launch");
}

extern "C" int arp_anneal_arpeggio_ws(const int32_t* Gt, const int32_t* h, const int32_t* T, int L, int q, int P, int R,
                                      int D, int B, const uint32_t* inst, uint32_t k0, uint32_t k1,
                                      const uint32_t* thr, int gate, int8_t* out_s, int64_t* out_de, int64_t* out_st,
                                      int8_t* out_used, void* stream) {
  if (P != 2 && P != 4 && P != 6) return fail(ARP_E_ARG, "arpeggio_ws: P = %d, expected 2, 4 or 6", P);
  if (D < 1 || D > 128) return fail(ARP_E_RANGE, "arpeggio_ws: D = %d outside 1..128", D);
  if (L < 1 || R < 1 || B < 0) return fail(ARP_E_ARG, "arpeggio_ws: L = %d, R = %d, B = %d", L, R, B);
  if (B == 0) return ok();
  if (!Gt || !h || !T || !inst || !thr || !out_s || !out_de || !out_st)
    return fail(ARP_E_ARG, "arpeggio_ws: null pointer");
  const int waves = arpeggio_waves(R, D);
  if (!block_fits(waves, ARP_ARPEGGIO_MAX_THREADS))
    return fail(ARP_E_ARG, "arpeggio_ws: %d anneals per workgroup (%d threads) exceed the kernel's bound of %d threads",
                waves, ARP_WAVE * waves, ARP_ARPEGGIO_MAX_THREADS);
  return hip_check(arp::launch_arpeggio_fx(Gt, h, T, L, q, P, R, D, B, inst, k0, k1, 0, 0, out_s, (long long*)out_de,
                                           (long long*)out_st, waves, (hipStream_t)stream, thr, gate ? 1 : 0,
                                           out_used),
                   "arpeggio_ws launch");
}

extern "C" int arp_anneal_arpeggio_batched(const arp_batch_class* classes, int n_classes, int R, uint32_t k0,
                                           uint32_t k1, uint32_t* counter, int workgroups, int waves, void* stream) {
  if (n_classes < 1 || n_classes > ARP_MAX_CLASSES)
    return fail(ARP_E_ARG, "arpeggio_batched: n_classes = %d outside 1..%d", n_classes, ARP_MAX_CLASSES);
  if (!classes || !counter) return fail(ARP_E_ARG, "arpeggio_batched: null pointer");
  if (R < 1) return fail(ARP_E_ARG, "arpeggio_batched: R = %d", R);
  if (!block_fits(waves, ARP_BATCHED_MAX_THREADS) || workgroups < 1)
    return fail(ARP_E_ARG, "arpeggio_batched: workgroups = %d, waves = %d (1..%d)", workgroups, waves,
                ARP_BATCHED_MAX_THREADS / ARP_WAVE);
  arp::ArpBatchArgs a = {};
  a.n_classes = n_classes;
  a.R = R;
  a.k0 = k0;
  a.k1 = k1;
  size_t lds = 0;
  int tiles = 0;
  for (int i = 0; i < n_classes; ++i) {
    const arp_batch_class& c = classes[i];
    if (c.D < 1 || c.D > 128) return fail(ARP_E_RANGE, "arpeggio_batched: class %d D = %d outside 1..128", i, c.D);
    if (c.apw < 0 || c.apw > 2) return fail(ARP_E_ARG, "arpeggio_batched: class %d apw = %d outside 0..2", i, c.apw);
    if (c.B < 0 || c.L < 1 || c.tile < 1 || (c.start != 0 && c.start != 1))
      return fail(ARP_E_ARG, "arpeggio_batched: class %d B = %d, L = %d, tile = %d, start = %d", i, c.B, c.L, c.tile,
                  c.start);
    if (c.B > 0 && (!c.Gt || !c.h || !c.inst || !c.T || !c.out_s || !c.out_de || !c.out_st))
      return fail(ARP_E_ARG, "arpeggio_batched: class %d null pointer", i);
    const size_t stride = (size_t)c.D * c.D + ARP_WAVE * (c.D <= ARP_WAVE ? 1 : 2);
    const int t = c.tile < c.B ? c.tile : (c.B > 0 ? c.B : 1);
    lds = lds > stride * t * 4 ? lds : stride * t * 4;
    arp::ArpBatchClass& d = a.c[i];
    d.Gt = c.Gt;
    d.h = c.h;
    d.inst = c.inst;
    d.T = c.T;
    d.out_s = c.out_s;
    d.out_de = (long long*)c.out_de;
    d.out_st = (long long*)c.out_st;
    d.D = c.D;
    d.B = c.B;
    d.L = c.L;
    d.q = c.q;
    d.start = c.start;
    d.tile = t;
    d.thr = c.thr;
    d.used = c.thr ? c.out_used : nullptr;
    d.gate = c.gate ? 1 : 0;
    d.apw = c.D <= 32 ? (c.apw ? c.apw : arp::anneals_per_wave(c.D)) : 1;
    d.first_tile = tiles;
    tiles += (c.B + t - 1) / t;
  }
  a.n_tiles = tiles;
  if (tiles == 0) return ok();
  if (lds > device_max_lds())
    return fail(ARP_E_RANGE, "arpeggio_batched: tiles need %zu bytes of LDS, the device allows %zu", lds,
                device_max_lds());
  return hip_check(arp::launch_arpeggio_batched(a, counter, workgroups, waves, lds, (hipStream_t)stream),
                   "arpeggio_batched launch");
}

extern "C" int arp_mlp_start_fused(const uint16_t* X, int M, int F, const uint16_t* W1t, const float* b1,
                                   const uint16_t* W2t, const float* b2, const uint16_t* W3t, const float* b3, int H,
                                   int N, float inv_temp, uint32_t* thr, void* stream) {
  if (M < 0 || F < 32 || F % 32 || F > 1016)
    return fail(ARP_E_ARG, "mlp_start_fused: F = %d, need a multiple of 32 in 32..1016 (M = %d)", F, M);
  if (H != 256) return fail(ARP_E_ARG, "mlp_start_fused: hidden width %d, the fused kernel is built for 256", H);
  if (N < 64 || N > 256 || N % 64) return fail(ARP_E_ARG, "mlp_start_fused: N = %d, need 64, 128, 192 or 256", N);
  if (M == 0) return ok();
  if (!X || !W1t || !b1 || !W2t || !b2 || !W3t || !b3 || !thr) return fail(ARP_E_ARG, "mlp_start_fused: null pointer");
  return hip_check(arp::launch_mlp_start_fused(X, M, F, W1t, b1, W2t, b2, W3t, b3, N, inv_temp, thr,
                                               (hipStream_t)stream),
                   "mlp_start_fused launch");
}

extern "C" int arp_mlp_gemm(const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, int M, int N,
                            int K, const float* bias, int act, float* Cf, float* Cpre, uint16_t* Cb, long ldc,
                            void* stream) {
  if (M < 0 || N < 0 || K < 1) return fail(ARP_E_ARG, "mlp_gemm: M = %d, N = %d, K = %d", M, N, K);
  if (act != 0 && act != 1) return fail(ARP_E_ARG, "mlp_gemm: act = %d", act);
  if (M == 0 || N == 0) return ok();
  if (!A || !B || (!Cf && !Cpre && !Cb)) return fail(ARP_E_ARG, "mlp_gemm: null pointer");
  return hip_check(arp::launch_gemm_bf16(A, sam, sak, B, sbk, sbn, M, N, K, bias, act, Cf, Cpre, Cb, ldc,
                                         (hipStream_t)stream),
                   "mlp_gemm launch");
}

extern "C" int arp_mlp_gelu_bwd(const float* dH, const float* Apre, long n, uint16_t* dAb, float* dAf, void* stream) {
  if (n < 0 || !dH || !Apre || (!dAb && !dAf)) return fail(ARP_E_ARG, "mlp_gelu_bwd: bad arguments");
  if (n == 0) return ok();
  return hip_check(arp::launch_gelu_bwd(dH, Apre, n, dAb, dAf, (hipStream_t)stream), "gelu_bwd launch");
}

extern "C" int arp_mlp_bce(const float* z, long ld, const int8_t* y, int M, int N, uint16_t* dzb, float* dzf,
                           float* loss, void* stream) {
  if (M < 0 || N < 0 || !z || !y) return fail(ARP_E_ARG, "mlp_bce: bad arguments");
  if (M == 0 || N == 0) return ok();
  return hip_check(arp::launch_bce(z, ld, y, M, N, dzb, dzf, loss, (hipStream_t)stream), "bce launch");
}

extern "C" int arp_mlp_colsum(const float* X, long ld, int M, int N, float* out, void* stream) {
  if (M < 0 || N < 0 || !X || !out) return fail(ARP_E_ARG, "mlp_colsum: bad arguments");
  if (N == 0) return ok();
  return hip_check(arp::launch_colsum(X, ld, M, N, out, (hipStream_t)stream), "colsum launch");
}

extern "C" int arp_mlp_adam(float* W, const float* G, float* m, float* v, long n, float lr, float b1, float b2,
                            float eps, int step, uint16_t* Wb, void* stream) {
  if (n < 0 || !W || !G || !m || !v || step < 1) return fail(ARP_E_ARG, "mlp_adam: bad arguments");
  if (n == 0) return ok();
  return hip_check(arp::launch_adam(W, G, m, v, n, lr, b1, b2, eps, step, Wb, (hipStream_t)stream), "adam launch");
}

extern "C" int arp_cast_bf16(const float* x, long n, uint16_t* y, void* stream) {
  if (n < 0 || !x || !y) return fail(ARP_E_ARG, "cast_bf16: bad arguments");
  if (n == 0) return ok();
  return hip_check(arp::launch_cast_bf16(x, n, y, (hipStream_t)stream), "cast launch");
}

extern "C" int arp_start_thresholds(const float* z, long ld, int M, int N, float inv_temp, uint32_t* thr,
                                    void* stream) {
  if (M < 0 || N < 0 || !z || !thr) return fail(ARP_E_ARG, "start_thresholds: bad arguments");
  if (M == 0 || N == 0) return ok();
  return hip_check(arp::launch_start_thr(z, ld, M, N, inv_temp, thr, (hipStream_t)stream), "start_thr launch");
}

extern "C" int arp_anneal_arpeggio_list(const int32_t* Gt, const int16_t* Gh, const int32_t* h, const int32_t* T, int L,
                                        int q, int P,
                                        int R, int D, int B, const uint32_t* inst, uint32_t k0, uint32_t k1, int start,
                                        const uint32_t* thr, int gate, int r0, const int32_t* list,
                                        const int32_t* count, int8_t* out_s, int64_t* out_de, int64_t* out_st,
                                        int8_t* out_used, void* stream) {
  if (P != 2 && P != 4 && P != 6) return fail(ARP_E_ARG, "arpeggio_list: P = %d, expected 2, 4 or 6", P);
  if (D < 1 || D > 128) return fail(ARP_E_RANGE, "arpeggio_list: D = %d outside 1..128", D);
  if (L < 1 || R < 1 || B < 0 || r0 < 0)
    return fail(ARP_E_ARG, "arpeggio_list: L = %d, R = %d, B = %d, r0 = %d", L, R, B, r0);
  if (start != 0 && start != 1) return fail(ARP_E_ARG, "arpeggio_list: start = %d", start);
  if (!list != !count) return fail(ARP_E_ARG, "arpeggio_list: list and count go together");
  if (B == 0) return ok();
  if (!Gt == !Gh) return fail(ARP_E_ARG, "arpeggio_list: exactly one of Gt (int32) and Gh (int16) must be given");
  if (!h || !T || !inst || !out_s || !out_de || !out_st) return fail(ARP_E_ARG, "arpeggio_list: null pointer");
  // a frame list is usually short (the open frames of a two-stage detection): more anneals in flight per frame
  const int waves = list && g_waves_override == 0 ? (R < 8 ? R : 8) : arpeggio_waves(R, D);
  if (!block_fits(waves, ARP_ARPEGGIO_MAX_THREADS))
    return fail(ARP_E_ARG, "arpeggio_list: %d anneals per workgroup (%d threads) exceed the kernel's bound of %d "
                "threads", waves, ARP_WAVE * waves, ARP_ARPEGGIO_MAX_THREADS);
  arp::AnnealExtra ex{thr, thr ? out_used : nullptr, list, count, gate ? 1 : 0, r0, nullptr, 0, Gh};
  return hip_check(arp::launch_arpeggio_ex(Gt, h, T, L, q, P, R, D, B, inst, k0, k1, thr ? 0 : start, 0, out_s,
                                           (long long*)out_de, (long long*)out_st, waves, ex, (hipStream_t)stream),
                   "arpeggio_list launch");
}

extern "C" int arp_decode_select_list(const float* H, const float* y, const int8_t* xg, const int8_t* spins, int B,
                                      int R, int P, int D, int Mr, int M, int delta, const int8_t* ltrue,
                                      const int32_t* list, const int32_t* count, const int8_t* prior_l,
                                      const float* prior_d, int8_t* lhat, int32_t* biterr, float* best_d,
                                      void* stream) {
  if (D < 1 || D > 128 || Mr < 1 || Mr > 256) return fail(ARP_E_RANGE, "decode_list: D = %d, Mr = %d", D, Mr);
  if (R < 0 || R > 256) return fail(ARP_E_RANGE, "decode_list: R = %d outside 0..256", R);
  if (M != 4 && M != 16) return fail(ARP_E_ARG, "decode_list: M = %d", M);
  if (delta && (P < 1 || P > 6)) return fail(ARP_E_ARG, "decode_list: P = %d", P);
  if (B < 0) return fail(ARP_E_ARG, "decode_list: B = %d < 0", B);
  if (!list != !count) return fail(ARP_E_ARG, "decode_list: list and count go together");
  if (!prior_l != !prior_d) return fail(ARP_E_ARG, "decode_list: prior levels and distances go together");
  if (B == 0) return ok();
  if (!H || !y || !xg || (R > 0 && !spins)) return fail(ARP_E_ARG, "decode_list: null input");
  if (biterr && !ltrue) return fail(ARP_E_ARG, "decode_list: biterr needs ltrue");
  arp::DecodeExtra ex{list, count, prior_l, prior_d, best_d};
  return hip_check(arp::launch_decode_ex(H, y, xg, spins, B, R, P, D, Mr, M, delta, ltrue, lhat, biterr, nullptr, ex,
                                         (hipStream_t)stream),
                   "decode_list launch");
}

extern "C" int arp_certify(const float* best_d, const float* N0, float N0_scalar, float thr_scale, int B,
                           int32_t* list, int32_t* count, void* stream) {
  if (B < 0) return fail(ARP_E_ARG, "certify: B = %d < 0", B);
  if (!list || !count || (B > 0 && !best_d)) return fail(ARP_E_ARG, "certify: null pointer");
  return hip_check(arp::launch_certify(best_d, N0, N0_scalar, thr_scale, B, list, count, (hipStream_t)stream),
                   "certify launch");
}

extern "C" int arp_mmse_sic(const float* H, const float* y, const float* N0, float N0_scalar, int B, int Mr, int D, int M,
                            int per_user, int8_t* xs, void* stream) {
  if (D < 2 || D > 128 || D % 2 || Mr < 1 || Mr > 128)
    return fail(ARP_E_RANGE, "mmse_sic: D = %d (even, 2..128), Mr = %d (1..128)", D, Mr);
  if (M != 4 && M != 16) return fail(ARP_E_ARG, "mmse_sic: M = %d", M);
  if (B < 0) return fail(ARP_E_ARG, "mmse_sic: B = %d < 0", B);
  if (B == 0) return ok();
  if (!H || !y || !xs) return fail(ARP_E_ARG, "mmse_sic: null pointer");
  return hip_check(arp::launch_mmse_sic(H, y, N0, N0_scalar, B, Mr, D, M, per_user ? 1 : 0, xs, (hipStream_t)stream),
                   "mmse_sic launch");
}

extern "C" int arp_mdi_radius(const float* H, const float* y, const int8_t* xg, int B, int Mr, int D, int M, int Rmax,
                              int32_t* R, float* Gam, float* dist, void* stream) {
  if (D < 1 || D > 256 || Mr < 1 || Mr > 256) return fail(ARP_E_RANGE, "mdi_radius: D = %d, Mr = %d (1..256)", D, Mr);
  if (M != 4 && M != 16) return fail(ARP_E_ARG, "mdi_radius: M = %d", M);
  if (B < 0 || Rmax < 0) return fail(ARP_E_ARG, "mdi_radius: B = %d, Rmax = %d", B, Rmax);
  if (B == 0) return ok();
  if (!H || !y || !xg) return fail(ARP_E_ARG, "mdi_radius: null pointer");
  return hip_check(arp::launch_mdi_radius(H, y, xg, B, Mr, D, M, Rmax, R, Gam, dist, (hipStream_t)stream),
                   "mdi_radius launch");
}

extern "C" int arp_mdi_build(const float* H, const float* y, const int8_t* xg, const int32_t* list, int n, int Mr,
                             int D, float Q, int32_t* Gi, int32_t* hi, void* stream) {
  if (D < 1 || D > 128 || Mr < 1 || Mr > 128) return fail(ARP_E_RANGE, "mdi_build: D = %d, Mr = %d (1..128)", D, Mr);
  if (n < 0) return fail(ARP_E_ARG, "mdi_build: n = %d < 0", n);
  if (n == 0) return ok();
  if (!H || !y || !xg || !list || !Gi || !hi) return fail(ARP_E_ARG, "mdi_build: null pointer");
  return hip_check(arp::launch_mdi_build(H, y, xg, list, n, Mr, D, Q, Gi, hi, (hipStream_t)stream),
                   "mdi_build launch");
}

extern "C" int arp_anneal_chord_lane(const int32_t* Gt, const int32_t* h, const int32_t* T, const int32_t* Cq, int L,
                                     int q, int P, int R, int D, int B, const uint32_t* inst, uint32_t k0,
                                     uint32_t k1, int rng, const int32_t* list, const int32_t* count, int8_t* out_s,
                                     int64_t* out_de, int64_t* out_st, void* stream) {
  if (P < 1 || P > 6 || P % 2) return fail(ARP_E_ARG, "chord_lane: P = %d, expected 2, 4 or 6", P);
  if (D < 1 || D > 128) return fail(ARP_E_RANGE, "chord_lane: D = %d outside 1..128", D);
  if (L < 1 || R < 1 || B < 0) return fail(ARP_E_ARG, "chord_lane: L = %d, R = %d, B = %d", L, R, B);
  if (rng != 0 && rng != 1) return fail(ARP_E_ARG, "chord_lane: rng = %d, expected 0 (Philox) or 1 (xorshift)", rng);
  if (!list != !count) return fail(ARP_E_ARG, "chord_lane: list and count go together");
  if (B == 0) return ok();
  if (!Gt || !h || !T || !inst || !out_s || !out_de || !out_st) return fail(ARP_E_ARG, "chord_lane: null pointer");
  const int waves = arpeggio_waves(R, D);
  if (!block_fits(waves, ARP_ARPEGGIO_MAX_THREADS))
    return fail(ARP_E_ARG, "chord_lane: %d anneals per workgroup (%d threads) exceed the kernel's bound of %d threads",
                waves, ARP_WAVE * waves, ARP_ARPEGGIO_MAX_THREADS);
  arp::AnnealExtra ex{nullptr, nullptr, list, count, 0, 0, Cq, rng};
  return hip_check(arp::launch_arpeggio_ex(Gt, h, T, L, q, P, R, D, B, inst, k0, k1, 0, 1, out_s, (long long*)out_de,
                                           (long long*)out_st, waves, ex, (hipStream_t)stream),
                   "chord_lane launch");
}
