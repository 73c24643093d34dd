/*
 * arpeggio_hip — C ABI of the MI355X (gfx950) kernels for the delta-encoded Ising annealing detector.
 *
 * Every entry point takes device pointers (memory the caller allocated, e.g. PyTorch-ROCm tensors), plain sizes and
 * a HIP stream (hipStream_t passed as void*; NULL = the default stream). Launches are asynchronous on that stream;
 * nothing is allocated or freed here. Each returns 0 on success or a negative ARP_E* code, with a message available
 * from arp_last_error() (per calling thread). Calls are thread-compatible: distinct streams may be driven from
 * distinct threads.
 *
 * Integer ranges (the reference kernels share them): fields are int32, so for the Arpeggio kernel
 * max_i(|h_i| + 2 P sum_j |G_ij|) + 2 max|G_jj| + |q| + 2 max T must stay below 2^31, and 0 <= T < 2^29.
 * The Python layer (arpeggio_amd) checks these before a launch.
 */
#ifndef ARPEGGIO_HIP_H
#define ARPEGGIO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARP_OK 0
#define ARP_E_ARG (-1)    /* invalid argument (sizes, pointers, unsupported P / D / N) */
#define ARP_E_HIP (-2)    /* a HIP runtime call failed */
#define ARP_E_RANGE (-3)  /* a size beyond what the kernel supports */

/* Message of the last failing call on this thread ("" if none). */
const char *arp_last_error(void);

/* ABI version of this library (incremented when a signature changes). */
int arp_abi_version(void);

/* Anneals (wave64s) per workgroup of the annealing entry points called afterwards on this thread: 0 restores the
 * library's choice (4 for D <= 64 and 8 beyond for the first-accept kernels, 4 or 16 for Chord); a choice whose
 * workgroup exceeds a kernel's compiled bound makes that entry point return ARP_E_ARG before anything is launched.
 * A tuning knob with no reference counterpart (the reference launches 128-thread blocks, arpeggiofx.py:39-42). */
int arp_set_anneal_waves(int waves);

/* Anneals per wave64 of the two-copy first-accept launches for frames of D <= 32 (arp_anneal_arpeggio, _ws, _list,
 * _batched, the adaptive chains) called afterwards on this thread: 2 runs two anneals per wave (lanes 0-31 anneal r,
 * lanes 32-63 anneal r + 1), 1 one anneal per wave (half the lanes idle at D = 32), 0 the library's choice. Results
 * are identical either way. The reference runs one anneal per thread (arpeggio_fx.cu:39-42); this packing has no
 * counterpart there. */
int arp_set_anneals_per_wave(int apw);

/* Front-end schedule for D, Mr <= 64 (no counterpart in the reference, which has one float front end): 0 (default)
 * the matrix-core block sweep, 1 the register sweep, 2 the complex-lift form for D = Mr in {32, 64} (other shapes as
 * 0): H must be the real lift of a complex channel, H = [[Re, -Im], [Im, Re]] (mimo.complex_to_real,
 * arpeggio/mimo/mimo.py:55-58), of which only the first D/2 columns are read; the Gram A + iB is Hermitian and the
 * inverse is swept in complex form, with about half the matrix-core work. With form 2 the two-copy decodes (D = Mr in
 * {32, 64}) also read only those columns (same distances bit for bit). All forms compute the same quantities; the
 * results agree to float32 rounding. Process-wide. */
int arp_set_front_end_form(int form);

/* Chip probes (no reference counterpart; measurement only). arp_probe_lds: conflict-free 16-byte LDS reads, `iters`
 * rounds of 8 KiB per wave64, 256-thread workgroups; arp_probe_issue: mode 0 eight independent FP32 MFMAs (16x16x4)
 * per iteration, mode 1 sixty-four independent FMAs, mode 2 both in each wave. out: workgroups * 256 floats. */
int arp_probe_lds(float *out, int workgroups, int iters, void *stream);
int arp_probe_issue(int mode, float *out, int workgroups, int iters, void *stream);

/* Philox4x32-10 known-answer kernel: in (n, 6) uint32 rows (counter[4], key[2]) -> out (n, 4).
 * Replaces the reference's device KAT `philox_kat` (arpeggio/mimo/chord_kernel.cu:26-32), driven by
 * Kernel.philox_kat (arpeggio/mimo/chordfx.py:180-187). */
int arp_philox_kat(const uint32_t *in, uint32_t *out, int n, void *stream);

/* Fixed-point Arpeggio first-accept annealer (mode 0 'arpeggio'; mode 1 'par', the engine's parallel pass).
 * Replaces `arpeggio_fx` (arpeggio/sim/arpeggio_fx.cu:35-37, same argument order and meaning; launched by
 * arpeggio/sim/arpeggiofx.py:23-44) and the 'par' mode of arpeggio/sim/local_code_a59/arpeggio_ref59.py:41-85.
 *   Gt    (B, D, D) int32, transposed Gram: Gt[b][j][i] = G_ij     h (B, D) int32      T (L) int32, one per pass
 *   q     pinning, Q units    P copies in {2, 4, 6}    R anneals per frame    1 <= D <= 128
 *   inst  (B) uint32 Philox instance words    k0, k1 = low, high 32 bits of the 64-bit key
 *   start 0 'rand', 1 'bal'
 *   out_s (B, R, P*D) int8 best spins, copy-major (spin p*D + j)   out_de (B, R) int64 best energy change (2 E0, Q)
 *   out_st (B, R, 2) int64 (flips, tests)
 * Fields and decision bounds are 32-bit here where the reference compares in 64 bits: 0 <= T < 2^29, and
 * max_i(|h_i| + 2 P sum_j |G_ij|) + 2 max|G_jj| + |q| + 2 max T < 2^31 (the front end's |G|, |h| <= Q keep this). */
int arp_anneal_arpeggio(const int32_t *Gt, const int32_t *h, const int32_t *T, int L, int q, int P, int R, int D,
                        int B, const uint32_t *inst, uint32_t k0, uint32_t k1, int start, int mode, int8_t *out_s,
                        int64_t *out_de, int64_t *out_st, void *stream);

/* The production engine's lane-shared Chord (chord_ref.chord_ref, arpeggio/sim/x3model/chord_ref.py:73-123; the
 * cycle model arpeggio/sim/x3model/engine_model.cpp computes the same spins and energies): the compressed dDI problem
 * (Gt (B, D, D) int32 transposed Gram = Gram, h (B, D)), P copies, L steps with T and Cq (Onsager c_k, may be NULL)
 * per step, pinning q; step k tests copy k mod P on one snapshot and commits every accepted lane. rng 1: the
 * xorshift32 lane streams of the production engine (seeded from k0, k1, inst, r by fmix32; RB = 16); rng 0: Philox
 * (then with Cq = NULL it equals arp_anneal_arpeggio's mode 1). list/count (both NULL, or both set): slot i < *count
 * anneals frame list[i] (problem and instance word at the frame's index, outputs at slot i), as
 * arp_anneal_arpeggio_list. Outputs as arp_anneal_arpeggio (out_st = flips, comparator passes). Replaces
 * engine_py.solve_batch (arpeggio/sim/x3model/engine_py.py:116-133) as the production-Chord solver of the MDI runs
 * (arpeggio/sim/x3_prod.py:97-112, run_a513.py:119-122); the engine's modeled cost per anneal is L + flips. */
int arp_anneal_chord_lane(const int32_t *Gt, const int32_t *h, const int32_t *T, const int32_t *Cq, int L, int q, int P,
                          int R, int D, int B, const uint32_t *inst, uint32_t k0, uint32_t k1, int rng,
                          const int32_t *list, const int32_t *count, int8_t *out_s, int64_t *out_de, int64_t *out_st,
                          void *stream);

/* Fixed-point Chord annealer. Replaces `chord_m{1,2,3,4,6,8}` (arpeggio/mimo/chord_kernel.cu:177-184, same argument
 * order and meaning; launched by Kernel.chord, arpeggio/mimo/chordfx.py:189-208), bit-exact to chordfx.chord_fx.
 *   J (B, N, N) int16 symmetric   h, h0 (B, N) int32   Tq, Cq (S) int32   q pinning   P phases (copies), N % P == 0
 *   R anneals   N <= 256, N % 32 == 0   k0, k1 seed words   inst0: frame b uses Philox instance inst0 + b
 *   mode 0: out_e (B, R) int32, out_s (B, R, N) int8;  mode 1: out_e (B,), out_s (B, N) (lowest energy, lowest r)
 *   stats: NULL or 4 uint64 accumulators (flips, energy evaluations, Philox blocks, steps x anneals) */
int arp_anneal_chord(const int16_t *J, const int32_t *h, const int32_t *h0, const int32_t *Tq, const int32_t *Cq,
                     int q, int S, int P, int R, int N, int B, uint32_t k0, uint32_t k1, int inst0, int mode,
                     int32_t *out_e, int8_t *out_s, uint64_t *stats, void *stream);

/* Fused front end, one workgroup per frame: LMMSE hard decision (for D, Mr <= 64 the inverse of
 * H^T H + (N0/2)/E[l^2] I by a four-pivot block sweep on the matrix cores, see arp_set_front_end_form; otherwise an
 * in-register symmetric sweep, one wave per frame for D <= 64, two for D <= 128),
 * delta-Ising build around it (dDI, 2K copies) on the matrix cores, and the fixed-point inputs of the annealers.
 * Replaces, on the kernel path, mimo.mmse / mmse_hard / _mu / quantize_levels (arpeggio/mimo/mimo.py:86-109, 176-182),
 * mimo.ising_delta_at (mimo.py:506-523), the J/h rounding of chordfx.to_fixed (arpeggio/mimo/chordfx.py:60-76) and
 * chord_ref.compress (arpeggio/sim/x3model/chord_ref.py:47-59).
 *   H (B, Mr, D) float32 real lift (Mr = 2 Nr, D = 2 Nt <= 128, Mr <= 128)    y (B, Mr) float32
 *   N0: (B) float32 per-frame noise density, or NULL to use N0_scalar for every frame    M: 4 (QPSK) or 16 (16-QAM)
 *   K: dDI radius (P = 2K copies)    Q: fixed-point scale (4096)
 * Outputs (each may be NULL): xg (B, D) int8 levels; Gi (B, D, D) int32 = rne(G / scale Q) (symmetric, so also the
 * transposed Gram the Arpeggio kernel takes); hi (B, D) int32; scale (B); Gf (B, D, D), hf (B, D) float32 G/scale,
 * h/scale; J (B, N, N) int16 and hN (B, N) int32, the expanded Chord inputs (N = 2 K D <= 256); z, mu (B, D);
 * Jf (B, N, N), hNf (B, N) float32, the expanded inputs of the FP32 Chord anneal; feat (B, feat_ld) bf16 (bit
 * pattern as uint16), the learned-start features [H^T y / g, G_jj / g - 1, x_g / lmax, z/mu - x_g (D each),
 * (N0/2) / g, zeros], g = mean_j G_jj, feat_ld >= 4 D + 1 (no reference counterpart); Gh (B, D, D) int16, the
 * same integers as Gi at half the bytes (|G_int| <= Q), which arp_anneal_arpeggio_list takes in place of Gt. */
int arp_front_end(const float *H, const float *y, const float *N0, float N0_scalar, int B, int Mr, int D, int M, int K,
                  float Q, int8_t *xg, int32_t *Gi, int32_t *hi, float *scale, float *Gf, float *hf, int16_t *J,
                  int32_t *hN, float *z, float *mu, float *Jf, float *hNf, uint16_t *feat, int feat_ld, int16_t *Gh,
                  void *stream);

/* Decode and select: candidate levels from anneal spins (delta = 1: clip(x_g + sum of the P copies); delta = 0: RI
 * spins), lowest ||y - H x||^2 over the R candidates then x_g (first minimum), Gray bit errors against ltrue.
 * Replaces mimo.spins_to_levels(_delta), mimo.dist, mimo.select, mimo.bit_errors (arpeggio/mimo/mimo.py:39-51,
 * 185-190, 416-431) and arpx.biterr_per_use (arpeggio/sim/arpx.py:63-66).
 *   spins (B, R, N) int8 (N = P D for delta = 1; D or 2 D for RI)    ltrue (B, D) int8 or NULL
 *   lhat (B, D) int8, biterr (B) int32 (needs ltrue), dist (B, R + 1) float32; each may be NULL. */
int arp_decode_select(const float *H, const float *y, const int8_t *xg, const int8_t *spins, int B, int R, int P, int D,
                      int Mr, int M, int delta, const int8_t *ltrue, int8_t *lhat, int32_t *biterr, float *dist,
                      void *stream);

/* FP32 Arpeggio anneal on float32 scaled couplings (Gt = G / scale transposed, h = h / scale, T and q scaled): the
 * float twin of arp_anneal_arpeggio, acceptance z < ((float)x 2^-32 - 1/2) 4T on the same Philox words. The
 * reference's float first-accept solver is arpeggio/sim/arpx55.py:29-99 (rule 'r', torch generator draws).
 *   out_de (B, R) float32 best energy change in E0 units (-h.s - s^T J s)   out_w (B, R, D) final fields or NULL */
int arp_anneal_arpeggio_f32(const float *Gt, const float *h, const float *T, int L, float q, int P, int R, int D,
                            int B, const uint32_t *inst, uint32_t k0, uint32_t k1, int start, int8_t *out_s,
                            float *out_de, int64_t *out_st, float *out_w, void *stream);

/* FP32 Chord anneal on float32 symmetric J (B, N, N), h, h0 (B, N), T, c (S) float32: the float twin of
 * arp_anneal_chord with Philox draws (the reference's float Chord is mimo.chord, arpeggio/mimo/mimo.py:204-249).
 *   out_e (B, R) best E0, out_s (B, R, N) int8, out_u (B, R, N) final fields or NULL, out_fl (B, R) flips or NULL */
int arp_anneal_chord_f32(const float *J, const float *h, const float *h0, const float *T, const float *c, float q, int S,
                         int P, int R, int N, int B, uint32_t k0, uint32_t k1, int inst0, float *out_e, int8_t *out_s,
                         float *out_u, int64_t *out_fl, void *stream);

/* Fixed-point first-accept anneal from a learned start (no reference counterpart; the reference's start is the random
 * one of arpeggio_fx.cu:48-51, kept here as the fallback). Arguments as arp_anneal_arpeggio, plus
 *   thr (B, P*D) uint32 per-spin start thresholds: spin +1 iff x + thr >= 2^32, x the random start's Philox word
 *       (thr = 2^31 reproduces the random start exactly)
 *   gate 1: per anneal, keep the learned start only if its Ising energy is not above the random start's
 *   out_used (B, R) int8 or NULL: 1 where the learned start was kept.
 * Same integer ranges as arp_anneal_arpeggio. */
int arp_anneal_arpeggio_ws(const int32_t *Gt, const int32_t *h, const int32_t *T, int L, int q, int P, int R, int D,
                           int B, const uint32_t *inst, uint32_t k0, uint32_t k1, const uint32_t *thr, int gate,
                           int8_t *out_s, int64_t *out_de, int64_t *out_st, int8_t *out_used, void *stream);

/* One class of a mixed batch (frames of one size D): device buffers with the layouts of arp_anneal_arpeggio, that
 * class's schedule T (L int32), q, start (0 rand, 1 bal) and `tile` = frames per unit of work (their staged Grams must
 * fit the workgroup's LDS together: tile * (D*D + 64*ceil(D/64)) * 4 bytes). thr (B, 2D) uint32 or NULL: the learned
 * start of arp_anneal_arpeggio_ws for this class (with the energy gate `gate`, kept flags to out_used (B, R) int8 or
 * NULL); NULL: the start `start`. apw: anneals per wave64 for a class of D <= 32 (2: two per wave, 1: one, 0: the
 * thread's arp_set_anneals_per_wave choice); ignored above D = 32. */
typedef struct {
  const int32_t *Gt;
  const int32_t *h;
  const uint32_t *inst;
  const int32_t *T;
  int8_t *out_s;
  int64_t *out_de;
  int64_t *out_st;
  int D, B, L, q, start, tile;
  const uint32_t *thr;
  int8_t *out_used;
  int gate;
  int apw;
} arp_batch_class;

/* Frames of uneven size in one persistent fixed-point Arpeggio launch (P = 2, first accept): `workgroups`
 * workgroups of `waves` wave64s pull tiles of frames from a device counter, in the order the classes are given (put
 * the largest first). Each anneal equals arp_anneal_arpeggio's on the same class buffers, stream for stream. The
 * reference launches one fixed shape per call (arpeggio_fx.cu:35-37); packing several shapes into one launch has no
 * counterpart there. counter: one device uint32 of scratch (cleared on the stream before the launch).
 * n_classes <= 8, D <= 128 per class, waves 1..16. Same integer ranges as arp_anneal_arpeggio per class. */
int arp_anneal_arpeggio_batched(const arp_batch_class *classes, int n_classes, int R, uint32_t k0, uint32_t k1,
                                uint32_t *counter, int workgroups, int waves, void *stream);

/* One class of a mixed batch for the whole-chain entry point: nt users x nt rx, constellation M, B frames. */
typedef struct arp_batch_frames {
  const float *H;         /* (B, 2nt, 2nt) real lift, float32 */
  const float *y;         /* (B, 2nt) */
  const float *N0;        /* (B) per-frame noise density, or NULL: N0_scalar */
  float N0_scalar;
  const uint32_t *inst;   /* (B) Philox instance words (the frames' arrival indices) */
  const int8_t *ltrue;    /* (B, 2nt) true levels, or NULL */
  int8_t *lhat;           /* (B, 2nt) decided levels (out) */
  int32_t *biterr;        /* (B) bit errors (out; needs ltrue) or NULL */
  int nt, M, B;
} arp_batch_frames;

/* Workspace bytes arp_detect_batched and arp_tune_batched need for these classes and R anneals per frame (0 for
 * invalid arguments). */
size_t arp_detect_batched_workspace_bytes(const arp_batch_frames *classes, int n_classes, int R);

/* The mixed batch's whole step in one call: per class the front end (LMMSE, dDI build, fixed point Q; per-frame N0),
 * one persistent first-accept anneal over every class (arp_anneal_arpeggio_batched: R anneals of L passes per frame,
 * schedule T (L) and q in Q units, P = 2, random start, Philox key k0/k1), per class decode and bit errors. Launch
 * shape: `workgroups` workgroups of `waves` waves, frames per tile from `lds_budget` bytes (arp_tune_batched picks
 * them); anneals per wave of the classes of D <= 32: the thread's arp_set_anneals_per_wave choice. Classes in the
 * order given (largest first packs best), at most 8, nt <= 64. Stream-ordered with no host
 * synchronisation and no memset, so a caller may capture it in a HIP graph. No reference counterpart (the reference
 * detects one shape per call, arpeggio/mimo/fxdetect.py:24-43). */
int arp_detect_batched(const arp_batch_frames *classes, int n_classes, const int32_t *T, int L, int q, int R, float Q,
                       uint32_t k0, uint32_t k1, int workgroups, int waves, int lds_budget, void *work,
                       size_t work_bytes, void *stream);

/* Launch-shape autotuner of the persistent anneal: runs the front ends once, then times `reps` persistent launches
 * for every candidate (LDS budget in {one frame of the largest class, twice that, 80 KiB, 160 KiB}, waves in {4, 8,
 * 16}, workgroups = CUs x workgroups resident per CU, anneals per wave in {1, 2} when a class has D <= 32) with HIP
 * events on the stream, and writes the fastest as best_shape = {workgroups, waves, lds_budget, anneals_per_wave}
 * (4 ints) and its time in ms. Synchronises the stream: not for graph capture. */
int arp_tune_batched(const arp_batch_frames *classes, int n_classes, const int32_t *T, int L, int q, int R, float Q,
                     uint32_t k0, uint32_t k1, int reps, void *work, size_t work_bytes, int *best_shape, float *best_ms,
                     void *stream);

/* ---- Adaptive (early-stop) detection: a short anneal for every frame, a certificate, and a longer anneal for the
 * frames the certificate leaves open (no reference counterpart for the two-stage form; the stopping rule is the
 * reference's first-pass stop t = ||y - H x||^2 / N0 <= Nr + c sqrt(Nr), arpeggio/sim/a56an.py:58-67). */

/* First-accept anneal over a frame list: workgroup slot i < *count anneals frame list[i] (arrays Gt, h, inst, thr
 * indexed by frame) with Philox streams r0 .. r0 + R - 1 and writes out_s (slots, R, P D), out_de (slots, R),
 * out_st (slots, R, 2), out_used (slots, R) at slot i; slots at or past *count do nothing. B = number of slots
 * launched (the list's capacity). list = count = NULL: every frame b < B in its own slot (then r0 = 0 equals
 * arp_anneal_arpeggio / arp_anneal_arpeggio_ws). thr = NULL: the start `start` (0 rand, 1 bal); else the learned start
 * with the energy gate `gate` (as arp_anneal_arpeggio_ws). Anneal r of a frame is the same stream in every launch, so
 * r0 = R1 continues a frame's R1 anneals of an earlier launch with new ones. Mode: first accept. The Gram comes as
 * Gt (int32) or as Gh (int16, the front end's Gh output): exactly one of the two is given. */
int arp_anneal_arpeggio_list(const int32_t *Gt, const int16_t *Gh, const int32_t *h, const int32_t *T, int L, int q,
                             int P, int R, int D, int B, const uint32_t *inst, uint32_t k0, uint32_t k1, int start,
                             const uint32_t *thr,
                             int gate, int r0, const int32_t *list, const int32_t *count, int8_t *out_s,
                             int64_t *out_de, int64_t *out_st, int8_t *out_used, void *stream);

/* Decode and select over a frame list (list/count as arp_anneal_arpeggio_list; NULL: every frame), with an optional
 * prior candidate per frame (prior_l (B, D) int8 levels, prior_d (B) float32 distance) that the best new candidate
 * must beat strictly. Writes lhat, biterr and best_d (the chosen candidate's ||y - H x||^2; each may be NULL) at the
 * frame's index. Spins are read from the slot. Replaces the same reference functions as arp_decode_select. */
int arp_decode_select_list(const float *H, const float *y, const int8_t *xg, const int8_t *spins, int B, int R, int P,
                           int D, int Mr, int M, int delta, const int8_t *ltrue, const int32_t *list,
                           const int32_t *count, const int8_t *prior_l, const float *prior_d, int8_t *lhat,
                           int32_t *biterr, float *best_d, void *stream);

/* Certificate: frames b < B with best_d[b] > N0_b * thr_scale (N0_b = N0[b], or N0_scalar if N0 is NULL; thr_scale =
 * Nr + c sqrt(Nr)) are written to list (order within a wave by frame, across waves unspecified); *count = their
 * number (cleared on the stream first). Graph-capturable: no memset node. */
int arp_certify(const float *best_d, const float *N0, float N0_scalar, float thr_scale, int B, int32_t *list,
                int32_t *count, void *stream);

/* Warm-start MLP weights for the whole-chain entry point (bf16 as uint16 bit patterns, n-major as
 * arp_mlp_start_fused takes them): F = the feature row (4 D + 1 rounded up to a multiple of 32), H = 256, N = 2 D. */
typedef struct {
  const uint16_t *W1t;
  const float *b1;
  const uint16_t *W2t;
  const float *b2;
  const uint16_t *W3t;
  const float *b3;
  int F, H;
  float inv_temp;
} arp_mlp_weights;

/* Bytes of workspace arp_detect_adaptive needs for B frames of D coordinates, R1 stage-1 and R2 stage-2 anneals per
 * frame (learned = 1 with the warm-start MLP); 0 for invalid sizes. */
size_t arp_detect_workspace_bytes(int B, int D, int R1, int R2, int learned);

/* The whole detection chain in one call (dDI K = 1, first accept), replacing the reference's host sequence
 * fxdetect.detect (arpeggio/mimo/fxdetect.py:24-43: MMSE, mapping, to_fixed, kernel, levels, select) for the
 * first-accept solver: front end -> [warm-start MLP (mlp != NULL) -> start thresholds, energy gate `gate`] -> R1
 * anneals of L1 passes (schedule T1) -> decode/select -> if R2 > 0: certificate d <= N0 thr_scale -> R2 more anneals
 * (streams R1..R1+R2-1) of L2 passes (T2) on the open frames -> decode against stage 1's choice. R2 = 0 is the one-stage
 * detector. H (B, Mr, D), y (B, Mr) float32; N0 (B) or NULL (N0_scalar); inst (B) Philox instance words; k0, k1 the
 * key; q pinning in Q units; ltrue (B, D) int8 or NULL; outputs lhat (B, D) int8 and biterr (B) int32 (needs
 * ltrue; may be NULL). work: device memory of at least arp_detect_workspace_bytes(B, D, R1, R2, mlp != NULL) bytes,
 * owned by the caller; stage buffers are carved from it (256-byte aligned), so two calls that overlap on one stream
 * order must not share it. Asynchronous on `stream`, graph-capturable (no allocation, no memset node, no host sync). */
int arp_detect_adaptive(const float *H, const float *y, const float *N0, float N0_scalar, int B, int Mr, int D, int M,
                        const int32_t *T1, int L1, int R1, const int32_t *T2, int L2, int R2, int q, float thr_scale,
                        const uint32_t *inst, uint32_t k0, uint32_t k1, const arp_mlp_weights *mlp, int gate,
                        const int8_t *ltrue, int8_t *lhat, int32_t *biterr, void *work, size_t work_bytes,
                        void *stream);

/* arp_detect_adaptive with no workspace argument and no allocation: the stage buffers live in a device array of the
 * library (32 MiB per device, taken with hipGetSymbolAddress on the first call on each device). No reference
 * counterpart: for callers that cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation). A shape whose arp_detect_workspace_bytes exceeds 32 MiB is refused with ARP_E_RANGE (about
 * 1,500 frames of 32 users at R1 + R2 = 20 fit). Every call on a device shares its array: calls that overlap on one
 * device (other streams, other host threads) race and must be serialised by the caller. */
int arp_detect_adaptive_noalloc(const float *H, const float *y, const float *N0, float N0_scalar, int B, int Mr, int D,
                                int M, const int32_t *T1, int L1, int R1, const int32_t *T2, int L2, int R2, int q,
                                float thr_scale, const uint32_t *inst, uint32_t k0, uint32_t k1,
                                const arp_mlp_weights *mlp, int gate, const int8_t *ltrue, int8_t *lhat,
                                int32_t *biterr, void *stream);

/* ---- Multi-stage delta-Ising (MDI) wrapper kernels (the host groups the sub-problems by radius, as the reference's
 * mimo.mdi, arpeggio/mimo/mimo.py:546-602, and run_a513.mdi_logged, arpeggio/sim/local_code_a513/run_a513.py:43-80). */

/* Ordered MMSE-SIC guess (arpeggio/mimo/mimo.py:466-503): per stage the stream (per_user = 0) or user (per_user = 1)
 * with the smallest MMSE error variance, its bias-corrected estimate quantised and cancelled. H (B, Mr, D), y (B, Mr)
 * float32, N0 (B) or NULL (N0_scalar); xs (B, D) int8 levels. D even, <= 128; Mr <= 128. */
int arp_mmse_sic(const float *H, const float *y, const float *N0, float N0_scalar, int B, int Mr, int D, int M,
                 int per_user, int8_t *xs, void *stream);

/* Search radius of a guess (arpeggio/mimo/mimo.py:526-543, 'mean' heuristic) and the guess's ||y - H x_g||^2:
 * R (B) int32 in [0, Rmax], Gam (B) float32, dist (B) float32 (each may be NULL). */
int arp_mdi_radius(const float *H, const float *y, const int8_t *xg, int B, int Mr, int D, int M, int Rmax, int32_t *R,
                   float *Gam, float *dist, void *stream);

/* Fixed-point dDI sub-problems around the guesses xg (B, D) of the n frames list[0..n): Gi (B, D, D) and hi (B, D) int32
 * written at each listed frame's index, equal to mimo.ising_delta_at -> chordfx.to_fixed -> chord_ref.compress
 * (mimo.py:506-523, chordfx.py:60-76, chord_ref.py:47-59) up to float32 rounding; Q the fixed-point scale (4096). */
int arp_mdi_build(const float *H, const float *y, const int8_t *xg, const int32_t *list, int n, int Mr, int D, float Q,
                  int32_t *Gi, int32_t *hi, void *stream);

/* Learned-start MLP kernels (no reference counterpart). bf16 tensors are passed as their uint16 bit patterns.
 * arp_mlp_gemm: C(m, n) = sum_k A(m, k) B(k, n) (+ bias[n]), act 0 none / 1 GELU (tanh form), on v_mfma_f32_16x16x32_bf16
 *   with fp32 accumulation; A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn]; outputs (each may be NULL): Cf
 *   fp32 after the activation, Cpre fp32 before it, Cb bf16 after it, row stride ldc. */
int arp_mlp_gemm(const uint16_t *A, long sam, long sak, const uint16_t *B, long sbk, long sbn, int M, int N, int K,
                 const float *bias, int act, float *Cf, float *Cpre, uint16_t *Cb, long ldc, void *stream);
/* dA = dH * GELU'(Apre) over n elements, to bf16 and/or fp32. */
int arp_mlp_gelu_bwd(const float *dH, const float *Apre, long n, uint16_t *dAb, float *dAf, void *stream);
/* Mean binary cross-entropy of logits z (M x N, row stride ld) against spins y (int8, +-1, M x N contiguous):
 * gradient (sigmoid(z) - (y+1)/2) / (M N) to bf16 and/or fp32; *loss += the mean loss (atomic). */
int arp_mlp_bce(const float *z, long ld, const int8_t *y, int M, int N, uint16_t *dzb, float *dzf, float *loss,
                void *stream);
/* out[n] = sum over m of X[m*ld + n] (bias gradients). */
int arp_mlp_colsum(const float *X, long ld, int M, int N, float *out, void *stream);
/* One Adam step (bias-corrected, step >= 1) on fp32 master weights W with gradient G and moments m, v; Wb (or NULL)
 * receives the bf16 copy. */
int arp_mlp_adam(float *W, const float *G, float *m, float *v, long n, float lr, float b1, float b2, float eps, int step,
                 uint16_t *Wb, void *stream);
/* fp32 -> bf16, round to nearest even. */
int arp_cast_bf16(const float *x, long n, uint16_t *y, void *stream);
/* Fused inference of the warm-start MLP (F -> H = 256 -> 256 -> N, GELU): features X (M x F bf16, F a multiple of 32
 * up to 1016) and n-major bf16 weights W1t (H x F), W2t (H x H), W3t (N x H) with fp32 biases -> start thresholds
 * (M x N uint32, N in {64, 128, 192, 256}), equal bit for bit to three arp_mlp_gemm calls and arp_start_thresholds. */
int arp_mlp_start_fused(const uint16_t *X, int M, int F, const uint16_t *W1t, const float *b1, const uint16_t *W2t,
                        const float *b2, const uint16_t *W3t, const float *b3, int H, int N, float inv_temp,
                        uint32_t *thr, void *stream);
/* Start thresholds from logits: thr = rint(sigmoid(z * inv_temp) 2^32) clamped to [1, 2^32 - 1]. */
int arp_start_thresholds(const float *z, long ld, int M, int N, float inv_temp, uint32_t *thr, void *stream);

/* RI-MIMO Ising build in fixed point: mimo.ising (arpeggio/mimo/mimo.py:132-149) and chordfx.to_fixed's J, h, h0
 * (arpeggio/mimo/chordfx.py:60-76), the problem fxdetect.detect anneals for its 'ri' schedules (fxdetect.py:17-18,
 * 28-30). H (B, Mr, D), y (B, Mr), z (B, D) the LMMSE output (float32). Spins: N = D (QPSK) or N = 2D (16-QAM, MSB
 * block first; l = 2 s_MSB + s_LSB). Out: J (B, N, N) int16 = rne(-offdiag(T^T H^T H T) / scale Q), h (B, N) int32 =
 * rne((h0 + 2 r s_P) Q), h0 (B, N) int32 = rne(2 T^T H^T y / scale Q), scale (B) float32 or NULL; s_P are the spins of
 * z's nearest levels (ties to the lower). D, Mr <= 128, N <= 256. */
int arp_ising_ri(const float *H, const float *y, const float *z, int B, int Mr, int D, int M, float r, float Q,
                 int16_t *J, int32_t *h, int32_t *h0, float *scale, void *stream);

/* ---- Soft output: max-log bit LLRs, positive favouring bit 1, (B, D nb) dimension-major with the MSB first (nb = 2
 * for 16-QAM, 1 for QPSK), as arpeggio/mimo/softlib.py:1-9 defines them. */

/* MMSE LLRs from the front end's z and mu (B, D) float32: z = mu l + e with Var(e) = E[l^2] mu (1 - mu); per bit,
 * (min over levels with the bit 0 - min with it 1) of (z - mu l)^2 / max(Var, 1e-12), halved. Replaces
 * mimo.mmse_soft_llr (arpeggio/mimo/mimo.py:111-129). */
int arp_soft_mmse_llr(const float *z, const float *mu, int B, int D, int M, float *llr, void *stream);

/* List LLRs: the candidates of arp_decode_select (R anneal outputs formed from spins as there, then the MMSE levels
 * xg), the decision (first minimum of ||y - H x||^2, written to lhat (B, D) int8 if not NULL; the C = R + 1 distances
 * to dist (B, C) if not NULL), and the decision's D nb single-Gray-bit neighbours: LLR = (min d with the bit 0 -
 * min d with it 1) / N0 clipped to +-clip, N0 per frame (N0, (B)) or N0_scalar. D, Mr <= 128, R <= 63. Replaces
 * mimo.select + mimo.list_llr (arpeggio/mimo/mimo.py:421-425, 434-462), the list detectors of softlib.detect_llr
 * (arpeggio/mimo/softlib.py:33-65). */
int arp_soft_list_llr(const float *H, const float *y, const int8_t *xg, const int8_t *spins, int B, int R, int P, int D,
                      int Mr, int M, int delta, const float *N0, float N0_scalar, float clip, int8_t *lhat,
                      float *dist, float *llr, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ARPEGGIO_HIP_H */
