#!/usr/bin/env python3
"""Detections/s at fixed BER on one MI355X, learned-start vs plain annealer: 32 users x 32 rx, 16-QAM, i.i.d.
Rayleigh, 65,536 frames per step, one-stage dDI (K = 1).

A step is one pass of the kernel path over the batch with H, y and the true levels already in HBM.
  plain:    front end (LMMSE + dDI Ising build + fixed point) -> fixed-point Arpeggio (first accept), 16 anneals x 64
            passes from the random start -> decode/select -> bit errors (the reference's 32-user pick
            abl_far_32_0.1_0.1_0_rand_*: 32 sweeps x 2 copies, T = 0.1, q = 0)
  learned:  the adaptive detector: front end with the warm-start features -> MLP -> start thresholds -> stage 1,
            learned-start anneals (budget picked in step 3) for every frame -> decode -> certificate ||y - Hx||^2 <= N0 (Nr + c
            sqrt(Nr)) (the reference's first-pass stop rule, arpeggio/sim/a56an.py:58-67) -> stage 2, the plain
            budget (16 anneals x 64 passes, new streams) for the open frames only -> decode against stage 1's choice

Procedure (fixed-BER comparison):
  1. the SNR where plain reaches BER 1e-3 is found on calibration frames (log-BER interpolated between bracketing
     SNRs; --snr skips this);
  2. c is picked by the reference's rule (a56an.py:109-118) on the calibration frames at that SNR: the largest c of
     its grid whose paired-bootstrap BER ratio to the same detector without the stop has an upper 95 % bound <= 1.10;
  3. stage 1's budget (anneals x passes) on the same frames: the fastest of S1_GRID whose paired BER ratio to plain
     has an upper 95 % bound <= 1.10; then stage 2's anneal count by the same rule (S2_GRID);
  4. on the measurement frames, the learned detector's BER must match plain's: paired-bootstrap ratio (a56an.py:74-95)
     with an upper 95 % bound <= 1.10 (the reference's equivalence cut-off, README.md:171);
  5. both pipelines are timed over K steps; value = learned detections/s, plain.value and speedup beside it, and the
     same adaptive detector from the random start (what the network adds) and the learned start without the stop.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--snr DB] [--frames B] [--no-cpu] [--with-chord]

Extra fields: roofline (the timed step's dominant kernel, HIP events on its own stream), stages (per-stage ms),
cpu_baseline (the C oracle of the plain chain on this host's cores, three timed segments)."""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X: 256 CUs x 4 SIMDs at 2.4 GHz (rocminfo); a wave64 vector instruction holds a SIMD for 4 clocks. LDS read
# bandwidth: measured by arpeggio_amd.probes (profiles/chip_probes.json: conflict-free 16-byte reads from every CU),
# 128 B per clock per CU where that record is absent
SIMDS, CLOCK_HZ = 256 * 4, 2.4e9


def _lds_peak():
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles', 'chip_probes.json')
    if os.path.exists(p):
        return float(json.load(open(p))['lds']['GBs'])
    return 128 * 256 * CLOCK_HZ / 1e9


LDS_PEAK_GBS = _lds_peak()
MFMA_FP32_TFLOPS = 157.3
NT, M = 32, 16
METRIC = 'detections/s at fixed BER (32u 16-QAM), learned-start vs plain annealer'


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--frames', type=int, default=65536)
    ap.add_argument('--calib-frames', type=int, default=32768)
    ap.add_argument('--snr', type=float, default=None, help='skip the calibration and measure at this SNR (dB)')
    ap.add_argument('--target-ber', type=float, default=1e-3)
    ap.add_argument('--copies', type=int, default=16)
    ap.add_argument('--passes', type=int, default=64, help='plain annealer pass budget')
    ap.add_argument('--stage1', default=None, help='adaptive stage 1: anneals,passes (default: picked on calibration '
                    'frames, the fastest equivalent budget of S1_GRID)')
    ap.add_argument('--c', type=float, default=None, help='certificate constant (default: picked on calibration '
                    'frames by the reference rule)')
    ap.add_argument('--weights', default=None)
    ap.add_argument('--seed', type=int, default=1)
    ap.add_argument('--plain-only', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--warm-seconds', type=float, default=1.0, help='minimum warm-up time before each timed loop')
    ap.add_argument('--with-chord', action='store_true')
    ap.add_argument('--no-batched', action='store_true', help='skip the mixed-size batch measurement')
    ap.add_argument('--batched-frames', type=int, default=1 << 17)
    ap.add_argument('--no-64u', action='store_true', help='skip the 64-user correlated-channel measurement')
    ap.add_argument('--frames-64u', type=int, default=1 << 16)
    ap.add_argument('--sweep-frames', type=int, default=16384, help='frames per SNR point of the BER sweeps')
    ap.add_argument('--no-sweep', action='store_true', help='skip the 32-user BER sweep over 12-22 dB')
    ap.add_argument('--copy-probe', action='store_true', help='run only a plain device copy of known size (PMC scale)')
    ap.add_argument('--no-mdi', action='store_true', help='skip the MDI wrapper against the reference fixtures')
    ap.add_argument('--mdi-batches', type=int, default=60, help='32-user fixture batches (1,000 frames each)')
    ap.add_argument('--mdi-large', type=int, default=16384, help='frames of the MDI throughput batch')
    ap.add_argument('--gate', action='store_true',
                    help='headline learned start with the energy gate (fallback to the random start per anneal); '
                         'the gated detector is measured beside it either way')
    ap.add_argument('--stage2-copies', type=int, default=None,
                    help='adaptive stage 2 anneals (default: picked on calibration frames by the equivalence rule)')
    ap.add_argument('--front-end-form', type=int, default=2, choices=(0, 1, 2),
                    help='2: the complex-lift front end (the frames are real lifts of complex channels), 0: general')
    return ap.parse_args()


def log(*a):
    print('#', *a, file=sys.stderr, flush=True)


def copy_probe():
    import torch
    n = 256 << 20
    a = torch.empty(n, dtype=torch.uint8, device='cuda')
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps(dict(copy_probe_bytes=n, launches=3)))


def key_of(seed):
    return (seed * 0x9E3779B97F4A7C15 + 0xA513) & 0xFFFFFFFFFFFFFFFF


def make_pipe(B, args, L, start, model=None):
    from arpeggio_amd.detector import DetectPipeline
    return DetectPipeline(B, 2 * NT, 2 * NT, M, args.copies, policy='arpeggio', cycles=L, T0=0.1, T1=0.1, q=0.0,
                          start=start, key=key_of(args.seed), inst0=0, model=model, gate=args.gate)


def frame_errors(pipe, b):
    import torch
    lt = b['l'].to(torch.int8).contiguous()
    _, be = pipe.run(b['H'], b['y'], b['N0'], lt)
    return be.cpu().numpy().astype(np.int64)


def calibrate(args):
    """SNR (dB) where the plain annealer's BER crosses the target, on calibration frames (draw index 1)."""
    from arpeggio_amd import frames
    Bc = args.calib_frames
    pipe = make_pipe(Bc, args, args.passes, 'rand')
    pts = []
    for snr in (22.0, 23.0, 24.0, 25.0, 26.0):
        b = frames.draw(args.seed, snr, 1, Bc, NT, M, dev='cuda')
        ber = frame_errors(pipe, b).sum() / (Bc * 4 * NT)
        log('calibration: plain BER %.3e at %.1f dB' % (ber, snr))
        pts.append((snr, ber))
        if ber <= args.target_ber:
            break
    hi = [p for p in pts if p[1] > args.target_ber]
    lo = [p for p in pts if 0 < p[1] <= args.target_ber]
    if not hi or not lo:
        raise RuntimeError('calibration did not bracket BER %g: %s' % (args.target_ber, pts))
    (s0, b0), (s1, b1) = hi[-1], lo[0]
    t = (math.log10(b0) - math.log10(args.target_ber)) / (math.log10(b0) - math.log10(b1))
    return round(s0 + t * (s1 - s0), 2), pts


C_GRID = (0.5, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0)   # the reference's grid of stop constants (a56an.py:10)
S2_GRID = (16, 12, 8)                               # stage-2 anneal counts tried, the plain count first
# stage-1 budgets (anneals, passes) tried; every pass count is a whole number of Philox blocks (8 passes at P = 2)
S1_GRID = ((2, 16), (2, 32), (3, 24), (3, 32), (4, 16), (4, 24), (4, 32), (4, 40), (6, 24))
S1_DEFAULT = '4,24'                                 # stage 1 while the stop constant is picked
EQUIV_LIMIT = 1.10                             # the reference's BER-equivalence cut-off (README.md:171)


def make_adaptive(B, args, start, model=None, c=3.0, r2=None, gate=None):
    from arpeggio_amd.detector import AdaptivePipeline
    R1, L1 = (int(x) for x in args.stage1.split(','))
    return AdaptivePipeline(B, 2 * NT, 2 * NT, M, R1=R1, L1=L1, R2=r2 or args.copies, L2=args.passes, c=c, start=start,
                            T0=0.1, T1=0.1, q=0.0, key=key_of(args.seed), inst0=0,
                            model=model if start == 'learned' else None,
                            gate=args.gate if gate is None else gate)


def pick_c(args, snr, model):
    """The reference's pick_c (arpeggio/sim/a56an.py:109-118) on calibration frames (draw index 1) at the measurement
    SNR: the largest c of its grid whose paired BER ratio to the same detector without the stop (every frame takes
    stage 2) has an upper 95 % bound <= 1.10."""
    from arpeggio_amd import frames
    from arpeggio_amd.stats import boot_ratio
    Bc = args.calib_frames
    b = frames.draw(args.seed, snr, 1, Bc, NT, M, dev='cuda')
    pipe = make_adaptive(Bc, args, 'learned', model, c=0.0)
    pipe.thr_scale = -1.0                       # no stop: every frame is open
    base = frame_errors(pipe, b)
    rows, best = [], None
    for c in C_GRID:
        pipe.c = c
        pipe.thr_scale = NT + c * NT ** 0.5
        be = frame_errors(pipe, b)
        r, lo, hi = boot_ratio(be, base)
        rows.append(dict(c=c, ratio=r, ci=[lo, hi], open=int(pipe.count.item()) / Bc))
        log('stop constant c = %.1f: BER ratio to no stop %.3f [%.3f, %.3f], open %.3f' % (c, r, lo, hi,
                                                                                          rows[-1]['open']))
        if hi is not None and hi <= EQUIV_LIMIT:
            best = c
    return best, rows


def pick_stage1(args, snr, model, c):
    """Stage 1's budget (anneals x passes, S1_GRID) on the calibration frames of pick_c: the fastest learned adaptive
    detector (stop constant c, stage 2 at the plain budget) whose paired BER ratio to the plain detector has an upper
    95 % bound <= 1.10 (README.md:171). Each candidate's step is timed on those frames after a warm-up."""
    from arpeggio_amd import frames
    from arpeggio_amd.stats import boot_ratio
    Bc = args.calib_frames
    b = frames.draw(args.seed, snr, 1, Bc, NT, M, dev='cuda')
    base = frame_errors(make_pipe(Bc, args, args.passes, 'rand'), b)
    keep = args.stage1
    rows, best = [], None
    for r1, l1 in S1_GRID:
        args.stage1 = '%d,%d' % (r1, l1)
        pipe = make_adaptive(Bc, args, 'learned', model, c)
        be = frame_errors(pipe, b)
        r, lo, hi = boot_ratio(be, base)
        ms = time_pipe(pipe, b, 10, 2, 0.3)['ms_per_step']
        ok = hi is not None and hi <= EQUIV_LIMIT
        rows.append(dict(anneals=r1, passes=l1, ratio=r, ci=[lo, hi], open=int(pipe.count.item()) / Bc, ms=ms,
                         equivalent=bool(ok)))
        log('stage 1 %d x %d: BER ratio to plain %.3f [%.3f, %.3f], %.3f ms per %d frames' % (r1, l1, r, lo, hi, ms,
                                                                                              Bc))
        del pipe
        if ok and (best is None or ms < best[1]):
            best = (args.stage1, ms)
    args.stage1 = best[0] if best else keep
    return args.stage1, rows


def pick_stage2(args, snr, model, c):
    """Stage 2's anneal count by the reference's equivalence rule (README.md:171; a56an.py:74-95) on the calibration
    frames of pick_c: the smallest count of S2_GRID whose learned adaptive detector (stop constant c, stage 1 as
    configured, stage 2 at the plain pass budget) has a paired BER ratio to the plain detector with an upper 95 % bound
    <= 1.10; the grid is walked from the plain count down and stops at the first miss."""
    from arpeggio_amd import frames
    from arpeggio_amd.stats import boot_ratio
    Bc = args.calib_frames
    b = frames.draw(args.seed, snr, 1, Bc, NT, M, dev='cuda')
    base = frame_errors(make_pipe(Bc, args, args.passes, 'rand'), b)
    rows, best = [], None
    for r2 in S2_GRID:
        pipe = make_adaptive(Bc, args, 'learned', model, c, r2)
        be = frame_errors(pipe, b)
        r, lo, hi = boot_ratio(be, base)
        rows.append(dict(copies=r2, ratio=r, ci=[lo, hi]))
        log('stage 2 with %d anneals: BER ratio to plain %.3f [%.3f, %.3f]' % (r2, r, lo, hi))
        del pipe
        if hi is None or hi > EQUIV_LIMIT:
            break
        best = r2
    # no count passes on these frames (the interval of the plain count itself can straddle the cut-off): the plain
    # count, whose equivalence the measurement frames then test
    return best or args.copies, rows


def stage_ms(pipe, b, reps=5):
    """Mean per-stage times (HIP events on the launch stream between the stages) over reps steps."""
    import torch
    lt = b['l'].to(torch.int8).contiguous()
    n = len(pipe.STAGES) + 1
    acc = np.zeros(n - 1)
    for _ in range(reps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
        pipe.run(b['H'], b['y'], b['N0'], lt, marks=ev)
        torch.cuda.synchronize()
        acc += [ev[i].elapsed_time(ev[i + 1]) for i in range(n - 1)]
    return {k: float(v) for k, v in zip(pipe.STAGES, acc / reps)}


def time_pipe(pipe, b, steps, warmup, min_warm_s=1.0):
    """K timed steps after at least W warm-up steps and at least min_warm_s seconds of warm-up work (the GPU's clocks
    settle under load); barrier = synchronize on both sides. Returns det/s, ms/step, mean anneal kernel ms (HIP
    events on the launch stream), BER over the timed steps."""
    import torch
    lt = b['l'].to(torch.int8).contiguous()
    H, y, N0 = b['H'], b['y'], b['N0']
    # the warm-up runs exactly the timed loop's work (the error sum included), so every kernel the loop uses is
    # loaded before the clock starts
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    errs = torch.zeros((), dtype=torch.int64, device='cuda')
    t_w, n_w = time.perf_counter(), 0
    while n_w < warmup or time.perf_counter() - t_w < min_warm_s:
        _, be = pipe.run(H, y, N0, lt, events=ev[0])
        errs += be.sum()
        torch.cuda.synchronize()
        n_w += 1
    errs.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g0.record()
    for i in range(steps):
        _, be = pipe.run(H, y, N0, lt, events=ev[i])
        errs += be.sum()
    g1.record()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    B = pipe.B
    return dict(value=B * steps / dt, ms_per_step=dt / steps * 1e3, gpu_ms_per_step=g0.elapsed_time(g1) / steps,
                anneal_ms=float(np.mean([a.elapsed_time(c) for a, c in ev])),
                ber=float(errs.item()) / (B * 2 * pipe.D * steps))


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline(b, pipe, args, seg_s=6.0, nseg=3):
    """The oracle's plain chain on the host: numpy float64 front end, the C fixed-point Arpeggio (one thread per core of
    the GPU's CPU share: 16 on the one-GPU box) on the GPU front end's integers, numpy decode/select. Timed as nseg
    segments of at least seg_s seconds each, each segment detecting chunks of 2,048 of the step's frames in turn (the
    reference's segment rule, arpeggio/mimo/bench_b.py:63-90); value = detections / wall time over all segments."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    cores = min(16, os.cpu_count() or 1)
    chunk = 2048
    H = b['H'].cpu().numpy()
    y = b['y'].cpu().numpy()
    G = pipe.Gi.cpu().numpy()
    h = pipe.hi.cpu().numpy()
    T = pipe.T.cpu().numpy()
    gs = pipe.spins.cpu().numpy()
    oracle.lib()
    B = H.shape[0]
    segs, agree, done, pos = [], True, 0, 0
    for _ in range(nseg):
        t0, n = time.perf_counter(), 0
        while time.perf_counter() - t0 < seg_s:
            sl = slice(pos, pos + chunk)
            z, Ainv, _ = oracle.mmse(H[sl], y[sl], b['N0'], M)
            xg = oracle.mmse_hard(H[sl], y[sl], b['N0'], M, z, Ainv)
            oracle.ising_delta_at(H[sl], y[sl], xg, 1)
            s, de, st = oracle.arpeggio_fx(G[sl], h[sl], 2, args.copies, T, pipe.q, np.arange(pos, pos + chunk),
                                           pipe.key, 'rand', nthreads=cores)
            Lc = oracle.spins_to_levels_delta(s, xg, 1, M)
            oracle.select(H[sl], y[sl], Lc, xg)
            agree &= bool((gs[sl] == s).all())
            n += chunk
            pos = (pos + chunk) % B
        segs.append(dict(frames=n, seconds=time.perf_counter() - t0))
        done += n
    wall = sum(x['seconds'] for x in segs)
    rates = [x['frames'] / x['seconds'] for x in segs]
    return dict(value=done / wall, unit='detections/s', cores=cores, kind='port',
                nproc=os.cpu_count(), cpu_model=cpu_model(), segments=segs, segment_rates=rates,
                sample='the plain step on the host: chunks of %d of its frames (float64 numpy front end, C oracle '
                       'Arpeggio with %d threads, %d passes, R = %d, on the GPU front end\'s integers, numpy decode) in '
                       '%d segments of >= %.0f s; %d detections in %.1f s; anneal spins equal to the GPU\'s: %s' %
                       (chunk, cores, args.passes, args.copies, nseg, seg_s, done, wall, agree))


def profile_record(name, key):
    """A kernel's record from a committed profile summary under profiles/ (None if absent)."""
    p = os.path.join(ROOT, 'profiles', name)
    if not os.path.exists(p):
        return None
    d = json.load(open(p))
    return (d.get('launches') or d.get('kernels') or {}).get(key)


def roofline_front_end(t_ms, B, learned, form=2):
    """The front end: LMMSE + Gram on the matrix cores. Algorithmic work per frame (the Gram, H^T H, and H^T y):
    2 (2Nr)(2Nt)^2 + 2 (2Nr)(2Nt) FLOP, the real-lift count, for either form; the bound is the FP32 MFMA peak.
    Algorithmic HBM bytes: H (half of it for the complex-lift form 2, which reads [Re; Im] only), y, and the outputs.
    traffic: the measured HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by a device copy's reading);
    counters: the matrix-core and vector busy fractions of the same kernel (SQ counters), which also count the
    inversion's MFMAs."""
    D = Mr = 2 * NT
    flop = B * (2 * Mr * D * D + 2 * Mr * D)
    achieved = flop / (t_ms * 1e-3) / 1e12
    hcols = D // 2 if form == 2 else D
    hbm = B * (4 * Mr * hcols + 4 * Mr + 2 * D * D + 4 * D + D + 4 + (2 * feat_len() if learned else 0))
    key = ('adaptive' if learned else 'plain') + '/front_end'
    tr = profile_record('detect_step_traffic.json', key)
    ct = profile_record('detect_step_counters.json', key)
    return dict(bound='mfma', achieved=achieved, peak=MFMA_FP32_TFLOPS, unit='TFLOP/s',
                frac=achieved / MFMA_FP32_TFLOPS, traffic=tr['hbm_bytes_per_launch'] if tr else None,
                kernel='front_end_cpx_kernel<64>' if form == 2 else 'front_end_mfma_kernel<64>', launch_ms=t_ms, algorithmic_flop=flop, hbm_algorithmic_bytes=hbm,
                hbm=dict(achieved_GBs=hbm / (t_ms * 1e-3) / 1e9, peak_GBs=HBM_PEAK_GBS,
                         frac=hbm / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS),
                counters=dict(mfma_busy=ct['mfma_busy'], vector_busy=ct['vector_busy'],
                              mfma_per_frame=ct['mfma_per_wave'], source='profiles/detect_step_counters.json')
                if ct else None,
                traffic_source='profiles/detect_step_traffic.json' if tr else None)


def roofline_anneal(t_ms, B, R, L, flips, learned, kernel):
    """A first-accept anneal launch over B frames, R anneals of L passes each. Algorithmic HBM bytes per launch: G
    (int16, D x D, the front end's hand-off) + h + inst per frame and T per pass in [+ start thresholds P D int32 per
    frame]; spins (R P D int8), best energy (int64), (flips, tests) (2 int64) [+ kept flag] per anneal out. The bound
    that holds is instruction issue; the LDS bytes (one Gram column per flip, the D x D initial-field sweep) are
    reported beside it."""
    D, P = 2 * NT, 2
    hbm = B * (2 * D * D + 4 * D + 4) + 4 * L + B * R * (P * D + 8 + 16)
    if learned:
        hbm += B * 4 * P * D + B * R
    achieved = hbm / (t_ms * 1e-3) / 1e9
    lds = B * R * 4 * D * (flips + D)
    issue = None
    ip = os.path.join(ROOT, 'profiles', 'arpeggio_anneal_issue.json')
    if os.path.exists(ip):
        recs = [r for r in json.load(open(ip)).get('launches', []) if r.get('start') == ('learned' if learned else 'rand')]
        same = [r for r in recs if r.get('copies') == R and r.get('passes') == L]
        for rec in (same or recs)[:1]:
                q, qs = rec['valu_quad_cycles_per_anneal_pass'], rec['salu_quad_cycles_per_anneal_pass']
                avail = t_ms * 1e-3 * SIMDS * CLOCK_HZ
                issue = dict(bound='instruction issue (vector and scalar)', valu_quad_cycles_per_anneal_pass=q,
                             salu_quad_cycles_per_anneal_pass=qs, frac=4 * q * B * R * L / avail,
                             salu_frac=4 * qs * B * R * L / avail,
                             counters_at=dict(frames=rec['frames'], copies=rec['copies'], passes=rec['passes']),
                             source='profiles/arpeggio_anneal_issue.json')
    tr = profile_record('detect_step_traffic.json', 'adaptive/arpeggio_fx_stage1' if learned else 'plain/arpeggio_fx')
    return dict(bound='hbm', achieved=achieved, peak=HBM_PEAK_GBS, unit='GB/s', frac=achieved / HBM_PEAK_GBS,
                traffic=tr['hbm_bytes_per_launch'] if tr else None, kernel=kernel, algorithmic_bytes=hbm,
                launch_ms=t_ms, flips_per_anneal=flips,
                lds=dict(algorithmic_bytes=lds, achieved_GBs=lds / (t_ms * 1e-3) / 1e9, peak_GBs=LDS_PEAK_GBS,
                         frac=lds / (t_ms * 1e-3) / 1e9 / LDS_PEAK_GBS),
                issue=issue, traffic_source='profiles/detect_step_traffic.json' if tr else None)


def feat_len():
    return (4 * 2 * NT + 1 + 31) // 32 * 32


def mdi(args):
    """The MDI wrapper with the fixed-point first-accept solver on the reference's own batches (its draws; seeds, SNRs
    and picks of data/results/a513_fixed, copied into tests/golden/a513_fixed.npz): per-use bit errors paired with the
    reference's (bootstrap CI of the BER ratio), anneals per K against its counters, and detections/s (GPU time of the
    detect calls, host grouping included; the draws are not timed). On the same batches, the production Chord engine
    (lane-shared Chord, xorshift32 streams, prod_fx_xs_<Ns>) in the same wrapper: the paired BER ratio of first-accept
    to Chord and x_F, first-accept's modeled cycles per anneal (tests + flips) over Chord's (L + flips), beside the
    reference's (data/results/a513/analysis.json, in tests/golden/a513_cycles.json). Then the throughput of one large
    batch."""
    import torch
    from arpeggio_amd import cycles, frames
    from arpeggio_amd.mdi import MDIDetector, parse_pick
    from arpeggio_amd.stats import boot_ratio
    gd = os.path.join(ROOT, 'tests', 'golden')
    z = np.load(os.path.join(gd, 'a513_fixed.npz'), allow_pickle=False)
    idx = json.load(open(os.path.join(gd, 'a513_fixed.json')))
    gcy = json.load(open(os.path.join(gd, 'a513_cycles.json')))
    point = {'x16': 'P1', 'x32': 'P2', 'x64': 'P4'}

    def run(det, rec, nb):
        errs, wall = [], 0.0
        for bi in range(nb):
            b = frames.draw(rec['seed'], rec['snr'], bi, rec['m'], rec['nt'], 16, dev='cuda')
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, _, be = det.detect(b['H'], b['y'], b['N0'], frames.draw_seed(rec['seed'], rec['snr'], bi), ltrue=b['l'])
            torch.cuda.synchronize()
            wall += time.perf_counter() - t0
            errs.append(be.cpu().numpy().astype(np.int64))
        return np.concatenate(errs), wall

    out = {}
    for tag, nb in (('x32', args.mdi_batches), ('x16', max(1, args.mdi_batches // 5)),
                    ('x64', max(1, args.mdi_batches // 2))):
        rec = idx[tag]
        nb = min(nb, rec['batches'])
        det = MDIDetector(rec['m'], rec['nt'], 16, rec['pick'])
        ours, wall = run(det, rec, nb)
        theirs = z['%s/biterr' % tag][:nb].reshape(-1).astype(np.int64)
        r, lo, hi = boot_ratio(ours, theirs)
        nbits = ours.size * 4 * rec['nt']
        cref = z['%s/cnt' % tag][:nb].sum(0)
        out[tag] = dict(users=rec['nt'], snr_db=rec['snr'], pick=rec['pick'], seed=rec['seed'], batches=nb,
                        uses=int(ours.size), ber=float(ours.sum()) / nbits, reference_ber=float(theirs.sum()) / nbits,
                        ratio_to_reference=r, ci=[lo, hi], per_use_equal=float((ours == theirs).mean()),
                        anneals_per_K=[int(det.cnt.get(K, [0, 0, 0])[2]) for K in (1, 2, 3)],
                        reference_anneals_per_K=cref[:, 2].tolist(),
                        flips_per_K=[int(det.cnt.get(K, [0, 0, 0])[0]) for K in (1, 2, 3)],
                        reference_flips_per_K=cref[:, 0].tolist(), value=ours.size / wall, unit='detections/s',
                        batch=rec['m'])
        log('MDI %s: BER %.3e vs reference %.3e, ratio %.3f [%.3f, %.3f], %.0f detections/s' % (
            tag, out[tag]['ber'], out[tag]['reference_ber'], r, lo, hi, out[tag]['value']))
        # the production Chord engine on the same batches: the comparison within Chord's modeled budget
        pm = 'prod_fx_xs_%d' % parse_pick(rec['pick'])['Ns']
        dc = MDIDetector(rec['m'], rec['nt'], 16, pm)
        ch, wall_c = run(dc, rec, nb)
        rc, lc, hc = boot_ratio(ours, ch)
        x, xk, kf, kp = cycles.x_f(det.cnt, dc.cnt)
        g = gcy[point[tag]]
        out[tag]['chord'] = dict(
            method=pm, ber=float(ch.sum()) / nbits, value=ch.size / wall_c, unit='detections/s',
            ratio_first_accept_to_chord=rc, ci=[lc, hc], x_F=x, x_F_per_K=xk, cycles_per_anneal_first_accept=kf,
            cycles_per_anneal_chord=kp, within_budget=cycles.within_budget(x),
            reference=dict(ratio=g['fx_prod']['ratio'], ci=[g['fx_prod']['lo'], g['fx_prod']['hi']], x_F=g['x_F'],
                           cycles_per_anneal_first_accept=g['cost_per_anneal']['fx'],
                           cycles_per_anneal_chord=g['cost_per_anneal']['prod'], uses=g['fx_prod']['uses'],
                           chord_ber=g['fx_prod']['b_ber']))
        log('MDI %s within Chord\'s modeled budget: first-accept / Chord BER %.3f [%.3f, %.3f] (reference %.3f '
            '[%.3f, %.3f]), x_F %.3f (reference %.3f), Chord %.0f detections/s' % (
                tag, rc, lc, hc, g['fx_prod']['ratio'], g['fx_prod']['lo'], g['fx_prod']['hi'], x, g['x_F'],
                ch.size / wall_c))
    # one large batch (throughput; the per-batch keys of a 65,536-frame batch have no reference counterpart)
    rec = idx['x32']
    Bl = args.mdi_large
    b = frames.draw(rec['seed'], rec['snr'], 9000, Bl, 32, 16, dev='cuda')
    det = MDIDetector(Bl, 32, 16, rec['pick'])
    det.detect(b['H'], b['y'], b['N0'], 1, ltrue=b['l'])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, be = det.detect(b['H'], b['y'], b['N0'], 2, ltrue=b['l'])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out['x32_large'] = dict(frames=Bl, value=Bl / dt, unit='detections/s', ms_per_batch=dt * 1e3,
                            ber=float(be.sum()) / (Bl * 128))
    log('MDI x32, %d frames in one batch: %.0f detections/s' % (Bl, Bl / dt))
    return out


def batched(args):
    """Mixed-size batch: 2^17 frames, 40 / 40 / 20 % of 16 / 32 / 64 users, QPSK and 16-QAM half each, i.i.d.
    Rayleigh, per-frame SNR uniform in 8-24 dB (GPU generator), Arpeggio R = 8, 64 passes for every class. One step =
    per-class front end + one persistent anneal launch (autotuned shape) + per-class decode, replayed from one HIP
    graph; K timed replays after warm-up."""
    import torch
    from arpeggio_amd.autotune import tune_batched
    from arpeggio_amd.batch import BatchedDetector, draw_mixed, nbits, prepare
    fr = prepare(draw_mixed(args.seed, args.batched_frames))
    det = BatchedDetector([dict(nt=f['nt'], M=f['M'], B=f['l'].shape[0], idx=f['idx']) for f in fr], R=8, passes=64,
                          key=key_of(args.seed))
    det.run(fr)
    torch.cuda.synchronize()
    best_ms, shape, tried = tune_batched(det, reps=2, log=log)
    g = det.capture(fr)
    t_w = time.perf_counter()
    while time.perf_counter() - t_w < args.warm_seconds:
        g.replay()
        torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    B = args.batched_frames
    return dict(value=B * args.steps / dt, unit='detections/s', ms_per_step=dt / args.steps * 1e3,
                gpu_ms_per_step=e0.elapsed_time(e1) / args.steps, frames=B,
                classes=[dict(users=c['nt'], M=c['M'], frames=c['B'], tile=c['tile']) for c in det.cls],
                shape=dict(workgroups=shape[0], waves=shape[1], lds_budget=shape[2], anneals_per_wave=shape[3]),
                anneal_ms=best_ms,
                autotune=[dict(ms=ms, workgroups=sh[0], waves=sh[1], lds_budget=sh[2], anneals_per_wave=sh[3])
                          for ms, sh in tried],
                ber=det.bit_errors() / nbits(fr), R=8, passes=64, snr_db=[8.0, 24.0], graph=True)


SWEEP_32 = (12.0, 14.0, 16.0, 18.0, 20.0, 22.0)     # the 32-user BER curve's SNR points (dB)
SWEEP_64 = (14.0, 18.0, 22.0, 26.0)                 # the 64-user correlated channel's (dB)
CHORD_32 = dict(T0=0.3, T1=0.2, q=0.4, kappa=0.3)   # the reference's x32 Chord schedule (fxdetect.py SCHED['x32'])


def level_bit_errors(x, l):
    """Per-frame Gray bit errors of 16-QAM real levels x against l (2 bits per level: sign and |x| == 1)."""
    import torch
    gb = lambda v: torch.stack([v > 0, v.abs() == 1], -1)
    return (gb(x.float()) != gb(l.float())).flatten(1).sum(1).cpu().numpy().astype(np.int64)


def sweep32(args, model, c, r2):
    """BER against SNR at 32 users x 32 rx, 16-QAM, i.i.d. Rayleigh (the reference's draw recipe, draw index 2, so
    none of these frames is a calibration or measurement frame): the LMMSE decision, plain first accept (16 x 64),
    the learned adaptive detector at the headline's picks with its paired BER ratio to plain, and Chord at the
    reference's x32 schedule (256 steps, 16 anneals)."""
    from arpeggio_amd import frames
    from arpeggio_amd.detector import DetectPipeline
    from arpeggio_amd.stats import boot_ratio
    Bs = args.sweep_frames
    nb = Bs * 4 * NT
    plain = make_pipe(Bs, args, args.passes, 'rand')
    ap = make_adaptive(Bs, args, 'learned', model, c, r2)
    ch = DetectPipeline(Bs, 2 * NT, 2 * NT, M, args.copies, policy='chord', cycles=256, key=key_of(args.seed), inst0=0,
                        **CHORD_32)
    rows = []
    for snr in SWEEP_32:
        b = frames.draw(args.seed, snr, 2, Bs, NT, M, dev='cuda')
        ep = frame_errors(plain, b)
        el = level_bit_errors(plain.xg, b['l'])
        ea = frame_errors(ap, b)
        ec = frame_errors(ch, b)
        r, lo, hi = boot_ratio(ea, ep)
        rows.append(dict(snr_db=snr, lmmse=float(el.sum()) / nb, plain=float(ep.sum()) / nb,
                         adaptive=float(ea.sum()) / nb, adaptive_ratio_to_plain=r, ci=[lo, hi],
                         open_fraction=int(ap.count.item()) / Bs, chord=float(ec.sum()) / nb))
        log('sweep 32 users %.0f dB: LMMSE %.3e, plain %.3e, adaptive %.3e (ratio %.3f [%.3f, %.3f]), Chord %.3e' % (
            snr, rows[-1]['lmmse'], rows[-1]['plain'], rows[-1]['adaptive'], r, lo, hi, rows[-1]['chord']))
    return dict(frames_per_snr=Bs, users=NT, M=M, plain='%d x %d' % (args.copies, args.passes),
                adaptive='stage 1 %s, stage 2 %d x %d, c %.1f' % (args.stage1, r2, args.passes, c),
                chord='x32: %d x 256 steps, t 0.3 -> 0.2, q 0.4, kappa 0.3' % args.copies, points=rows)


SNR_64 = 28.0                                       # where the annealers beat LMMSE on this channel
SCHED_64 = dict(T0=0.08, T1=0.03, q=0.0)            # first accept, tuned at 26-28 dB on 4,096 frames
CHORD_64 = dict(T0=0.2, T1=0.2, q=0.4, kappa=0.3)   # Chord re-tuned on the same frames (256 steps)
S1_GRID_64 = ((4, 16), (8, 16), (8, 32), (8, 64), (16, 32))   # 64-user stage-1 budgets (anneals, passes)
CALIB_64 = 16384                                    # 64-user calibration frames (a separate draw)


def adaptive64(args, model, c, stage1, L2, B, nt=64):
    from arpeggio_amd.detector import AdaptivePipeline
    R1, L1 = stage1
    return AdaptivePipeline(B, 2 * nt, 2 * nt, 16, R1=R1, L1=L1, R2=32, L2=L2, c=c, start='learned',
                            key=key_of(args.seed), inst0=0, model=model, gate=False, **SCHED_64)


def pick_adaptive64(args, model, L2, nt=64):
    """The 64-user adaptive detector's stop constant and stage-1 budget on calibration frames (draw seed + 65, 28 dB):
    c by the reference's rule against the same detector without the stop (a56an.py:109-118), at stage 1 8 x 32; then
    the fastest stage-1 budget of S1_GRID_64 whose paired BER ratio to plain at 32 x 128 has an upper 95 % bound
    <= 1.10 (README.md:171)."""
    import torch
    from arpeggio_amd import frames
    from arpeggio_amd.detector import DetectPipeline
    from arpeggio_amd.stats import boot_ratio
    Bc = CALIB_64
    bc = frames.draw_device(args.seed + 65, SNR_64, SNR_64, Bc, nt, 16, rho_rx=0.5, rho_tx=0.5)
    pl = DetectPipeline(Bc, 2 * nt, 2 * nt, 16, 32, key=key_of(args.seed), inst0=0, policy='arpeggio', cycles=128,
                        start='rand', **SCHED_64)
    ref = frame_errors(pl, bc)
    del pl
    pipe = adaptive64(args, model, 0.0, (8, 32), L2, Bc, nt)
    pipe.thr_scale = -1.0                       # no stop: every frame is open
    base = frame_errors(pipe, bc)
    c, crows = C_GRID[0], []
    for cc in C_GRID:
        pipe.c, pipe.thr_scale = cc, nt + cc * nt ** 0.5
        r, lo, hi = boot_ratio(frame_errors(pipe, bc), base)
        crows.append(dict(c=cc, ratio=r, ci=[lo, hi]))
        if hi is not None and hi <= EQUIV_LIMIT:
            c = cc
    del pipe
    best, srows = None, []
    for s1 in S1_GRID_64:
        pipe = adaptive64(args, model, c, s1, L2, Bc, nt)
        r, lo, hi = boot_ratio(frame_errors(pipe, bc), ref)
        ms = time_pipe(pipe, bc, 5, 1, 0.3)['ms_per_step']
        ok = hi is not None and hi <= EQUIV_LIMIT
        srows.append(dict(anneals=s1[0], passes=s1[1], ratio=r, ci=[lo, hi], open=int(pipe.count.item()) / Bc, ms=ms,
                          equivalent=bool(ok)))
        log('64 users, stage 1 %d x %d: BER ratio to plain 128 %.3f [%.3f, %.3f], %.2f ms per %d frames' % (
            s1[0], s1[1], r, lo, hi, ms, Bc))
        del pipe
        torch.cuda.empty_cache()
        if ok and (best is None or ms < best[1]):
            best = (s1, ms)
    return c, (best[0] if best else (8, 32)), dict(c_grid=crows, stage1_grid=srows, frames=Bc)


def users64(args):
    """64 users x 64 rx, 16-QAM, Kronecker exponential correlation rho = 0.5 at both ends (GPU generator), 28 dB (where
    the annealers beat the LMMSE decision), 2^16 frames, 32 anneals per frame, first accept with T 0.08 -> 0.03: plain
    (random start) at 64 and 128 passes; the learned start (the 64-user network) at 32 and 64 passes, each with its
    paired BER ratio to plain at 128 passes, and the speedup of the smallest learned budget whose upper 95 % bound is
    <= 1.10; the adaptive two-stage detector (stage 2 = that learned budget; c and stage 1 picked on calibration
    frames by pick_adaptive64) with its ratio and speedup; Chord re-tuned (256 steps, t 0.2, q 0.4, kappa 0.3) on 4,096
    of the frames; the LMMSE decision's BER."""
    import torch
    from arpeggio_amd import frames
    from arpeggio_amd.detector import DetectPipeline, front_end
    from arpeggio_amd.learned import WEIGHTS_64, WarmStartMLP
    from arpeggio_amd.stats import boot_ratio
    B, nt = args.frames_64u, 64
    b = frames.draw_device(args.seed + 64, SNR_64, SNR_64, B, nt, 16, rho_rx=0.5, rho_tx=0.5)
    lt = b['l'].to(torch.int8).contiguous()
    nbits = B * 4 * nt
    o = front_end(b['H'], b['y'], b['N0'], 16, 1, want=('xg',))
    gb = lambda v: torch.stack([v > 0, v.abs() == 1], -1)
    lmmse = float((gb(o['xg'].float()) != gb(b['l'])).sum()) / nbits
    del o
    model = WarmStartMLP.load(WEIGHTS_64)
    out, errs = {}, {}
    for name, start, L in (('plain_64', 'rand', 64), ('plain_128', 'rand', 128), ('learned_32', 'learned', 32),
                           ('learned_64', 'learned', 64)):
        pipe = DetectPipeline(B, 2 * nt, 2 * nt, 16, 32, key=key_of(args.seed), inst0=0, policy='arpeggio', cycles=L,
                              start=start, model=model if start == 'learned' else None, **SCHED_64)
        _, be = pipe.run(b['H'], b['y'], b['N0'], lt)
        errs[name] = be.cpu().numpy().astype(np.int64)
        t = time_pipe(pipe, b, args.steps, args.warmup, args.warm_seconds)
        out[name] = dict(value=t['value'], ms_per_step=t['ms_per_step'], anneal_ms=t['anneal_ms'],
                         ber=float(errs[name].sum()) / nbits, passes=L, start=start)
        del pipe
    ref = errs['plain_128']
    chosen = None
    for name in ('learned_32', 'learned_64'):
        r, lo, hi = boot_ratio(errs[name], ref)
        out[name].update(ratio_to_plain_128=r, ci=[lo, hi], ber_matched=bool(hi <= EQUIV_LIMIT))
        if chosen is None and hi <= EQUIV_LIMIT:
            chosen = name
    log('64 users at %.0f dB: LMMSE %.3e, plain 64/128 passes %.3e / %.3e, learned 32/64 %.3e / %.3e' % (
        SNR_64, lmmse, out['plain_64']['ber'], out['plain_128']['ber'], out['learned_32']['ber'],
        out['learned_64']['ber']))
    # the two-stage detector: stage 2 is the one-stage learned budget that matched plain 128 (64 passes otherwise)
    L2 = out[chosen]['passes'] if chosen else 64
    c64, s1, picks = pick_adaptive64(args, model, L2, nt)
    pipe = adaptive64(args, model, c64, s1, L2, B, nt)
    _, be = pipe.run(b['H'], b['y'], b['N0'], lt)
    ea = be.cpu().numpy().astype(np.int64)
    r, lo, hi = boot_ratio(ea, ref)
    t = time_pipe(pipe, b, args.steps, args.warmup, args.warm_seconds)
    out['adaptive'] = dict(value=t['value'], ms_per_step=t['ms_per_step'], ber=float(ea.sum()) / nbits,
                           ratio_to_plain_128=r, ci=[lo, hi], ber_matched=bool(hi <= EQUIV_LIMIT),
                           open_fraction=int(pipe.count.item()) / B, stage1='%d,%d' % s1, stage2='32,%d' % L2, c=c64,
                           start='learned', stages=stage_ms(pipe, b, reps=3), picks=picks)
    log('64 users, adaptive %d x %d then 32 x %d (c = %.1f): BER %.3e, ratio to plain 128 %.3f [%.3f, %.3f], %.2f ms'
        % (s1[0], s1[1], L2, c64, out['adaptive']['ber'], r, lo, hi, t['ms_per_step']))
    del pipe
    # BER against SNR over 14-26 dB on separate draws: LMMSE, plain 32 x 128, the adaptive detector at the picks above
    Bs, nb = args.sweep_frames, args.sweep_frames * 4 * nt
    pl = DetectPipeline(Bs, 2 * nt, 2 * nt, 16, 32, key=key_of(args.seed), inst0=0, policy='arpeggio', cycles=128,
                        start='rand', **SCHED_64)
    ad = adaptive64(args, model, c64, s1, L2, Bs, nt)
    sweep = []
    for i, snr in enumerate(SWEEP_64):
        bs = frames.draw_device(args.seed + 640 + i, snr, snr, Bs, nt, 16, rho_rx=0.5, rho_tx=0.5)
        ep = frame_errors(pl, bs)
        el = level_bit_errors(pl.xg, bs['l'])
        ea = frame_errors(ad, bs)
        r, lo, hi = boot_ratio(ea, ep)
        sweep.append(dict(snr_db=snr, lmmse=float(el.sum()) / nb, plain_128=float(ep.sum()) / nb,
                          adaptive=float(ea.sum()) / nb, adaptive_ratio_to_plain_128=r, ci=[lo, hi],
                          open_fraction=int(ad.count.item()) / Bs))
        log('sweep 64 users %.0f dB: LMMSE %.3e, plain 128 %.3e, adaptive %.3e (ratio %.3f [%.3f, %.3f])' % (
            snr, sweep[-1]['lmmse'], sweep[-1]['plain_128'], sweep[-1]['adaptive'], r, lo, hi))
    out['sweep'] = dict(frames_per_snr=Bs, points=sweep)
    del pl, ad
    Bc = min(B, 4096)
    bc = {k: (v[:Bc] if torch.is_tensor(v) and v.dim() > 0 and v.shape[0] == B else v) for k, v in b.items()}
    pipe = DetectPipeline(Bc, 2 * nt, 2 * nt, 16, 32, key=key_of(args.seed), inst0=0, policy='chord', cycles=256,
                          **CHORD_64)
    t = time_pipe(pipe, bc, 3, 1, 0.0)
    out['chord_256'] = dict(value=t['value'], ms_per_step=t['ms_per_step'], anneal_ms=t['anneal_ms'], ber=t['ber'],
                            frames=Bc, **CHORD_64)
    del pipe
    res = dict(unit='detections/s', frames=B, users=nt, rx=nt, M=16, copies=32, snr_db=SNR_64, rho=0.5,
               schedule=SCHED_64, lmmse_ber=lmmse, data='GPU generator, Kronecker-correlated Rayleigh', **out)
    if chosen:
        res['learned_fixed_ber'] = dict(budget=chosen, value=out[chosen]['value'],
                                        speedup_vs_plain_128=out[chosen]['value'] / out['plain_128']['value'])
    if out['adaptive']['ber_matched']:
        res['adaptive_fixed_ber'] = dict(value=out['adaptive']['value'],
                                         speedup_vs_plain_128=out['adaptive']['value'] / out['plain_128']['value'])
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py measures one GPU: every box here has one MI355X; run with --gpus 1', file=sys.stderr)
        sys.exit(2)
    import torch
    if args.copy_probe:
        copy_probe()
        return
    from arpeggio_amd import _lib, frames
    # every frame here is the real lift of a complex channel (frames.draw, as mimo.complex_to_real builds it), so the
    # front end may take its complex-lift form: Hermitian Gram and complex sweep from half of H
    _lib.check(_lib.load().arp_set_front_end_form(args.front_end_form), 'front end form')
    from arpeggio_amd.stats import boot_ratio
    B, R, L = args.frames, args.copies, args.passes
    if args.snr is None:
        snr, calib = calibrate(args)
    else:
        snr, calib = args.snr, None
    log('measuring at %.2f dB; drawing %d frames (CPU generator)' % (snr, B))
    b = frames.draw(args.seed, snr, 0, B, NT, M, dev='cuda')
    plain = make_pipe(B, args, L, 'rand')
    be_plain = frame_errors(plain, b)
    nbits = B * 4 * NT
    out = dict(metric=METRIC, unit='detections/s', n_gpus=1, steps=args.steps, warmup=args.warmup,
               higher_is_better=True, scaling='weak', vs_baseline=None, dtype='int32', data='synthetic')
    model = None
    log('timing plain (%d anneals x %d passes)' % (R, L))
    tp = time_pipe(plain, b, args.steps, args.warmup, args.warm_seconds)
    plain_rec = dict(value=tp['value'], ms_per_step=tp['ms_per_step'], gpu_ms_per_step=tp['gpu_ms_per_step'],
                     passes=L, copies=R, start='rand', ber=float(be_plain.sum()) / nbits, anneal_ms=tp['anneal_ms'],
                     stages=stage_ms(plain, b))
    matched = False
    s1_given = args.stage1 is not None
    if not s1_given:
        args.stage1 = S1_DEFAULT   # for pick_c; pick_stage1 then replaces it
    if not args.plain_only:
        from arpeggio_amd.learned import DEFAULT_WEIGHTS, WarmStartMLP
        model = WarmStartMLP.load(args.weights or DEFAULT_WEIGHTS)
        if args.c is None:
            c, crows = pick_c(args, snr, model)
            if c is None:
                c = C_GRID[0]
        else:
            c, crows = args.c, None
        log('stop constant c = %.1f' % c)
        s1rows = None
        if not s1_given:
            _, s1rows = pick_stage1(args, snr, model, c)
        log('stage 1: %s (anneals,passes)' % args.stage1)
        if args.stage2_copies is None:
            r2, srows = pick_stage2(args, snr, model, c)
        else:
            r2, srows = args.stage2_copies, None
        log('stage 2: %d anneals x %d passes' % (r2, L))
        recs = {}
        for name, start, gate in (('learned', 'learned', args.gate), ('learned_other_gate', 'learned', not args.gate),
                                  ('rand', 'rand', False)):
            ap = make_adaptive(B, args, start, model, c, r2, gate)
            be = frame_errors(ap, b)
            r, lo, hi = boot_ratio(be, be_plain)
            open_frac = int(ap.count.item()) / B
            log('adaptive, %s: BER %.3e, ratio to plain %.3f [%.3f, %.3f], open %.3f' % (
                name, be.sum() / nbits, r, lo, hi, open_frac))
            t = time_pipe(ap, b, args.steps, args.warmup, args.warm_seconds)
            recs[name] = dict(value=t['value'], ms_per_step=t['ms_per_step'], gpu_ms_per_step=t['gpu_ms_per_step'],
                               ber=t['ber'], ratio_to_plain=r, ci=[lo, hi], ber_matched=bool(hi <= EQUIV_LIMIT),
                               open_fraction=open_frac, stage1=args.stage1, stage2='%d,%d' % (r2, L), c=c,
                               speedup_vs_plain=t['value'] / tp['value'], stages=stage_ms(ap, b),
                               gate=bool(gate) if start == 'learned' else None)
            if name == 'learned':
                lp = ap
                tl = t
            else:
                del ap
        # the learned start at the plain budget without the stop (what the start alone does)
        ns = make_pipe(B, args, L, 'learned', model)
        be = frame_errors(ns, b)
        r, lo, hi = boot_ratio(be, be_plain)
        recs['learned_no_stop'] = dict(ber=float(be.sum()) / nbits, ratio_to_plain=r, ci=[lo, hi], passes=L,
                                       copies=R, kept_learned=float(ns.used.double().mean()))
        del ns
        learned = recs['learned']
        matched = learned['ber_matched']
        out['learned'] = learned
        out['adaptive_random_start'] = recs['rand']
        out['learned_gated' if not args.gate else 'learned_ungated'] = recs['learned_other_gate']
        out['learned_no_stop'] = recs['learned_no_stop']
        out['stop_constant'] = dict(c=c, grid=crows, rule='largest c with paired BER ratio to no stop <= 1.10 at the '
                                    'upper 95 % bound (arpeggio/sim/a56an.py:109-118)')
        out['stage1_budget'] = dict(stage1=args.stage1, grid=s1rows,
                                    rule='fastest stage-1 budget (anneals x passes) with paired BER ratio to plain <= '
                                    '1.10 at the upper 95 % bound on the calibration frames (README.md:171)')
        out['stage2_budget'] = dict(copies=r2, passes=L, grid=srows,
                                    rule='smallest stage-2 anneal count with paired BER ratio to plain <= 1.10 at the '
                                    'upper 95 % bound on the calibration frames (README.md:171)')
    if matched:
        out.update(value=tl['value'], ms_per_step=tl['ms_per_step'])
        out['speedup_vs_plain'] = tl['value'] / tp['value']
        st = out['learned']['stages']
        timed = 'learned (adaptive)'
        fe_roof = roofline_front_end(st['front_end'], B, True, args.front_end_form)
        out['learned']['front_end_roofline'] = fe_roof
        if st['front_end'] >= st['anneal_1']:
            out['roofline'] = fe_roof
        else:
            fl = float(lp.st1.double()[..., 0].mean())
            out['roofline'] = roofline_anneal(st['anneal_1'], B, lp.R1, lp.L1, fl, True,
                                              'arpeggio_fx_kernel<2,1,0,true,64> (stage 1, learned start)')
    else:
        # no matched learned detector (or --plain-only): the plain annealer is the measured step
        out.update(value=tp['value'], ms_per_step=tp['ms_per_step'])
        timed = 'plain'
        fl = float(plain.st.double()[..., 0].mean())
        out['roofline'] = roofline_anneal(plain_rec['stages']['anneal'], B, R, L, fl, False,
                                          'arpeggio_fx_kernel<2,1,0,true,64>')
    out['plain'] = plain_rec
    out['config'] = dict(workload='one-stage dDI detection, 32x32 16-QAM i.i.d. Rayleigh, %d frames/step; plain: R=%d '
                                  'anneals x %d passes; learned: adaptive, stage 1 %s (anneals,passes), stage 2 %d '
                                  'anneals (picked by the equivalence rule) x %d passes on open frames; %s step timed' % (
                                      B, R, L, args.stage1, out.get('stage2_budget', {}).get('copies', R), L, timed),
                         frames=B, users=NT, rx=NT, M=M, copies=R, plain_passes=L, stage1=args.stage1,
                         policy='arpeggio', snr_db=snr, target_ber=args.target_ber, calibration=calib,
                         T0=0.1, T1=0.1, q=0.0,
                         front_end='complex lift (Hermitian Gram, half of H read)' if args.front_end_form == 2
                         else ('register sweep' if args.front_end_form == 1 else 'general matrix-core sweep'))
    if not args.no_cpu:
        log('cpu baseline (3 segments)')
        out['cpu_baseline'] = cpu_baseline(b, plain, args)
    if not args.no_sweep and model is not None:
        log('32-user BER sweep (%d frames per SNR)' % args.sweep_frames)
        out['sweep32'] = sweep32(args, model, c, r2)
    if not args.no_mdi:
        log('MDI wrapper against the reference fixtures')
        out['mdi'] = mdi(args)
    if not args.no_64u:
        log('64 users, correlated channel (%d frames)' % args.frames_64u)
        out['users64'] = users64(args)
    if not args.no_batched:
        log('mixed-size batch (%d frames)' % args.batched_frames)
        out['batched'] = batched(args)
    if args.with_chord:
        from arpeggio_amd.detector import DetectPipeline
        lt = b['l'].to(torch.int8).contiguous()
        cp = DetectPipeline(B, 2 * NT, 2 * NT, M, R, policy='chord', cycles=256, T0=0.3, T1=0.2, q=0.4, kappa=0.3,
                            key=key_of(args.seed), inst0=0)
        cp.run(b['H'], b['y'], b['N0'], lt)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ce = torch.zeros((), dtype=torch.int64, device='cuda')
        for _ in range(3):
            cp.run(b['H'], b['y'], b['N0'], lt)
            ce += cp.biterr.sum()
        torch.cuda.synchronize()
        out['chord'] = dict(value=3 * B / (time.perf_counter() - t1), S=256, ber=float(ce.item()) / (3 * nbits))
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
