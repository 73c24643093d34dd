#!/usr/bin/env python3
"""Detections/s at fixed BER on one MI355X, learned-start vs plain annealer: 32 users x 32 rx, 16-QAM, i.i.d.
Rayleigh, 16 anneals per frame, 65,536 frames per step, one-stage dDI (K = 1).

A step is one pass of the kernel path over the batch with H, y and the true levels already in HBM:
front end (LMMSE + dDI Ising build + fixed point [+ warm-start features]) -> [MLP -> start thresholds] ->
fixed-point Arpeggio anneal (first accept) -> decode/select -> bit errors.

Procedure (fixed-BER comparison):
  1. plain = random start, 64 passes (32 sweeps x 2 copies), T = 0.1, q = 0 (the reference's 32-user pick
     abl_far_32_0.1_0.1_0_rand_*); the SNR where plain reaches BER 1e-3 is found on calibration frames
     (log-BER interpolated between bracketing SNRs; --snr skips this);
  2. on the measurement frames at that SNR, the learned start runs at each budget of the ladder; the chosen
     budget is the smallest whose paired-bootstrap BER ratio to plain (95 % CI) reaches down to 1.0;
  3. both pipelines are timed over K steps; value = learned-start detections/s at the chosen budget,
     plain.value and speedup beside it.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--snr DB] [--frames B] [--no-cpu] [--with-chord]

Extra fields: roofline (the timed step's anneal kernel, HIP events on its own stream), cpu_baseline (the C oracle
of the plain chain on a bounded sample, on this host's cores)."""
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
# MI355X: 256 CUs x 4 SIMDs at 2.4 GHz (rocminfo); a wave64 vector instruction holds a SIMD for 4 clocks; LDS
# delivers 128 B per clock per CU
SIMDS, CLOCK_HZ = 256 * 4, 2.4e9
LDS_PEAK_GBS = 128 * 256 * CLOCK_HZ / 1e9
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
    ap.add_argument('--ladder', default='16,32,48,56,64,96,128,256',
                    help='learned-start pass budgets tried, ascending (multiples of 8)')
    ap.add_argument('--weights', default=None)
    ap.add_argument('--seed', type=int, default=1)
    ap.add_argument('--plain-only', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-frames', type=int, default=65536)
    ap.add_argument('--warm-seconds', type=float, default=1.0, help='minimum warm-up time before each timed loop')
    ap.add_argument('--with-chord', action='store_true')
    ap.add_argument('--no-batched', action='store_true', help='skip the mixed-size batch measurement')
    ap.add_argument('--batched-frames', type=int, default=1 << 17)
    ap.add_argument('--no-64u', action='store_true', help='skip the 64-user correlated-channel measurement')
    ap.add_argument('--frames-64u', type=int, default=1 << 16)
    ap.add_argument('--copy-probe', action='store_true', help='run only a plain device copy of known size (PMC scale)')
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
                          start=start, key=key_of(args.seed), inst0=0, model=model)


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


def cpu_baseline(b, pipe, B_cpu, args):
    """The oracle's plain chain on the first B_cpu frames on the host cores: numpy float64 front end, the C
    fixed-point Arpeggio (threads = min(16, cores)) on the GPU front end's integers, numpy decode. Wall time ->
    frames/s."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    cores = min(16, os.cpu_count() or 1)
    H = b['H'][:B_cpu].cpu().numpy()
    y = b['y'][:B_cpu].cpu().numpy()
    G = pipe.Gi[:B_cpu].cpu().numpy()
    h = pipe.hi[:B_cpu].cpu().numpy()
    T = pipe.T.cpu().numpy()
    inst = np.arange(B_cpu)
    oracle.lib()
    t0 = time.perf_counter()
    z, Ainv, _ = oracle.mmse(H, y, b['N0'], M)
    xg = oracle.mmse_hard(H, y, b['N0'], M, z, Ainv)
    oracle.ising_delta_at(H, y, xg, 1)
    s, de, st = oracle.arpeggio_fx(G, h, 2, args.copies, T, pipe.q, inst, pipe.key, 'rand', nthreads=cores)
    Lc = oracle.spins_to_levels_delta(s, xg, 1, M)
    oracle.select(H, y, Lc, xg)
    dt = time.perf_counter() - t0
    gs = pipe.spins[:B_cpu].cpu().numpy()
    agree = bool((gs == s).all())
    return dict(value=B_cpu / dt, unit='detections/s', cores=cores, kind='port',
                sample='first %d frames of the plain step: float64 numpy front end, C oracle Arpeggio (%d threads, '
                       '%d passes, R=%d) on the same integer inputs, numpy decode/select; %.1f s wall; anneal spins '
                       'equal the GPU: %s' % (B_cpu, cores, args.passes, args.copies, dt, agree))


def roofline(pipe, t, B, L, R, learned):
    """Dominant kernel of the timed step = the anneal. Algorithmic HBM bytes per launch: G (int32, D x D) + h + inst
    per frame and T per pass in [+ start thresholds P D int32 per frame]; spins (R P D int8), best energy (int64),
    (flips, tests) (2 int64) [+ used (int8)] per anneal out."""
    D, P = 2 * NT, 2
    hbm = B * (4 * D * D + 4 * D + 4) + 4 * L + B * R * (P * D + 8 + 16)
    if learned:
        hbm += B * 4 * P * D + B * R
    ms = t['anneal_ms']
    achieved = hbm / (ms * 1e-3) / 1e9
    st = pipe.st.double()
    flips = float(st[..., 0].mean())
    lds = B * R * 4 * D * (flips + D)
    traffic = None
    kern = 'arpeggio_fx_kernel<2,1,0,true,64>'
    tp = os.path.join(ROOT, 'profiles', 'arpeggio_anneal_traffic.json')
    if os.path.exists(tp):
        for rec in json.load(open(tp)).get('launches', []):
            if (rec.get('frames'), rec.get('passes'), rec.get('copies'), rec.get('learned')) == (B, L, R, learned):
                traffic = rec.get('hbm_bytes_per_launch')
    # the bound that holds: vector-instruction issue. Vector quad-cycles per anneal-pass come from the SQ counters
    # of the same launch shape (SQ_ACTIVE_INST_VALU / anneal-passes); the SIMD cycles are this launch's own time.
    issue = None
    ip = os.path.join(ROOT, 'profiles', 'arpeggio_anneal_issue.json')
    if os.path.exists(ip):
        for rec in json.load(open(ip)).get('launches', []):
            if rec.get('start') == ('learned' if learned else 'rand'):
                q = rec['valu_quad_cycles_per_anneal_pass']
                qs = rec['salu_quad_cycles_per_anneal_pass']
                busy = 4 * q * B * R * L
                avail = ms * 1e-3 * SIMDS * CLOCK_HZ
                issue = dict(bound='instruction issue (vector and scalar)', valu_quad_cycles_per_anneal_pass=q,
                             salu_quad_cycles_per_anneal_pass=qs,
                             achieved_simd_cycles=busy, available_simd_cycles=avail, frac=busy / avail,
                             salu_frac=4 * qs * B * R * L / avail,
                             counters_at=dict(frames=rec['frames'], copies=rec['copies'], passes=rec['passes']),
                             source='profiles/arpeggio_anneal_issue.json')
    lds_gbs = lds / (ms * 1e-3) / 1e9
    return dict(bound='hbm', achieved=achieved, peak=HBM_PEAK_GBS, unit='GB/s', frac=achieved / HBM_PEAK_GBS,
                traffic=traffic, kernel=kern + (' (learned start)' if learned else ''), algorithmic_bytes=hbm,
                launch_ms=ms, flips_per_anneal=flips,
                lds=dict(algorithmic_bytes=lds, achieved_GBs=lds_gbs, peak_GBs=LDS_PEAK_GBS,
                         frac=lds_gbs / LDS_PEAK_GBS),
                issue=issue,
                note='the anneal is an on-chip loop bound by instruction issue (Philox4x32-10, the 64-bit decision, '
                     'the ballot scan with its scalar find-first-set chain): issue.frac and issue.salu_frac are the '
                     'fractions of SIMD cycles spent issuing its vector and scalar instructions; HBM and LDS bytes '
                     'are the algorithmic ones')


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
                shape=dict(workgroups=shape[0], waves=shape[1], lds_budget=shape[2]), anneal_ms=best_ms,
                autotune=[dict(ms=ms, workgroups=sh[0], waves=sh[1], lds_budget=sh[2]) for ms, sh in tried],
                ber=det.bit_errors() / nbits(fr), R=8, passes=64, snr_db=[8.0, 24.0], graph=True)


def users64(args):
    """64 users x 64 rx, 16-QAM, Kronecker exponential correlation rho = 0.5 at both ends (GPU generator), 20 dB,
    2^16 frames, 32 anneals per frame: plain Arpeggio at 64 and 128 passes, K timed steps each after warm-up; and
    Chord at 256 steps (t 0.3 -> 0.2, q 0.4, kappa 0.3) on the first 4,096 of the frames, 3 timed steps."""
    import torch
    from arpeggio_amd import frames
    from arpeggio_amd.detector import DetectPipeline
    B, nt = args.frames_64u, 64
    b = frames.draw_device(args.seed + 64, 20.0, 20.0, B, nt, 16, rho_rx=0.5, rho_tx=0.5)
    out = {}
    for name, kw in (('arpeggio_64', dict(policy='arpeggio', cycles=64, T0=0.1, T1=0.1, q=0.0)),
                     ('arpeggio_128', dict(policy='arpeggio', cycles=128, T0=0.1, T1=0.1, q=0.0))):
        pipe = DetectPipeline(B, 2 * nt, 2 * nt, 16, 32, key=key_of(args.seed), inst0=0, **kw)
        t = time_pipe(pipe, b, args.steps, args.warmup, args.warm_seconds)
        out[name] = dict(value=t['value'], ms_per_step=t['ms_per_step'], anneal_ms=t['anneal_ms'], ber=t['ber'])
        del pipe
    Bc = min(B, 4096)
    bc = {k: (v[:Bc] if torch.is_tensor(v) and v.dim() > 0 and v.shape[0] == B else v) for k, v in b.items()}
    pipe = DetectPipeline(Bc, 2 * nt, 2 * nt, 16, 32, key=key_of(args.seed), inst0=0, policy='chord', cycles=256,
                          T0=0.3, T1=0.2, q=0.4, kappa=0.3)
    t = time_pipe(pipe, bc, 3, 1, 0.0)
    out['chord_256'] = dict(value=t['value'], ms_per_step=t['ms_per_step'], anneal_ms=t['anneal_ms'], ber=t['ber'],
                            frames=Bc)
    del pipe
    return dict(unit='detections/s', frames=B, users=nt, rx=nt, M=16, copies=32, snr_db=20.0, rho=0.5,
                data='GPU generator, Kronecker-correlated i.i.d. Rayleigh', **out)


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py measures one GPU: every box here has one MI355X; run with --gpus 1', file=sys.stderr)
        sys.exit(2)
    import torch
    if args.copy_probe:
        copy_probe()
        return
    from arpeggio_amd import frames
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
    ladder, chosen, model = [], None, None
    if not args.plain_only:
        from arpeggio_amd.learned import DEFAULT_WEIGHTS, WarmStartMLP
        model = WarmStartMLP.load(args.weights or DEFAULT_WEIGHTS)
        for Ll in [int(x) for x in args.ladder.split(',')]:
            pl = make_pipe(B, args, Ll, 'learned', model)
            be = frame_errors(pl, b)
            r, lo, hi = boot_ratio(be, be_plain)
            rec = dict(passes=Ll, ber=float(be.sum()) / nbits, ratio=r, ci=[lo, hi],
                       kept_learned=float(pl.used.double().mean()))
            log('learned start, %d passes: BER %.3e, ratio to plain %.3f [%.3f, %.3f]' % (Ll, rec['ber'], r, lo, hi))
            ladder.append(rec)
            del pl
            if lo <= 1.0:
                chosen = Ll
                break
    log('timing plain (%d passes)' % L)
    tp = time_pipe(plain, b, args.steps, args.warmup, args.warm_seconds)
    plain_rec = dict(value=tp['value'], ms_per_step=tp['ms_per_step'], gpu_ms_per_step=tp['gpu_ms_per_step'],
                     passes=L, start='rand', ber=float(be_plain.sum()) / nbits, anneal_ms=tp['anneal_ms'])
    if chosen is not None:
        log('timing learned start (%d passes)' % chosen)
        lp = make_pipe(B, args, chosen, 'learned', model)
        tl = time_pipe(lp, b, args.steps, args.warmup, args.warm_seconds)
        out.update(value=tl['value'], ms_per_step=tl['ms_per_step'])
        out['learned'] = dict(passes=chosen, ms_per_step=tl['ms_per_step'], gpu_ms_per_step=tl['gpu_ms_per_step'],
                              ber=tl['ber'], anneal_ms=tl['anneal_ms'], ber_matched=True)
        out['speedup_vs_plain'] = tl['value'] / tp['value']
        out['roofline'] = roofline(lp, tl, B, chosen, R, True)
        timed = 'learned'
    else:
        # no budget matched (or --plain-only): the plain annealer is the measured step
        out.update(value=tp['value'], ms_per_step=tp['ms_per_step'])
        if not args.plain_only:
            out['learned'] = dict(passes=None, ber_matched=False)
        out['roofline'] = roofline(plain, tp, B, L, R, False)
        timed = 'plain'
    out['plain'] = plain_rec
    out['ladder'] = ladder
    out['config'] = dict(workload='one-stage dDI detection, 32x32 16-QAM i.i.d. Rayleigh, %d frames/step, R=%d '
                                  'anneals/frame; %s step timed' % (B, R, timed),
                         frames=B, users=NT, rx=NT, M=M, copies=R, plain_passes=L, learned_passes=chosen,
                         policy='arpeggio', snr_db=snr, target_ber=args.target_ber, calibration=calib,
                         T0=0.1, T1=0.1, q=0.0)
    if not args.no_cpu:
        log('cpu baseline on %d frames' % args.cpu_frames)
        out['cpu_baseline'] = cpu_baseline(b, plain, min(args.cpu_frames, B), args)
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
