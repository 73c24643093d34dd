#!/usr/bin/env python3
"""Detections/s of the one-stage delta-Ising detector on one MI355X: 32 users x 32 rx, 16-QAM, i.i.d. Rayleigh,
16 anneals per frame, 65,536 frames per step.

A step is one pass of the kernel path over the batch, with H, y and the true levels already in HBM:
front end (LMMSE + dDI Ising build + fixed point) -> fixed-point Arpeggio anneal (first-accept, 64 passes) ->
decode/select -> bit errors. `value` = frames / s over the K timed steps.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--snr DB] [--frames B] [--no-cpu] [--with-chord]

Extra fields: roofline (the anneal kernel, timed with HIP events on its own stream), cpu_baseline (the C oracle
of the same chain on a bounded sample, on this host's cores), ber over the timed steps."""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = 'one-stage dDI detection, 32x32 16-QAM i.i.d. Rayleigh, Arpeggio fixed point, R=16, L=64 passes'


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--frames', type=int, default=65536)
    ap.add_argument('--snr', type=float, default=24.0)
    ap.add_argument('--copies', type=int, default=16)
    ap.add_argument('--passes', type=int, default=64)
    ap.add_argument('--seed', type=int, default=1)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-frames', type=int, default=4096)
    ap.add_argument('--with-chord', action='store_true')
    ap.add_argument('--copy-probe', action='store_true', help='run only a plain device copy of known size (PMC scale)')
    return ap.parse_args()


def copy_probe():
    import torch
    n = 256 << 20
    a = torch.empty(n, dtype=torch.uint8, device='cuda')
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps(dict(copy_probe_bytes=n, launches=3)))


def cpu_baseline(b, pipe, B_cpu, args):
    """The oracle's chain on the first B_cpu frames on the host cores: numpy float64 front end, the C fixed-point
    Arpeggio (threads = min(16, cores)) on the GPU front end's integers, numpy decode. Wall time -> frames/s."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    import torch
    cores = min(16, os.cpu_count() or 1)
    H = b['H'][:B_cpu].cpu().numpy()
    y = b['y'][:B_cpu].cpu().numpy()
    G = pipe.Gi[:B_cpu].cpu().numpy()
    h = pipe.hi[:B_cpu].cpu().numpy()
    T = pipe.T.cpu().numpy()
    inst = np.arange(B_cpu)
    oracle.lib()
    t0 = time.perf_counter()
    z, Ainv, _ = oracle.mmse(H, y, b['N0'], 16)
    xg = oracle.mmse_hard(H, y, b['N0'], 16, z, Ainv)
    oracle.ising_delta_at(H, y, xg, 1)
    s, de, st = oracle.arpeggio_fx(G, h, 2, args.copies, T, pipe.q, inst, pipe.key, 'rand', nthreads=cores)
    Lc = oracle.spins_to_levels_delta(s, xg, 1, 16)
    oracle.select(H, y, Lc, xg)
    dt = time.perf_counter() - t0
    gs = pipe.spins[:B_cpu].cpu().numpy()
    agree = bool((gs == s).all())
    return dict(value=B_cpu / dt, unit='detections/s', cores=cores, kind='port',
                sample='first %d frames of the step: float64 numpy front end, C oracle Arpeggio (%d threads) on the '
                       'same integer inputs, numpy decode/select; %.1f s wall; anneal spins equal the GPU: %s'
                       % (B_cpu, cores, dt, agree))


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
    from arpeggio_amd.detector import DetectPipeline
    B, nt, M, R, L = args.frames, 32, 16, args.copies, args.passes
    print('# drawing %d frames (CPU generator)' % B, file=sys.stderr, flush=True)
    b = frames.draw(args.seed, args.snr, 0, B, nt, M, dev='cuda')
    ltrue = b['l'].to(torch.int8).contiguous()
    H, y = b['H'].contiguous(), b['y'].contiguous()
    key = (args.seed * 0x9E3779B97F4A7C15 + 0xA513) & 0xFFFFFFFFFFFFFFFF
    pipe = DetectPipeline(B, 2 * nt, 2 * nt, M, R, policy='arpeggio', cycles=L, T0=0.1, T1=0.1, q=0.0, start='rand',
                          key=key, inst0=0)
    stream = torch.cuda.current_stream()
    # the anneal kernel's own duration: events on the stream it is launched on
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    L_ = pipe.lib
    from arpeggio_amd import _lib

    def step(i=None):
        P_ = _lib.ptr
        s = _lib.stream_ptr(stream)
        _lib.check(L_.arp_front_end(P_(H), P_(y), None, float(b['N0']), B, 2 * nt, 2 * nt, M, 1, pipe.Q, P_(pipe.xg),
                                    P_(pipe.Gi), P_(pipe.hi), P_(pipe.scale), None, None, None, None, None, None, None, None,
                                    None, 0, s),
                   'front_end')
        if i is not None:
            ev[i][0].record(stream)
        _lib.check(L_.arp_anneal_arpeggio(P_(pipe.Gi), P_(pipe.hi), P_(pipe.T), pipe.T.numel(), pipe.q, 2, R, 2 * nt, B,
                                          P_(pipe.inst), key & 0xFFFFFFFF, key >> 32, 0, 0, P_(pipe.spins), P_(pipe.de),
                                          P_(pipe.st), s), 'arpeggio')
        if i is not None:
            ev[i][1].record(stream)
        _lib.check(L_.arp_decode_select(P_(H), P_(y), P_(pipe.xg), P_(pipe.spins), B, R, 2, 2 * nt, 2 * nt, M, 1,
                                        P_(ltrue), P_(pipe.lhat), P_(pipe.biterr), None, s), 'decode')

    print('# warmup %d' % args.warmup, file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    errs = torch.zeros((), dtype=torch.int64, device='cuda')
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
        errs += pipe.biterr.sum()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    value = B * args.steps / dt
    ann_ms = float(np.mean([a.elapsed_time(c) for a, c in ev]))
    nbits = B * 2 * nt * 2 * args.steps
    ber = float(errs.item()) / nbits
    st = pipe.st.double()
    flips, tests = float(st[..., 0].mean()), float(st[..., 1].mean())
    D = 2 * nt
    # algorithmic HBM bytes of one anneal launch: G (int32) + h + inst + T in; spins, de, (flips, tests) out
    hbm_bytes = B * (4 * D * D + 4 * D + 4) + 4 * L + B * R * (2 * D + 8 + 16)
    achieved = hbm_bytes / (ann_ms * 1e-3) / 1e9
    # algorithmic LDS bytes: one int32 Gram column per committed flip plus the D x D initial-field sweep per anneal
    lds_bytes = B * R * 4 * D * (flips + D)
    traffic = None
    tp = os.path.join(ROOT, 'profiles', 'arpeggio_anneal_traffic.json')
    if os.path.exists(tp):
        t = json.load(open(tp))
        if t.get('frames') == B and t.get('passes') == L and t.get('copies') == R:
            traffic = t.get('hbm_bytes_per_launch')
    out = dict(metric='detections/s at fixed BER (32u 16-QAM), learned-start vs plain annealer', value=value,
               unit='detections/s', n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=ms_step,
               higher_is_better=True, scaling='weak', vs_baseline=None, dtype='int32', data='synthetic',
               config=dict(workload=WORKLOAD, frames=B, users=nt, rx=nt, M=M, copies=R, passes=L, policy='arpeggio',
                           start='rand', snr_db=args.snr, T0=0.1, T1=0.1, q=0.0),
               ber=ber, anneal=dict(ms=ann_ms, flips_per_anneal=flips, tests_per_anneal=tests,
                                    share_of_step=ann_ms / ms_step),
               roofline=dict(bound='hbm', achieved=achieved, peak=HBM_PEAK_GBS, unit='GB/s', frac=achieved / HBM_PEAK_GBS,
                             traffic=traffic, kernel='arpeggio_fx_kernel<2,1>',
                             lds=dict(algorithmic_bytes=lds_bytes, achieved_GBs=lds_bytes / (ann_ms * 1e-3) / 1e9),
                             note='the anneal is bound by integer VALU issue (Philox4x32-10 and the 64-bit decision), '
                                  'not by HBM or the matrix cores'))
    if not args.no_cpu:
        print('# cpu baseline on %d frames' % args.cpu_frames, file=sys.stderr, flush=True)
        out['cpu_baseline'] = cpu_baseline(b, pipe, min(args.cpu_frames, B), args)
    if args.with_chord:
        cp = DetectPipeline(B, 2 * nt, 2 * nt, M, R, policy='chord', cycles=256, T0=0.3, T1=0.2, q=0.4, kappa=0.3,
                            key=key, inst0=0)
        cp.run(H, y, b['N0'], ltrue)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ce = torch.zeros((), dtype=torch.int64, device='cuda')
        for _ in range(3):
            cp.run(H, y, b['N0'], ltrue)
            ce += cp.biterr.sum()
        torch.cuda.synchronize()
        out['chord'] = dict(value=3 * B / (time.perf_counter() - t1), S=256, ber=float(ce.item()) / (3 * B * 128))
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
