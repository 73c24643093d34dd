#!/usr/bin/env python3
"""Profiling driver for the detector's kernels, meant to run under rocprofv3.

    python3 -m arpeggio_amd.profile_anneal run [--start rand|learned] [--passes 64] [--frames 65536] [--copies 16]
                                               [--launches 3] [--snr 23.68] [--pipeline plain|adaptive]
                                               [--stage1 4,40] [--c 6] [--front-end-form 2] [--gate]
        draws the 32x32 16-QAM frames once, then runs the full step `launches` times; every kernel of the path appears
        `launches` times in a trace. plain: front end [+ MLP] + anneal + decode. adaptive: the two-stage detector
        (front end [+ MLP], stage-1 anneal and decode, certificate, stage-2 anneal and decode of the open frames).
        --front-end-form: 2 (default, as the bench) the complex-lift front end, 0 the general matrix-core form.
    python3 -m arpeggio_amd.profile_anneal chord [--frames 4096] [--launches 2]
        64 users x 64 rx, 16-QAM, Kronecker rho = 0.5, 28 dB: the bench's Chord detection (32 anneals of 256 steps,
        t 0.2 -> 0.2, q 0.4, kappa 0.3) `launches` times, for the snapshot kernel's counters.
    python3 -m arpeggio_amd.profile_anneal users64 [--frames 65536] [--launches 2] [--stage1 8,32]
        64 users x 64 rx, 16-QAM, Kronecker rho = 0.5, 28 dB (GPU-drawn frames): the bench's 64-user two-stage
        detector (learned start from the 64-user network, T 0.08 -> 0.03, stage 2 32 anneals x 64 passes, c = 6)
        `launches` times, for the counters of its front end, MLP, anneals and decodes.
    python3 -m arpeggio_amd.profile_anneal copy
        three device copies of 256 MiB: the known-size reference for scaling FETCH_SIZE / WRITE_SIZE.
    python3 -m arpeggio_amd.profile_anneal traffic FETCH_DB WRITE_DB COPY_FETCH_DB COPY_WRITE_DB [--match arpeggio]
        HBM bytes per dispatch of the matching kernel from two counter runs (FETCH_SIZE, WRITE_SIZE, KiB per
        dispatch), each scaled by what the same counter reads for the copy (bytes copied / counter bytes); prints
        one JSON object."""
import argparse
import json
import sqlite3
import sys

COPY_BYTES = 256 << 20


def run(a):
    import torch
    from . import _lib, frames
    _lib.check(_lib.load().arp_set_front_end_form(a.front_end_form), 'front end form')
    from .detector import DetectPipeline
    model = None
    if a.start == 'learned':
        from .learned import DEFAULT_WEIGHTS, WarmStartMLP
        model = WarmStartMLP.load(DEFAULT_WEIGHTS)
    b = frames.draw(1, a.snr, 0, a.frames, 32, 16, dev='cuda')
    lt = b['l'].to(torch.int8).contiguous()
    key = (0x9E3779B97F4A7C15 + 0xA513) & 0xFFFFFFFFFFFFFFFF
    if a.pipeline == 'adaptive':
        from .detector import AdaptivePipeline
        R1, L1 = (int(x) for x in a.stage1.split(','))
        pipe = AdaptivePipeline(a.frames, 64, 64, 16, R1=R1, L1=L1, R2=a.copies, L2=a.passes, c=a.c, start=a.start,
                                key=key, model=model, gate=a.gate)
    else:
        pipe = DetectPipeline(a.frames, 64, 64, 16, a.copies, cycles=a.passes, start=a.start, key=key, model=model,
                              gate=a.gate)
    for i in range(a.launches):
        pipe.run(b['H'], b['y'], b['N0'], lt)
        torch.cuda.synchronize()
        print('# step %d done' % i, file=sys.stderr, flush=True)
    print(json.dumps(dict(frames=a.frames, passes=a.passes, copies=a.copies, start=a.start,
                          bit_errors=int(pipe.biterr.sum()))))


def copy():
    import torch
    x = torch.empty(COPY_BYTES, dtype=torch.uint8, device='cuda')
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps(dict(copy_bytes=COPY_BYTES, launches=3)))


def counter_means(db, counter, match):
    c = sqlite3.connect(db)
    vals = {}
    for name, cn, v, disp in c.execute('select kernel_name, counter_name, value, dispatch_id from counters_collection'):
        if cn == counter and match in name:
            vals.setdefault(disp, 0.0)
            vals[disp] += v
    if not vals:
        raise SystemExit('%s: no %s for kernels matching %r' % (db, counter, match))
    v = sorted(vals.values())
    return sum(v) / len(v), len(v)


def traffic(a):
    f_k, nf = counter_means(a.dbs[0], 'FETCH_SIZE', a.match)
    w_k, nw = counter_means(a.dbs[1], 'WRITE_SIZE', a.match)
    cf, _ = counter_means(a.dbs[2], 'FETCH_SIZE', a.copy_match)
    cw, _ = counter_means(a.dbs[3], 'WRITE_SIZE', a.copy_match)
    sf, sw = COPY_BYTES / (cf * 1024.0), COPY_BYTES / (cw * 1024.0)
    out = dict(kernel_match=a.match, dispatches=[nf, nw], fetch_kib=f_k, write_kib=w_k,
               copy_fetch_kib=cf, copy_write_kib=cw, fetch_scale=sf, write_scale=sw,
               hbm_bytes_per_launch=f_k * 1024.0 * sf + w_k * 1024.0 * sw)
    print(json.dumps(out))


def chord(a):
    import torch
    from . import _lib, frames
    from .detector import DetectPipeline
    _lib.check(_lib.load().arp_set_front_end_form(2), 'front end form')
    b = frames.draw_device(7, 28.0, 28.0, a.frames, 64, 16, rho_rx=0.5, rho_tx=0.5)
    pipe = DetectPipeline(a.frames, 128, 128, 16, 32, key=1, policy='chord', cycles=256, T0=0.2, T1=0.2, q=0.4,
                          kappa=0.3)
    lt = b['l'].to(torch.int8).contiguous()
    for _ in range(a.launches):
        pipe.run(b['H'], b['y'], b['N0'], lt)
    torch.cuda.synchronize()
    print('chord launches', a.launches, 'bit errors', int(pipe.biterr.sum()))


def users64(a):
    import torch
    from . import _lib, frames
    from .detector import AdaptivePipeline
    from .learned import WEIGHTS_64, WarmStartMLP
    _lib.check(_lib.load().arp_set_front_end_form(2), 'front end form')
    b = frames.draw_device(65, 28.0, 28.0, a.frames, 64, 16, rho_rx=0.5, rho_tx=0.5)
    R1, L1 = (int(x) for x in a.stage1.split(','))
    pipe = AdaptivePipeline(a.frames, 128, 128, 16, R1=R1, L1=L1, R2=32, L2=64, c=6.0, start='learned', key=1,
                            model=WarmStartMLP.load(WEIGHTS_64), gate=False, T0=0.08, T1=0.03, q=0.0)
    lt = b['l'].to(torch.int8).contiguous()
    for i in range(a.launches):
        pipe.run(b['H'], b['y'], b['N0'], lt)
        torch.cuda.synchronize()
        print('# step %d done' % i, file=sys.stderr, flush=True)
    print(json.dumps(dict(frames=a.frames, stage1=a.stage1, open=int(pipe.count.item()),
                          bit_errors=int(pipe.biterr.sum()))))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest='cmd', required=True)
    r = sub.add_parser('run')
    r.add_argument('--start', default='rand', choices=('rand', 'learned'))
    r.add_argument('--passes', type=int, default=64)
    r.add_argument('--frames', type=int, default=65536)
    r.add_argument('--copies', type=int, default=16)
    r.add_argument('--launches', type=int, default=3)
    r.add_argument('--snr', type=float, default=23.68)
    r.add_argument('--pipeline', default='plain', choices=('plain', 'adaptive'))
    r.add_argument('--stage1', default='4,40', help='adaptive stage 1: anneals,passes (the bench\'s pick)')
    r.add_argument('--c', type=float, default=6.0, help='adaptive certificate constant')
    r.add_argument('--front-end-form', type=int, default=2, choices=(0, 1, 2))
    r.add_argument('--gate', action='store_true', help='learned start with the energy gate (the bench: without)')
    c = sub.add_parser('chord')
    c.add_argument('--frames', type=int, default=4096)
    c.add_argument('--launches', type=int, default=2)
    u = sub.add_parser('users64')
    u.add_argument('--frames', type=int, default=65536)
    u.add_argument('--launches', type=int, default=2)
    u.add_argument('--stage1', default='8,32', help='stage 1: anneals,passes (the bench\'s 64-user pick)')
    sub.add_parser('copy')
    t = sub.add_parser('traffic')
    t.add_argument('dbs', nargs=4)
    t.add_argument('--match', default='arpeggio_fx_kernel')
    t.add_argument('--copy-match', default='copy')
    a = ap.parse_args()
    if a.cmd == 'run':
        run(a)
    elif a.cmd == 'chord':
        chord(a)
    elif a.cmd == 'users64':
        users64(a)
    elif a.cmd == 'copy':
        copy()
    else:
        traffic(a)


if __name__ == '__main__':
    main()
