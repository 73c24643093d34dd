#!/usr/bin/env python3
"""Trains the learned warm start on the GPU and evaluates it against the plain random start.

    python -m arpeggio_amd.train_warmstart [--frames 1048576] [--out arpeggio_amd/weights/....npz]

Training frames: --users users x as many rx (32 by default), 16-QAM, Rayleigh (Kronecker exponential correlation --rho
at both ends), per-frame SNR uniform in [snr-lo, snr-hi] dB, drawn on the GPU. Labels: the detection of a 1024-pass,
8-anneal first-accept run with the schedule --T0 -> --T1 (learned.label_batch). Calibration: the logit scale of the
start thresholds (inv_temp) that gives the lowest BER at --cal-passes on held-out frames, stored with the weights.
Evaluation: other held-out frames at one SNR; bit error rates of the plain start and the learned start at several
pass budgets, R anneals per frame. Held-out frames of 32 users on an i.i.d. channel follow the reference's host draw
recipe (seed 7, draw index 3 for calibration, 0 for evaluation); other shapes use the GPU generator (seeds 7003 and
7000). Prints one JSON line.

    python -m arpeggio_amd.train_warmstart --users 64 --rho 0.5 --T0 0.08 --T1 0.03 --snr-lo 24 --snr-hi 32 \
        --cal-snr 28 --eval-snr 28 --R 32 --out arpeggio_amd/weights/warmstart_64x64_16qam_rho05.npz

    python -m arpeggio_amd.train_warmstart --calibrate-only [--weights IN] [--out OUT]   (calibrate existing weights)"""
import argparse
import json
import os
import sys
import time

import torch

from . import frames
from .detector import DetectPipeline
from .learned import DEFAULT_WEIGHTS, WarmStartMLP, label_batch, train


def held_out(a, snr, index, B):
    """Held-out frames: the reference's host recipe for 32 i.i.d. users, the GPU generator otherwise."""
    if a.users == 32 and a.rho == 0.0:
        return frames.draw(7, snr, index, B, 32, 16, dev='cuda')
    return frames.draw_device(7000 + index, snr, snr, B, a.users, 16, rho_rx=a.rho, rho_tx=a.rho)


def calibrate(model, a, log, grid=(0.25, 0.35, 0.5, 0.7, 1.0, 1.4, 2.0, 3.0)):
    """Sets model.inv_temp to the grid value with the lowest BER of the learned start at a.cal_passes passes on
    held-out frames (draw index 3, a.cal_snr dB); returns the table, with the random start beside it."""
    B = a.cal_frames
    b = held_out(a, a.cal_snr, 3, B)
    lt = b['l'].to(torch.int8)
    D = 2 * a.users
    res = {}
    for it in (None,) + tuple(grid):
        pipe = DetectPipeline(B, D, D, 16, a.R, policy='arpeggio', cycles=a.cal_passes, T0=a.T0, T1=a.T1, q=0.0,
                              start='rand' if it is None else 'learned', key=1234,
                              model=None if it is None else model, inv_temp=it)
        _, be = pipe.run(b['H'], b['y'], b['N0'], lt)
        ber = float(be.sum()) / (B * 2 * D)
        res['rand' if it is None else 'inv_temp_%g' % it] = ber
        log('calibration', 'rand' if it is None else 'inv_temp %g' % it, '%.3e' % ber)
    best = min(grid, key=lambda it: res['inv_temp_%g' % it])
    model.inv_temp = float(best)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--frames', type=int, default=1 << 20)
    ap.add_argument('--users', type=int, default=32)
    ap.add_argument('--rho', type=float, default=0.0, help='Kronecker exponential correlation at both ends')
    ap.add_argument('--T0', type=float, default=0.1, help='schedule of the labelling, calibration and evaluation runs')
    ap.add_argument('--T1', type=float, default=0.1)
    ap.add_argument('--eval-passes', default='8,16,32,64')
    ap.add_argument('--chunk', type=int, default=65536)
    ap.add_argument('--snr-lo', type=float, default=20.0)
    ap.add_argument('--snr-hi', type=float, default=28.0)
    ap.add_argument('--passes', type=int, default=1024)
    ap.add_argument('--label-R', type=int, default=8)
    ap.add_argument('--epochs', type=int, default=4)
    ap.add_argument('--batch', type=int, default=8192)
    ap.add_argument('--lr', type=float, default=2e-3)
    ap.add_argument('--seed', type=int, default=2024)
    ap.add_argument('--out', default=DEFAULT_WEIGHTS)
    ap.add_argument('--eval-snr', type=float, default=24.0)
    ap.add_argument('--eval-frames', type=int, default=65536)
    ap.add_argument('--R', type=int, default=16)
    ap.add_argument('--cal-snr', type=float, default=23.7)
    ap.add_argument('--cal-passes', type=int, default=64)
    ap.add_argument('--cal-frames', type=int, default=262144)
    ap.add_argument('--weights', default=None, help='weights to calibrate (default: --out)')
    ap.add_argument('--calibrate-only', action='store_true')
    a = ap.parse_args()
    log = lambda *x: print('#', *x, file=sys.stderr, flush=True)
    if a.calibrate_only:
        model = WarmStartMLP.load(a.weights or a.out)
        cal = calibrate(model, a, log)
        model.save(a.out)
        print(json.dumps(dict(calibration=cal, inv_temp=model.inv_temp, out=a.out)), flush=True)
        return
    t0 = time.time()
    feats, labels = [], []
    for c in range(0, a.frames, a.chunk):
        B = min(a.chunk, a.frames - c)
        b = frames.draw_device(a.seed * 7919 + c, a.snr_lo, a.snr_hi, B, a.users, 16, rho_rx=a.rho, rho_tx=a.rho)
        f, l, be = label_batch(b['H'], b['y'], b['N0'], b['l'].to(torch.int8), R=a.label_R, passes=a.passes,
                               T0=a.T0, T1=a.T1, key=a.seed * 0x9E3779B9 + c, inst0=c)
        feats.append(f)
        labels.append(l)
        if c == 0:
            log('labels: BER of the labelling run on chunk 0: %.3e' % (float(be.sum()) / (B * 4 * a.users)))
    feats = torch.cat(feats)
    labels = torch.cat(labels)
    torch.cuda.synchronize()
    t_label = time.time() - t0
    log('labelled %d frames in %.1f s' % (feats.shape[0], t_label))
    model = WarmStartMLP(2 * a.users, seed=a.seed)
    t1 = time.time()
    losses = train(model, feats, labels, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
    torch.cuda.synchronize()
    t_train = time.time() - t1
    cal = calibrate(model, a, log)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    model.save(a.out)
    log('trained in %.1f s, saved %s' % (t_train, a.out))
    # held-out evaluation on the reference's draw recipe
    b = held_out(a, a.eval_snr, 0, a.eval_frames)
    lt = b['l'].to(torch.int8)
    D = 2 * a.users
    nbits = a.eval_frames * 2 * D
    res = {}
    for L in (int(x) for x in a.eval_passes.split(',')):
        for start in ('rand', 'learned'):
            pipe = DetectPipeline(a.eval_frames, D, D, 16, a.R, policy='arpeggio', cycles=L, T0=a.T0, T1=a.T1, q=0.0,
                                  start=start, key=99, model=model if start == 'learned' else None)
            _, be = pipe.run(b['H'], b['y'], b['N0'], lt)
            torch.cuda.synchronize()
            r = dict(ber=float(be.sum()) / nbits, flips=float(pipe.st[..., 0].double().mean()))
            if start == 'learned':
                r['learned_kept'] = float(pipe.used.double().mean())
            res['%s_L%d' % (start, L)] = r
            log(start, L, r)
    print(json.dumps(dict(frames=a.frames, users=a.users, rho=a.rho, T0=a.T0, T1=a.T1, label_passes=a.passes,
                          label_R=a.label_R, epochs=a.epochs,
                          inv_temp=model.inv_temp, calibration=cal,
                          loss_first=losses[0], loss_last=losses[-1], label_s=t_label, train_s=t_train,
                          eval_snr=a.eval_snr, eval_frames=a.eval_frames, R=a.R, results=res)), flush=True)


if __name__ == '__main__':
    main()
