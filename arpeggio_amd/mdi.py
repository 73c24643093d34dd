"""Multi-stage delta-Ising (MDI) detection on the GPU: the caller behind the reference's published Arpeggio BERs.

Per batch: the MMSE guess (front-end kernel) and the ordered MMSE-SIC guess (arp_mmse_sic); then up to S stages, each
of which computes every guess's search radius R_g (arp_mdi_radius), groups the frames by (guess g, K <= R_g, total
radius t = sum_g R_g) exactly as the reference's loops do, and for each group builds the fixed-point dDI sub-problems
around the guess (arp_mdi_build), anneals them with ceil(Ns / t) first-accept anneals (arp_anneal_arpeggio_list over
the group's frame list, 2K copies, per-call Philox key) and keeps a candidate only where it is strictly closer than the
guess's best so far (arp_decode_select_list with the guess's levels and distance as the prior). After the stages, the
better guess (first on ties) is the decision. All arrays stay on the device; the host reads the radii once per stage.
With the production Chord engine as the solver (pick prod_fx_xs_<Ns>), the same sub-problems go to
arp_anneal_chord_lane (Sk 2K steps, xorshift32 lane streams, key without the tag), as the reference's x3_prod runs.

Reference: arpeggio/mimo/mimo.py:546-602 (mdi), arpeggio/sim/local_code_a513/run_a513.py:43-130 (mdi_logged and its
fixed-point Arpeggio solver: key (seed_base + ci KEY_STEP + KEY_TAG) mod 2^64 for solver call ci, instance word = frame
index in the batch, schedule S sweeps x 2K passes; its 'prod' solver, run_a513.py:97-122, and x3_prod.py:90-112),
arpeggio/mimo/mdi_solvers.py:16-20 (Rmax, S, guesses, MDICHORD)."""
import math

import numpy as np
import torch

from . import _lib
from .arpeggiofx import schedule_check
from .chordfx import QSCALE, to_fixed
from .detector import MASK, _f32, _n0_arg, arpeggio_schedule

M64 = 0xFFFFFFFFFFFFFFFF
_DUMMY = dict(J=torch.zeros(1, 2, 2), h=torch.zeros(1, 2), h0=torch.zeros(1, 2))
KEY_MUL = 0x9E3779B97F4A7C15     # seed base = draw seed x this (run_a513.py:99)
KEY_STEP = 0xD1B54A32D192ED03    # per solver call (run_a513.py:28)
KEY_TAG = 0xA513                 # first-accept stream tag (run_a513.py:29); the Chord engine's keys have none
# Singh et al. 2024 Table I as the reference configures it (mdi_solvers.py:16-20): (users, M) -> Rmax, stages, guesses
MDI_CFG = {(16, 16): dict(Rmax=3, S=8, guesses=('mmse', 'sic')), (32, 16): dict(Rmax=3, S=8, guesses=('mmse', 'sic')),
           (64, 16): dict(Rmax=3, S=8, guesses=('mmse', 'sic')), (32, 4): dict(Rmax=1, S=1, guesses=('mmse',))}
# the production Chord engine's schedule in MDI (mdi_solvers.py:18): Sk steps per copy, T from t0 to t1, pinning q
MDICHORD = dict(Sk=128, t0=0.3, t1=0.3, q=0.4, kappa=0.0)


def parse_pick(pick):
    """abl_far_<S>_<T0>_<T1>_<q>_<start>_<Ns> (the reference's method names) -> dict (run_a513.py:34-38);
    prod_fx_xs_<Ns> (the production Chord engine, x3_prod.py:129) -> dict(solver='prod', Ns)."""
    if pick.startswith('prod_fx_xs_'):
        return dict(solver='prod', Ns=int(pick[len('prod_fx_xs_'):]))
    a = pick.replace('arpeggiofx_', 'abl_far_').split('_')
    if len(a) != 8 or a[0] != 'abl' or a[1] != 'far':
        raise ValueError('pick must look like abl_far_<S>_<T0>_<T1>_<q>_<start>_<Ns>: %r' % pick)
    return dict(solver='arpeggio', S=int(a[2]), T0=float(a[3]), T1=float(a[4]), q=float(a[5]), start=a[6],
                Ns=int(a[7]))


class MDIDetector:
    """MDI for batches of B frames of nt users (M-QAM) with the fixed-point first-accept solver or the production
    Chord engine (lane-shared Chord with xorshift32 lane streams).

    pick: abl_far_* / arpeggiofx_* string or dict (S, T0, T1, q, start, Ns) for first-accept; prod_fx_xs_<Ns> or
    dict(solver='prod', Ns) for the production Chord (schedule MDICHORD, x3_prod.py:80-84). cfg: Rmax, S (stages),
    guesses (default: MDI_CFG). Counters per K accumulate in self.cnt over detect() calls: [flips, tests, anneals]
    for first-accept, [steps, flips, anneals] for Chord (the sum of the first two is the modeled cycles, see
    arpeggio_amd.cycles); self.ann_k (B, Rmax) holds the last batch's anneals per frame and K."""

    def __init__(self, B, nt, M, pick, cfg=None, Q=QSCALE):
        self.B, self.nt, self.M, self.D, self.Mr = B, nt, M, 2 * nt, 2 * nt
        self.pk = parse_pick(pick) if isinstance(pick, str) else dict(pick)
        self.pk.setdefault('solver', 'arpeggio')
        self.cfg = dict(cfg or MDI_CFG[(nt, M)])
        self.Q = float(Q)
        self.lib = _lib.load()
        dev, D = 'cuda', self.D
        self.T, self.C = {}, {}
        for K in range(1, self.cfg['Rmax'] + 1):
            if self.pk['solver'] == 'prod':
                mc = MDICHORD
                F = to_fixed(_DUMMY, mc['Sk'] * 2 * K, mc['t0'], mc['t1'], mc['q'], mc['kappa'], Q=Q)
                T, qi = F['T'].cpu().tolist(), F['q']
                self.C[K] = F['c'].to(device=dev, dtype=torch.int32).contiguous()
            else:
                T, qi = arpeggio_schedule(self.pk['S'], self.pk['T0'], self.pk['T1'], self.pk['q'], K, Q)
            schedule_check(T, qi, 2 * K, D, Q)
            self.T[K] = torch.tensor(T, dtype=torch.int32, device=dev)
            self.q = qi
        self.Gi = torch.empty(B, D, D, dtype=torch.int32, device=dev)
        self.hi = torch.empty(B, D, dtype=torch.int32, device=dev)
        self.inst = (torch.arange(B, dtype=torch.int64) & MASK).to(torch.int32).to(dev)
        self.R = torch.empty(B, dtype=torch.int32, device=dev)
        self.dist = torch.empty(B, dtype=torch.float32, device=dev)
        self.count = torch.zeros(1, dtype=torch.int32, device=dev)
        self.cnt = {}
        self.calls = 0

    def guesses(self, H, y, N0):
        """Starting guesses: dict name -> (B, D) int8 levels."""
        L, s, P_ = self.lib, _lib.stream_ptr(), _lib.ptr
        B, Mr, D, M = self.B, self.Mr, self.D, self.M
        n0t, n0s = _n0_arg(N0, B)
        out = {}
        for g in self.cfg['guesses']:
            x = torch.empty(B, D, dtype=torch.int8, device='cuda')
            if g == 'mmse':
                _lib.check(L.arp_front_end(P_(H), P_(y), P_(n0t), n0s, B, Mr, D, M, 1, self.Q, P_(x), None, None, None,
                                           None, None, None, None, None, None, None, None, None, 0, None, s),
                           'front_end')
            else:
                _lib.check(L.arp_mmse_sic(P_(H), P_(y), P_(n0t), n0s, B, Mr, D, M, 0, P_(x), s), 'mmse_sic')
            out[g] = x
        return out

    def radius(self, H, y, x):
        """(R (B,) int32 device, ||y - H x||^2 (B,) float32 device) of guess levels x."""
        R = torch.empty(self.B, dtype=torch.int32, device='cuda')
        d = torch.empty(self.B, dtype=torch.float32, device='cuda')
        _lib.check(self.lib.arp_mdi_radius(_lib.ptr(H), _lib.ptr(y), _lib.ptr(x), self.B, self.Mr, self.D, self.M,
                                           self.cfg['Rmax'], _lib.ptr(R), None, _lib.ptr(d), _lib.stream_ptr()),
                   'mdi_radius')
        return R, d

    def solve(self, H, y, xg, idx, K, na, key):
        """One solver call: sub-problems around xg for frames idx (device int32), na anneals each; returns device
        spins (n, na, 2K D) in list order."""
        L, s, P_ = self.lib, _lib.stream_ptr(), _lib.ptr
        n, D = idx.numel(), self.D
        _lib.check(L.arp_mdi_build(P_(H), P_(y), P_(xg), P_(idx), n, self.Mr, D, self.Q, P_(self.Gi), P_(self.hi), s),
                   'mdi_build')
        self.count.fill_(n)
        spins = torch.empty(n, na, 2 * K * D, dtype=torch.int8, device='cuda')
        de = torch.empty(n, na, dtype=torch.int64, device='cuda')
        st = torch.empty(n, na, 2, dtype=torch.int64, device='cuda')
        T = self.T[K]
        if self.pk['solver'] == 'prod':
            _lib.check(L.arp_anneal_chord_lane(P_(self.Gi), P_(self.hi), P_(T), P_(self.C[K]), T.numel(), self.q, 2 * K,
                                               na, D, n, P_(self.inst), key & MASK, (key >> 32) & MASK, 1, P_(idx),
                                               P_(self.count), P_(spins), P_(de), P_(st), s), 'Chord (MDI)')
        else:
            _lib.check(L.arp_anneal_arpeggio_list(P_(self.Gi), None, P_(self.hi), P_(T), T.numel(), self.q, 2 * K, na, D,
                                                  n, P_(self.inst), key & MASK, (key >> 32) & MASK,
                                                  0 if self.pk['start'] == 'rand' else 1, None, 0, 0, P_(idx),
                                                  P_(self.count), P_(spins), P_(de), P_(st), None, s), 'anneal (MDI)')
        return spins, st

    def detect(self, H, y, N0, draw_seed, ltrue=None, counters=True):
        """One batch: returns (levels (B, D) int8, anneals per frame (B,) float64 numpy, bit errors (B,) int32 or
        None). draw_seed: the batch's draw seed (seed_base = draw_seed x KEY_MUL mod 2^64, run_a513.py:99)."""
        H, y = _f32(H), _f32(y)
        assert H.shape == (self.B, self.Mr, self.D) and y.shape == (self.B, self.Mr)
        B, D, M = self.B, self.D, self.M
        cfg = self.cfg
        seed_base = (int(draw_seed) * KEY_MUL) & M64
        ci = 0
        prod = self.pk['solver'] == 'prod'
        cur = self.guesses(H, y, N0)
        ann = np.zeros(B)
        self.ann_k = np.zeros((B, cfg['Rmax']))
        for st in range(cfg['S']):
            Rd = {}
            for g in cfg['guesses']:
                Rd[g] = self.radius(H, y, cur[g])
            Rh = {g: Rd[g][0].cpu().numpy().astype(np.int64) for g in cfg['guesses']}
            tot = sum(Rh[g] for g in cfg['guesses'])
            if int(tot.max()) == 0:
                break
            nxt = {}
            for g in cfg['guesses']:
                bl = cur[g].clone()
                bd = Rd[g][1].clone()
                for K in range(1, cfg['Rmax'] + 1):
                    need = Rh[g] >= K
                    if not need.any():
                        continue
                    for t in np.unique(tot[need]).tolist():
                        idx_h = np.flatnonzero(need & (tot == t))
                        na = math.ceil(self.pk['Ns'] / t)
                        key = (seed_base + ci * KEY_STEP + (0 if prod else KEY_TAG)) & M64
                        ci += 1
                        idx = torch.as_tensor(idx_h.astype(np.int32)).cuda()
                        spins, stt = self.solve(H, y, cur[g], idx, K, na, key)
                        self.count.fill_(len(idx_h))
                        _lib.check(self.lib.arp_decode_select_list(
                            _lib.ptr(H), _lib.ptr(y), _lib.ptr(cur[g]), _lib.ptr(spins), len(idx_h), na, 2 * K, D,
                            self.Mr, M, 1, None, _lib.ptr(idx), _lib.ptr(self.count), _lib.ptr(bl), _lib.ptr(bd),
                            _lib.ptr(bl), None, _lib.ptr(bd), _lib.stream_ptr()), 'decode (MDI)')
                        ann[idx_h] += na
                        self.ann_k[idx_h, K - 1] += na
                        if counters:
                            v = self.cnt.setdefault(K, [0, 0, 0])
                            sts = stt.sum((0, 1)).cpu()
                            if prod:   # [steps, flips, anneals]: L steps per anneal
                                v[0] += self.T[K].numel() * na * len(idx_h)
                                v[1] += int(sts[0])
                            else:      # [flips, tests, anneals]
                                v[0] += int(sts[0])
                                v[1] += int(sts[1])
                            v[2] += na * len(idx_h)
                nxt[g] = bl
            cur = nxt
        self.calls += ci
        # the better guess, first on ties
        names = list(cfg['guesses'])
        dists = torch.stack([self.radius(H, y, cur[g])[1] for g in names], 1)
        pick = dists.argmin(1)
        allc = torch.stack([cur[g] for g in names], 1)
        lhat = allc[torch.arange(B, device='cuda'), pick].contiguous()
        be = None
        if ltrue is not None:
            from .frames import levels_to_bits
            lt = ltrue.to(device='cuda')
            be = (levels_to_bits(lt.to(torch.int64), M) != levels_to_bits(lhat.to(torch.int64), M)).sum(-1).to(
                torch.int32)
        return lhat, ann, be
