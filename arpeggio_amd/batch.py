"""Frames of uneven size detected in one step: 16/32/64 users, QPSK and 16-QAM, per-frame noise.

A mixed batch is kept in HBM as one packed group per class (users, constellation); `inst` of every frame is its
arrival index, so the Philox streams do not depend on the grouping. One step is

    per class: front end (LMMSE + dDI build + fixed point, per-frame N0)
    all classes: ONE persistent fixed-point Arpeggio launch (arp_anneal_arpeggio_batched)
    per class: decode/select + bit errors

launched back to back on one stream with no host synchronisation, so the whole step can be captured once in a HIP
graph (BatchedDetector.capture) and replayed. The persistent launch's shape (workgroups, waves per workgroup, frames
per tile) is chosen by arpeggio_amd.autotune.

The reference detects one fixed shape per call with a batch-scalar N0 (arpeggio/mimo/mimo.py:61-82); this module is
an extension beside the drop-in path, and every anneal in it equals arp_anneal_arpeggio's on the same class buffers."""
import ctypes
import math

import torch

from . import _lib
from .arpeggiofx import schedule_check
from .detector import MASK, QSCALE, _n0_arg, arpeggio_schedule
from .frames import PAM, complex_to_real

LDS_PAD = 64


def gram_stride(D):
    """int32 words one staged frame takes in the persistent launch's LDS."""
    return D * D + LDS_PAD * (1 if D <= 64 else 2)


def draw_mixed(seed, B, mix=((16, 0.4), (32, 0.4), (64, 0.2)), mods=(4, 16), snr_lo=8.0, snr_hi=24.0,
               dev='cuda'):
    """B frames on the GPU generator: users by `mix` fractions, constellation uniform over `mods`, i.i.d. Rayleigh,
    per-frame SNR uniform in [snr_lo, snr_hi] dB, in random arrival order. Returns a list of class dicts
    (nt, M, H, y, l, N0 (per frame), idx = arrival indices int64), largest frames first."""
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    counts = []
    for nt, f in mix:
        for m in mods:
            counts.append(((nt, m), int(round(B * f / len(mods)))))
    counts[-1] = (counts[-1][0], B - sum(c for _, c in counts[:-1]))
    labels = torch.cat([torch.full((c,), i, dtype=torch.int64) for i, (_, c) in enumerate(counts)])
    order = labels[torch.randperm(B, generator=torch.Generator().manual_seed(seed))]
    out = []
    for i, ((nt, M), c) in enumerate(counts):
        if c == 0:
            continue
        idx = torch.nonzero(order == i).flatten()
        a = PAM[M]['a']
        lv = torch.tensor(PAM[M]['levels'], dtype=torch.float32, device=dev)
        l = lv[torch.randint(0, len(lv), (c, 2 * nt), generator=g, device=dev)]
        Hc = torch.complex(torch.randn(c, nt, nt, generator=g, device=dev),
                           torch.randn(c, nt, nt, generator=g, device=dev)) / math.sqrt(2)
        H = (complex_to_real(Hc) * a).contiguous()
        snr = snr_lo + (snr_hi - snr_lo) * torch.rand(c, generator=g, device=dev)
        N0 = (nt / (10 ** (snr / 10))).contiguous()
        n = torch.randn(c, 2 * nt, generator=g, device=dev) * torch.sqrt(N0 / 2)[:, None]
        y = (torch.bmm(H, l[:, :, None])[:, :, 0] + n).contiguous()
        out.append(dict(nt=nt, M=M, H=H, y=y, l=l, N0=N0, idx=idx.to(dev)))
    out.sort(key=lambda d: -d['nt'])
    return out


class BatchedDetector:
    """Preallocated buffers for one mixed batch (the class list of draw_mixed: nt, M, B per class, largest first).
    shape = (workgroups, waves, lds_budget_bytes) of the persistent launch; None -> a default, or tune with
    arpeggio_amd.autotune.tune_batched. models: {nt: learned.WarmStartMLP} for the classes with a learned start (the
    front end also writes their features, the MLP their start thresholds, and the persistent launch starts those
    anneals from them, with the energy gate `gate`); classes of other sizes start at random."""

    def __init__(self, classes, R=8, passes=64, T0=0.1, T1=0.1, q=0.0, key=0, shape=None, Q=QSCALE, models=None,
                 gate=True):
        if not 1 <= len(classes) <= 8:
            raise ValueError('a mixed batch holds 1..8 classes')
        self.R, self.P, self.K, self.Q, self.key = R, 2, 1, float(Q), key
        self.lib = _lib.load()
        T, self.q = arpeggio_schedule(passes // 2, T0, T1, q)
        schedule_check(T, self.q, self.P, max(2 * c['nt'] for c in classes), Q)
        self.T = torch.tensor(T, dtype=torch.int32, device='cuda')
        self.L = len(T)
        self.cls = []
        for c in classes:
            nt, M, B = c['nt'], c['M'], int(c['B'] if 'B' in c else c['H'].shape[0])
            D = 2 * nt
            idx = c['idx'] if 'idx' in c else torch.arange(B, device='cuda')
            self.cls.append(dict(
                nt=nt, M=M, B=B, D=D, Mr=D,
                inst=(idx.to(torch.int64) & MASK).to(torch.int32).to('cuda').contiguous(),
                xg=torch.empty(B, D, dtype=torch.int8, device='cuda'),
                Gi=torch.empty(B, D, D, dtype=torch.int32, device='cuda'),
                hi=torch.empty(B, D, dtype=torch.int32, device='cuda'),
                scale=torch.empty(B, device='cuda'),
                spins=torch.empty(B, R, 2 * D, dtype=torch.int8, device='cuda'),
                de=torch.empty(B, R, dtype=torch.int64, device='cuda'),
                st=torch.empty(B, R, 2, dtype=torch.int64, device='cuda'),
                lhat=torch.empty(B, D, dtype=torch.int8, device='cuda'),
                biterr=torch.empty(B, dtype=torch.int32, device='cuda')))
            m = (models or {}).get(nt)
            if m is not None:
                from .learned import feat_ld
                if m.D != D or m.P != self.P:
                    raise ValueError('learned start for %d users needs a model with D = %d, P = 2' % (nt, D))
                self.cls[-1].update(model=m, feat=torch.empty(B, feat_ld(D), dtype=torch.bfloat16, device='cuda'),
                                    thr=torch.empty(B, self.P * D, dtype=torch.int32, device='cuda'),
                                    used=torch.empty(B, R, dtype=torch.int8, device='cuda'))
        self.gate = gate
        self.counter = torch.zeros(1, dtype=torch.int32, device='cuda')
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        self.set_shape(shape or (2 * cus, 8, 66048))

    def set_shape(self, shape):
        """(workgroups, waves, lds budget in bytes[, anneals per wave]): frames per tile = budget // staged frame size
        (at least 1); anneals per wave (1 or 2; 0 or absent: the library's choice) applies to the classes of D <= 32."""
        wg, waves, budget = shape[:3]
        apw = int(shape[3]) if len(shape) > 3 else 0
        self.shape = (int(wg), int(waves), int(budget), apw)
        arr = (_lib.BatchClass * len(self.cls))()
        p = lambda t: t.data_ptr()
        for i, c in enumerate(self.cls):
            tile = max(1, budget // (4 * gram_stride(c['D'])))
            c['tile'] = tile
            learned = 'thr' in c
            arr[i] = _lib.BatchClass(p(c['Gi']), p(c['hi']), p(c['inst']), p(self.T), p(c['spins']), p(c['de']),
                                     p(c['st']), c['D'], c['B'], self.L, self.q, 0, tile,
                                     p(c['thr']) if learned else None, p(c['used']) if learned else None,
                                     1 if learned and self.gate else 0, apw)
        self._arr = arr

    def anneal(self, stream=None):
        """The persistent launch alone (on the front end's integers already in the class buffers)."""
        wg, waves = self.shape[:2]
        _lib.check(self.lib.arp_anneal_arpeggio_batched(ctypes.cast(self._arr, ctypes.c_void_p), len(self.cls),
                                                        self.R, self.key & MASK,
                                                        (self.key >> 32) & MASK, _lib.ptr(self.counter), wg, waves,
                                                        _lib.stream_ptr(stream)), 'arpeggio_batched')

    def run(self, frames, with_errors=True, stream=None):
        """frames: the class list (same order and sizes as at construction) with H, y, N0 (and l for bit errors)."""
        L, s = self.lib, _lib.stream_ptr(stream)
        P_ = _lib.ptr
        for c, f in zip(self.cls, frames):
            n0t, n0s = _n0_arg(f['N0'], c['B'])
            feat = c.get('feat')
            _lib.check(L.arp_front_end(P_(f['H']), P_(f['y']), P_(n0t), n0s, c['B'], c['Mr'], c['D'], c['M'], 1,
                                       self.Q, P_(c['xg']), P_(c['Gi']), P_(c['hi']), P_(c['scale']), None, None,
                                       None, None, None, None, None, None, P_(feat),
                                       feat.shape[1] if feat is not None else 0, None, s), 'front_end')
            if feat is not None:
                c['model'].thresholds(feat, c['model'].inv_temp, out=c['thr'], stream=stream)
        self.anneal(stream)
        for c, f in zip(self.cls, frames):
            lt = f.get('lt') if with_errors else None
            _lib.check(L.arp_decode_select(P_(f['H']), P_(f['y']), P_(c['xg']), P_(c['spins']), c['B'], self.R,
                                           self.P, c['D'], c['Mr'], c['M'], 1, P_(lt), P_(c['lhat']),
                                           P_(c['biterr']) if lt is not None else None, None, s), 'decode_select')

    def _c_classes(self, frames, with_errors):
        """The arp_batch_frames table of the whole-chain C entry point (decisions into c['lhat_c'], errors into
        c['biterr_c'])."""
        arr = (_lib.BatchFrames * len(self.cls))()
        keep = []
        for i, (c, f) in enumerate(zip(self.cls, frames)):
            n0t, n0s = _n0_arg(f['N0'], c['B'])
            keep.append(n0t)
            if 'lhat_c' not in c:
                c['lhat_c'] = torch.empty_like(c['lhat'])
                c['biterr_c'] = torch.empty_like(c['biterr'])
            lt = f.get('lt') if with_errors else None
            p = lambda t: t.data_ptr() if t is not None else None
            arr[i] = _lib.BatchFrames(p(f['H']), p(f['y']), p(n0t), n0s if n0t is None else 0.0, p(c['inst']), p(lt),
                                      p(c['lhat_c']), p(c['biterr_c']) if lt is not None else None, c['nt'], c['M'],
                                      c['B'])
        return arr, keep

    def run_c(self, frames, with_errors=True, stream=None):
        """The same step through the single C entry point arp_detect_batched (its own workspace, this detector's launch
        shape); decisions and bit errors land in each class's 'lhat_c' and 'biterr_c'. The C chain starts every anneal
        at random: classes with a learned start go through run()."""
        if any('thr' in c for c in self.cls):
            raise ValueError('run_c: the whole-chain C entry point has no learned start; use run()')
        arr, keep = self._c_classes(frames, with_errors)
        need = self.lib.arp_detect_batched_workspace_bytes(ctypes.cast(arr, ctypes.c_void_p), len(self.cls), self.R)
        if getattr(self, '_work', None) is None or self._work.numel() < need:
            self._work = torch.empty(need, dtype=torch.uint8, device='cuda')
        wg, waves, budget = self.shape[:3]
        _lib.check(self.lib.arp_detect_batched(ctypes.cast(arr, ctypes.c_void_p), len(self.cls), _lib.ptr(self.T),
                                               self.L, self.q, self.R, self.Q, self.key & MASK, (self.key >> 32) & MASK,
                                               wg, waves, budget, _lib.ptr(self._work), self._work.numel(),
                                               _lib.stream_ptr(stream)), 'detect_batched')
        self._keep = (arr, keep)

    def tune_c(self, frames, reps=2):
        """The C autotuner (arp_tune_batched): returns ((workgroups, waves, lds_budget, anneals per wave), ms) and
        sets the shape."""
        arr, keep = self._c_classes(frames, False)
        need = self.lib.arp_detect_batched_workspace_bytes(ctypes.cast(arr, ctypes.c_void_p), len(self.cls), self.R)
        work = torch.empty(need, dtype=torch.uint8, device='cuda')
        shape = (ctypes.c_int * 4)()
        ms = ctypes.c_float()
        _lib.check(self.lib.arp_tune_batched(ctypes.cast(arr, ctypes.c_void_p), len(self.cls), _lib.ptr(self.T),
                                             self.L, self.q, self.R, self.Q, self.key & MASK, (self.key >> 32) & MASK,
                                             reps, _lib.ptr(work), work.numel(), shape, ctypes.byref(ms),
                                             _lib.stream_ptr()), 'tune_batched')
        self.set_shape(tuple(shape))
        return tuple(shape), ms.value

    def capture(self, frames, with_errors=True):
        """Captures run() once into a HIP graph (after one eager run that also settles launch attributes);
        returns the torch.cuda.CUDAGraph, replayed with .replay()."""
        self.run(frames, with_errors)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run(frames, with_errors)
        return g

    def bit_errors(self):
        return sum(int(c['biterr'].sum()) for c in self.cls)


def prepare(frames):
    """Adds the int8 true levels (key 'lt') the decode kernel compares against."""
    for f in frames:
        f['lt'] = f['l'].to(torch.int8).contiguous()
    return frames


def nbits(frames):
    return sum(f['l'].shape[0] * f['l'].shape[1] * PAM[f['M']]['bits'] for f in frames)

