"""The detector's batched kernel path on MI355X: front end -> anneal -> decode/select, all HIP kernels.

  front_end(H, y, N0, M, K)            LMMSE hard decision + dDI Ising build + fixed point (arp_front_end)
  decode_select(H, y, xg, spins, ...)  candidates -> lowest ||y - Hx||^2 -> bits (arp_decode_select)
  DetectPipeline                       preallocated buffers; run() launches the three kernels back to back on one
                                       stream with no host synchronisation (capturable in a HIP graph)
  detect(H, y, sigma2, policy, copies, cycles, seed)
                                       the one-call detector API (sigma2 = N0/2 per real dimension)

Reference call stack this replaces: arpeggio/mimo/fxdetect.py:24-43 (MMSE -> mapping -> to_fixed -> kernel ->
levels -> select) with the fixed-point Arpeggio solver of arpeggio/sim/local_code_a513/run_a513.py:107-128."""
import ctypes

import torch

from . import _lib
from .arpeggiofx import schedule_check
from .chordfx import QSCALE, geom, to_fixed

MASK = 0xFFFFFFFF


def _f32(x):
    return x.to(device='cuda', dtype=torch.float32).contiguous()


def _n0_arg(N0, B):
    if torch.is_tensor(N0) and N0.numel() > 1:
        t = N0.to(device='cuda', dtype=torch.float32).contiguous()
        assert t.numel() == B
        return t, 0.0
    return None, float(N0)


def front_end(H, y, N0, M, K=1, Q=QSCALE, want=('xg', 'Gi', 'hi', 'scale')):
    """H (B, Mr, D), y (B, Mr) device float tensors; N0 scalar or (B,). Returns a dict with the requested outputs:
    xg (B, D) int8, Gi (B, D, D) int32, hi (B, D) int32, scale (B,), Gf, hf (float G/scale, h/scale), J (B, N, N)
    int16 and hN (B, N) int32 (N = 2K D), z, mu (B, D), Gh (B, D, D) int16 (Gi at half the bytes)."""
    H, y = _f32(H), _f32(y)
    B, Mr, D = H.shape
    N = 2 * K * D
    dev = H.device
    shapes = dict(xg=((B, D), torch.int8), Gi=((B, D, D), torch.int32), hi=((B, D), torch.int32),
                  scale=((B,), torch.float32), Gf=((B, D, D), torch.float32), hf=((B, D), torch.float32),
                  J=((B, N, N), torch.int16), hN=((B, N), torch.int32), z=((B, D), torch.float32),
                  mu=((B, D), torch.float32), Jf=((B, N, N), torch.float32), hNf=((B, N), torch.float32),
                  Gh=((B, D, D), torch.int16))
    out = {k: torch.empty(*shapes[k][0], dtype=shapes[k][1], device=dev) for k in want}
    n0t, n0s = _n0_arg(N0, B)
    order = ('xg', 'Gi', 'hi', 'scale', 'Gf', 'hf', 'J', 'hN', 'z', 'mu', 'Jf', 'hNf')
    L = _lib.load()
    _lib.check(L.arp_front_end(_lib.ptr(H), _lib.ptr(y), _lib.ptr(n0t), n0s, B, Mr, D, int(M), int(K), float(Q),
                               *[_lib.ptr(out.get(k)) for k in order], None, 0, _lib.ptr(out.get('Gh')),
                               _lib.stream_ptr()), 'front_end')
    return out


def decode_select(H, y, xg, spins, M, P=2, delta=True, ltrue=None, want_dist=False):
    """Returns (lhat (B, D) int8, biterr (B,) int32 or None, dist (B, R+1) or None)."""
    H, y = _f32(H), _f32(y)
    B, Mr, D = H.shape
    R = spins.shape[1]
    lhat = torch.empty(B, D, dtype=torch.int8, device=H.device)
    be = torch.empty(B, dtype=torch.int32, device=H.device) if ltrue is not None else None
    dist = torch.empty(B, R + 1, dtype=torch.float32, device=H.device) if want_dist else None
    lt = ltrue.to(device='cuda', dtype=torch.int8).contiguous() if ltrue is not None else None
    L = _lib.load()
    xg, spins = xg.contiguous(), spins.contiguous()        # held until the launch is queued behind their copies
    _lib.check(L.arp_decode_select(_lib.ptr(H), _lib.ptr(y), _lib.ptr(xg), _lib.ptr(spins),
                                   B, R, P, D, Mr, int(M), 1 if delta else 0, _lib.ptr(lt), _lib.ptr(lhat),
                                   _lib.ptr(be), _lib.ptr(dist), _lib.stream_ptr()), 'decode_select')
    return lhat, be, dist


def arpeggio_schedule(S, T0, T1, q, K=1, Q=QSCALE):
    """One integer T per pass, S sweeps x 2K passes, and q in Q units (run_a513.py:83-87)."""
    Tsw = [int(round(v * Q)) for v in geom(T0, T1, S).double().tolist()]
    return [t for t in Tsw for _ in range(2 * K)], int(round(q * Q))


class DetectPipeline:
    """One-stage dDI (K = 1) detection of a fixed-shape batch with preallocated device buffers.

    policy 'arpeggio': first-accept anneal, cycles = L passes (S = L / 2 sweeps), schedule (T0, T1, q, start).
    policy 'chord': snapshot anneal on the expanded N = 2D problem, cycles = S steps, schedule (t0, t1, q, kappa)."""

    def __init__(self, B, Mr, D, M, R, policy='arpeggio', cycles=64, T0=0.1, T1=0.1, q=0.0, kappa=0.0,
                 start='rand', key=0, inst0=0, mode='arpeggio', Q=QSCALE, model=None, gate=True, inv_temp=None):
        """start 'learned' (Arpeggio only) needs model (a learned.WarmStartMLP): the front end also writes the
        features, the MLP turns them into per-spin start thresholds, and the anneal starts from them (gate: fall back
        to the random start per anneal when its energy is lower). inv_temp scales the logits (None: the model's
        calibrated value)."""
        self.B, self.Mr, self.D, self.M, self.R = B, Mr, D, M, R
        self.policy, self.K, self.P = policy, 1, 2
        self.key, self.inst0, self.start, self.mode = key, inst0, start, mode
        dev = 'cuda'
        self.lib = _lib.load()
        self.xg = torch.empty(B, D, dtype=torch.int8, device=dev)
        self.scale = torch.empty(B, dtype=torch.float32, device=dev)
        self.lhat = torch.empty(B, D, dtype=torch.int8, device=dev)
        self.biterr = torch.empty(B, dtype=torch.int32, device=dev)
        N = self.P * D
        if policy == 'arpeggio':
            assert cycles % self.P == 0, 'arpeggio cycles = passes, a multiple of P'
            T, qi = arpeggio_schedule(cycles // self.P, T0, T1, q, self.K, Q)
            schedule_check(T, qi, self.P, D, Q)
            self.T = torch.tensor(T, dtype=torch.int32, device=dev)
            self.q = qi
            # first accept reads the Gram as int16 (the front end's Gh); the parallel pass takes the int32 form
            self.Gi = torch.empty(B, D, D, dtype=torch.int16 if mode == 'arpeggio' else torch.int32, device=dev)
            self.hi = torch.empty(B, D, dtype=torch.int32, device=dev)
            self.inst = ((torch.arange(B, dtype=torch.int64) + inst0) & MASK).to(torch.int32).to(dev)
            self.spins = torch.empty(B, R, N, dtype=torch.int8, device=dev)
            self.de = torch.empty(B, R, dtype=torch.int64, device=dev)
            self.st = torch.empty(B, R, 2, dtype=torch.int64, device=dev)
            if start == 'learned':
                from .learned import feat_ld
                assert model is not None and model.D == D and model.P == self.P, 'learned start needs a matching model'
                self.model, self.gate = model, gate
                self.inv_temp = model.inv_temp if inv_temp is None else inv_temp
                self.feat = torch.empty(B, feat_ld(D), dtype=torch.bfloat16, device=dev)
                self.thr = torch.empty(B, N, dtype=torch.int32, device=dev)
                self.used = torch.empty(B, R, dtype=torch.int8, device=dev)
        elif policy == 'chord':
            assert N <= 256
            dummy = dict(J=torch.zeros(1, 2, 2), h=torch.zeros(1, 2), h0=torch.zeros(1, 2))
            F = to_fixed(dummy, cycles, T0, T1, q, kappa)
            self.T, self.c, self.q, self.S = F['T'].to(dev), F['c'].to(dev), F['q'], cycles
            self.J = torch.empty(B, N, N, dtype=torch.int16, device=dev)
            self.hN = torch.empty(B, N, dtype=torch.int32, device=dev)
            self.spins = torch.empty(B, R, N, dtype=torch.int8, device=dev)
            self.e = torch.empty(B, R, dtype=torch.int32, device=dev)
        else:
            raise ValueError('policy must be arpeggio or chord')
        self.Q = float(Q)

    def _check_inputs(self, H, y, N0, ltrue):
        """The kernels read the buffers by raw pointer: refuse anything but the pipeline's own shapes, float32 / int8,
        contiguous, on the GPU."""
        B, Mr, D = self.B, self.Mr, self.D
        for name, t, shape, dt in (('H', H, (B, Mr, D), torch.float32), ('y', y, (B, Mr), torch.float32),
                                   ('ltrue', ltrue, (B, D), torch.int8)):
            if t is None:
                continue
            if tuple(t.shape) != shape or t.dtype != dt or not t.is_cuda or not t.is_contiguous():
                raise ValueError('DetectPipeline.run: %s must be a contiguous %s GPU tensor of shape %s, got %s %s%s'
                                 % (name, dt, shape, t.dtype, tuple(t.shape), '' if t.is_cuda else ' on the CPU'))
        if torch.is_tensor(N0) and N0.numel() > 1:
            if N0.numel() != B or N0.dtype != torch.float32 or not N0.is_cuda or not N0.is_contiguous():
                raise ValueError('DetectPipeline.run: per-frame N0 must be a contiguous float32 GPU tensor of %d '
                                 'values' % B)

    STAGES = ('front_end', 'mlp', 'anneal', 'decode')

    def run(self, H, y, N0, ltrue=None, stream=None, events=None, marks=None):
        """Launches front end, anneal and decode on one stream (default: torch's current stream; the MLP of the learned
        start runs on the same stream); returns (lhat, biterr or None) device tensors.
        events: optional (start, end) torch.cuda.Event pair recorded on the same stream around the anneal launch.
        marks: optional list of len(STAGES) + 1 events recorded before the front end and after each stage."""
        self._check_inputs(H, y, N0, ltrue)
        L, s = self.lib, _lib.stream_ptr(stream)
        ts = stream if stream is not None else torch.cuda.current_stream()

        def mark(i):
            if events is not None:
                events[i].record(ts)

        def stage(i):
            if marks is not None:
                marks[i].record(ts)
        stage(0)
        B, Mr, D, M, R, P = self.B, self.Mr, self.D, self.M, self.R, self.P
        n0t, n0s = _n0_arg(N0, B)
        P_ = _lib.ptr
        g16 = self.policy == 'arpeggio' and self.mode == 'arpeggio'
        if self.policy == 'arpeggio' and self.start == 'learned':
            _lib.check(L.arp_front_end(P_(H), P_(y), P_(n0t), n0s, B, Mr, D, M, self.K, self.Q, P_(self.xg),
                                       None if g16 else P_(self.Gi), P_(self.hi), P_(self.scale), None, None, None,
                                       None, None, None, None, None, P_(self.feat), self.feat.shape[1],
                                       P_(self.Gi) if g16 else None, s), 'front_end')
            stage(1)
            self.model.thresholds(self.feat, self.inv_temp, out=self.thr, stream=stream)
            stage(2)
            mark(0)
            if g16:
                _lib.check(L.arp_anneal_arpeggio_list(None, P_(self.Gi), P_(self.hi), P_(self.T), self.T.numel(),
                                                      self.q, P, R, D, B, P_(self.inst), self.key & MASK,
                                                      (self.key >> 32) & MASK, 0, P_(self.thr), 1 if self.gate else 0,
                                                      0, None, None, P_(self.spins), P_(self.de), P_(self.st),
                                                      P_(self.used), s), 'arpeggio_list')
            else:
                _lib.check(L.arp_anneal_arpeggio_ws(P_(self.Gi), P_(self.hi), P_(self.T), self.T.numel(), self.q, P, R,
                                                    D, B, P_(self.inst), self.key & MASK, (self.key >> 32) & MASK,
                                                    P_(self.thr), 1 if self.gate else 0, P_(self.spins), P_(self.de),
                                                    P_(self.st), P_(self.used), s), 'arpeggio_ws')
            mark(1)
        elif self.policy == 'arpeggio':
            _lib.check(L.arp_front_end(P_(H), P_(y), P_(n0t), n0s, B, Mr, D, M, self.K, self.Q, P_(self.xg),
                                       None if g16 else P_(self.Gi), P_(self.hi), P_(self.scale), None, None, None,
                                       None, None, None, None, None, None, 0, P_(self.Gi) if g16 else None, s),
                       'front_end')
            stage(1)
            stage(2)
            mark(0)
            if g16:
                _lib.check(L.arp_anneal_arpeggio_list(None, P_(self.Gi), P_(self.hi), P_(self.T), self.T.numel(),
                                                      self.q, P, R, D, B, P_(self.inst), self.key & MASK,
                                                      (self.key >> 32) & MASK, 0 if self.start == 'rand' else 1, None,
                                                      0, 0, None, None, P_(self.spins), P_(self.de), P_(self.st), None,
                                                      s), 'arpeggio_list')
            else:
                _lib.check(L.arp_anneal_arpeggio(P_(self.Gi), P_(self.hi), P_(self.T), self.T.numel(), self.q, P, R, D,
                                                 B, P_(self.inst), self.key & MASK, (self.key >> 32) & MASK,
                                                 0 if self.start == 'rand' else 1, 1, P_(self.spins), P_(self.de),
                                                 P_(self.st), s), 'arpeggio')
            mark(1)
        else:
            _lib.check(L.arp_front_end(P_(H), P_(y), P_(n0t), n0s, B, Mr, D, M, self.K, self.Q, P_(self.xg), None,
                                       None, P_(self.scale), None, None, P_(self.J), P_(self.hN), None, None, None,
                                       None, None, 0, None, s), 'front_end')
            stage(1)
            stage(2)
            mark(0)
            _lib.check(L.arp_anneal_chord(P_(self.J), P_(self.hN), P_(self.hN), P_(self.T), P_(self.c), self.q, self.S,
                                          P, R, P * D, B, self.key & MASK, (self.key >> 32) & MASK, self.inst0, 0,
                                          P_(self.e), P_(self.spins), None, s), 'chord')
            mark(1)
        stage(3)
        _lib.check(L.arp_decode_select(P_(H), P_(y), P_(self.xg), P_(self.spins), B, R, P, D, Mr, M, 1, P_(ltrue),
                                       P_(self.lhat), P_(self.biterr) if ltrue is not None else None, None, s),
                   'decode_select')
        stage(4)
        return self.lhat, (self.biterr if ltrue is not None else None)


def detect(H, y, sigma2, policy='arpeggio', copies=16, cycles=64, seed=0, M=16, T0=0.1, T1=0.1, q=0.0, kappa=0.0,
           start='rand', inst0=0, ltrue=None):
    """One-call detector. H (B, 2Nr, 2Nt) real lift, y (B, 2Nr), sigma2 = N0/2 per real dimension (scalar or (B,)).
    Returns levels (B, 2Nt) int8 (and per-frame bit errors if ltrue is given)."""
    H, y = _f32(H), _f32(y)
    B, Mr, D = H.shape
    N0 = 2 * sigma2 if not torch.is_tensor(sigma2) else (2 * sigma2).to('cuda', torch.float32)
    pipe = DetectPipeline(B, Mr, D, M, copies, policy=policy, cycles=cycles, T0=T0, T1=T1, q=q, kappa=kappa,
                          start=start, key=seed, inst0=inst0)
    lt = ltrue.to(device='cuda', dtype=torch.int8).contiguous() if ltrue is not None else None
    lhat, be = pipe.run(H, y, N0, lt)
    return (lhat, be) if ltrue is not None else lhat


class AdaptivePipeline:
    """Two-stage detection with an early-stop certificate (no reference counterpart for the two-stage form).

    Stage 1 anneals every frame briefly (R1 anneals x L1 passes, learned or random start) and decodes it. A frame whose
    chosen candidate already explains y within the noise, ||y - H x||^2 <= N0 (Nr + c sqrt(Nr)), is final: the
    reference's first-pass stop rule (arpeggio/sim/a56an.py:58-67). Stage 2 anneals only the open frames, from a
    device-side list, with R2 more anneals (Philox streams R1 .. R1 + R2 - 1) of L2 passes, and its best candidate
    replaces stage 1's only where it is strictly closer. Every launch is stream-ordered with no host synchronisation,
    so run() can be captured in a HIP graph.

    start 'learned' needs model (a learned.WarmStartMLP for this D); the MLP runs once for both stages."""

    def __init__(self, B, Mr, D, M, R1=4, L1=16, R2=16, L2=64, c=3.0, start='learned', T0=0.1, T1=0.1, q=0.0,
                 key=0, inst0=0, Q=QSCALE, model=None, gate=True, inv_temp=None):
        self.B, self.Mr, self.D, self.M = B, Mr, D, M
        self.R1, self.L1, self.R2, self.L2, self.c = R1, L1, R2, L2, float(c)
        self.P, self.K, self.key, self.start, self.Q = 2, 1, key, start, float(Q)
        self.lib = _lib.load()
        dev = 'cuda'
        N = self.P * D
        sched = []
        for L in (L1, L2):
            assert L % self.P == 0, 'passes must be a multiple of P'
            T, qi = arpeggio_schedule(L // self.P, T0, T1, q, self.K, Q)
            schedule_check(T, qi, self.P, D, Q)
            sched.append(torch.tensor(T, dtype=torch.int32, device=dev))
        self.T1t, self.T2t, self.q = sched[0], sched[1], qi
        self.xg = torch.empty(B, D, dtype=torch.int8, device=dev)
        self.Gi = torch.empty(B, D, D, dtype=torch.int16, device=dev)   # the front end's int16 Gram (Gh)
        self.hi = torch.empty(B, D, dtype=torch.int32, device=dev)
        self.scale = torch.empty(B, dtype=torch.float32, device=dev)
        self.inst = ((torch.arange(B, dtype=torch.int64) + inst0) & MASK).to(torch.int32).to(dev)
        self.spins1 = torch.empty(B, R1, N, dtype=torch.int8, device=dev)
        self.de1 = torch.empty(B, R1, dtype=torch.int64, device=dev)
        self.st1 = torch.empty(B, R1, 2, dtype=torch.int64, device=dev)
        self.spins2 = torch.empty(B, R2, N, dtype=torch.int8, device=dev)
        self.de2 = torch.empty(B, R2, dtype=torch.int64, device=dev)
        self.st2 = torch.empty(B, R2, 2, dtype=torch.int64, device=dev)
        self.d1 = torch.empty(B, dtype=torch.float32, device=dev)
        self.d2 = torch.empty(B, dtype=torch.float32, device=dev)
        self.lhat = torch.empty(B, D, dtype=torch.int8, device=dev)
        self.biterr = torch.empty(B, dtype=torch.int32, device=dev)
        self.list = torch.empty(B, dtype=torch.int32, device=dev)
        self.count = torch.zeros(1, dtype=torch.int32, device=dev)
        self.thr = None
        if start == 'learned':
            from .learned import feat_ld
            assert model is not None and model.D == D and model.P == self.P, 'learned start needs a matching model'
            self.model, self.gate = model, gate
            self.inv_temp = model.inv_temp if inv_temp is None else inv_temp
            self.feat = torch.empty(B, feat_ld(D), dtype=torch.bfloat16, device=dev)
            self.thr = torch.empty(B, N, dtype=torch.int32, device=dev)
            self.used1 = torch.empty(B, R1, dtype=torch.int8, device=dev)
            self.used2 = torch.empty(B, R2, dtype=torch.int8, device=dev)
        elif start != 'rand':
            raise ValueError("start must be 'learned' or 'rand'")
        nr = Mr // 2
        self.thr_scale = nr + self.c * nr ** 0.5

    _check_inputs = DetectPipeline._check_inputs

    STAGES = ('front_end', 'mlp', 'anneal_1', 'decode_1', 'certify', 'anneal_2', 'decode_2')

    def run(self, H, y, N0, ltrue=None, stream=None, events=None, marks=None):
        """Both stages on one stream; returns (lhat, biterr or None). events: optional (start, end) pair recorded around
        the two anneal stages (certificate included). marks: optional list of len(STAGES) + 1 events recorded before the
        front end and after each stage (for per-stage times)."""
        self._check_inputs(H, y, N0, ltrue)
        L, s = self.lib, _lib.stream_ptr(stream)
        ts = stream if stream is not None else torch.cuda.current_stream()

        def mark(i):
            if marks is not None:
                marks[i].record(ts)
        mark(0)
        B, Mr, D, M, P = self.B, self.Mr, self.D, self.M, self.P
        n0t, n0s = _n0_arg(N0, B)
        P_ = _lib.ptr
        learned = self.thr is not None
        _lib.check(L.arp_front_end(P_(H), P_(y), P_(n0t), n0s, B, Mr, D, M, self.K, self.Q, P_(self.xg), None,
                                   P_(self.hi), P_(self.scale), None, None, None, None, None, None, None, None,
                                   P_(self.feat) if learned else None, self.feat.shape[1] if learned else 0,
                                   P_(self.Gi), s), 'front_end')
        mark(1)
        if learned:
            self.model.thresholds(self.feat, self.inv_temp, out=self.thr, stream=stream)
        mark(2)
        if events is not None:
            events[0].record(ts)
        k0, k1 = self.key & MASK, (self.key >> 32) & MASK
        gate = 1 if learned and self.gate else 0
        _lib.check(L.arp_anneal_arpeggio_list(None, P_(self.Gi), P_(self.hi), P_(self.T1t), self.T1t.numel(), self.q,
                                              P, self.R1, D, B, P_(self.inst), k0, k1, 0, P_(self.thr), gate, 0, None,
                                              None, P_(self.spins1), P_(self.de1), P_(self.st1),
                                              P_(self.used1) if learned else None, s), 'anneal stage 1')
        mark(3)
        lt = P_(ltrue)
        _lib.check(L.arp_decode_select_list(P_(H), P_(y), P_(self.xg), P_(self.spins1), B, self.R1, P, D, Mr, M, 1,
                                            lt, None, None, None, None, P_(self.lhat),
                                            P_(self.biterr) if ltrue is not None else None, P_(self.d1), s),
                   'decode stage 1')
        mark(4)
        _lib.check(L.arp_certify(P_(self.d1), P_(n0t), n0s, self.thr_scale, B, P_(self.list), P_(self.count), s),
                   'certify')
        mark(5)
        _lib.check(L.arp_anneal_arpeggio_list(None, P_(self.Gi), P_(self.hi), P_(self.T2t), self.T2t.numel(), self.q,
                                              P, self.R2, D, B, P_(self.inst), k0, k1, 0, P_(self.thr), gate, self.R1,
                                              P_(self.list), P_(self.count), P_(self.spins2), P_(self.de2),
                                              P_(self.st2), P_(self.used2) if learned else None, s), 'anneal stage 2')
        mark(6)
        if events is not None:
            events[1].record(ts)
        _lib.check(L.arp_decode_select_list(P_(H), P_(y), P_(self.xg), P_(self.spins2), B, self.R2, P, D, Mr, M, 1,
                                            lt, P_(self.list), P_(self.count), P_(self.lhat), P_(self.d1),
                                            P_(self.lhat), P_(self.biterr) if ltrue is not None else None,
                                            P_(self.d2), s), 'decode stage 2')
        mark(7)
        return self.lhat, (self.biterr if ltrue is not None else None)

    def open_frames(self):
        """Host copy of the frame indices stage 2 annealed (sorted)."""
        n = int(self.count.item())
        return torch.sort(self.list[:n]).values.cpu()


def detect_chain(H, y, N0, M, R1=4, L1=24, R2=16, L2=64, c=3.0, model=None, gate=True, key=0, inst0=0, T0=0.1,
                 T1=0.1, q=0.0, ltrue=None, Q=QSCALE, work=None, noalloc=False):
    """The whole chain through the single C entry point arp_detect_adaptive (R2 = 0: one stage). Same results as
    AdaptivePipeline (DetectPipeline for R2 = 0) on the same arguments. noalloc: arp_detect_adaptive_noalloc, whose
    stage buffers are the library's own device array (no workspace). Returns (lhat, biterr or None)."""
    H, y = _f32(H), _f32(y)
    B, Mr, D = H.shape
    L = _lib.load()
    T1l, qi = arpeggio_schedule(L1 // 2, T0, T1, q, 1, Q)
    schedule_check(T1l, qi, 2, D, Q)
    T1t = torch.tensor(T1l, dtype=torch.int32, device='cuda')
    T2t = None
    if R2 > 0:
        T2l, _ = arpeggio_schedule(L2 // 2, T0, T1, q, 1, Q)
        schedule_check(T2l, qi, 2, D, Q)
        T2t = torch.tensor(T2l, dtype=torch.int32, device='cuda')
    inst = ((torch.arange(B, dtype=torch.int64) + inst0) & MASK).to(torch.int32).to('cuda')
    n0t, n0s = _n0_arg(N0, B)
    need = L.arp_detect_workspace_bytes(B, D, R1, R2, 1 if model is not None else 0)
    if work is None and not noalloc:
        work = torch.empty(need, dtype=torch.uint8, device='cuda')
    lhat = torch.empty(B, D, dtype=torch.int8, device='cuda')
    lt = ltrue.to(device='cuda', dtype=torch.int8).contiguous() if ltrue is not None else None
    be = torch.empty(B, dtype=torch.int32, device='cuda') if lt is not None else None
    mw = None
    if model is not None:
        W = model._weights_nmajor()
        P_ = model.params
        p = lambda t: t.data_ptr()
        mw = _lib.MlpWeights(p(W['W1']), p(P_['b1']), p(W['W2']), p(P_['b2']), p(W['W3']), p(P_['b3']), model.F,
                             model.H, float(model.inv_temp))
    nr = Mr // 2
    P_ = _lib.ptr
    if noalloc:
        _lib.check(L.arp_detect_adaptive_noalloc(P_(H), P_(y), P_(n0t), n0s, B, Mr, D, int(M), P_(T1t), L1, R1, P_(T2t),
                                                 L2, R2, qi, float(nr + c * nr ** 0.5), P_(inst), key & MASK,
                                                 (key >> 32) & MASK, ctypes.byref(mw) if mw is not None else None,
                                                 1 if gate else 0, P_(lt), P_(lhat), P_(be), _lib.stream_ptr()),
                   'detect_adaptive_noalloc')
        return lhat, be
    _lib.check(L.arp_detect_adaptive(P_(H), P_(y), P_(n0t), n0s, B, Mr, D, int(M), P_(T1t), L1, R1, P_(T2t), L2, R2,
                                     qi, float(nr + c * nr ** 0.5), P_(inst), key & MASK, (key >> 32) & MASK,
                                     ctypes.byref(mw) if mw is not None else None, 1 if gate else 0, P_(lt),
                                     P_(lhat), P_(be), P_(work), work.numel(), _lib.stream_ptr()), 'detect_adaptive')
    return lhat, be
