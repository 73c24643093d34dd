"""Soft output: max-log bit LLRs on the GPU, the reference's soft-output detectors for this path.

LLRs are raw (positive = bit 1), (B, D nb) dimension-major with the MSB first within each real dimension, as
arpeggio/mimo/softlib.py:1-9 defines them, and the hard decision comes with them.

    detect_llr(b, method, key=0, cfg=None) -> (llr (B, D nb) float32, l_hat (B, D) int8)

methods (arpeggio/mimo/softlib.py:19-65):
  'mmse'    the Gaussian max-log LLRs of the LMMSE output (mimo.mmse_soft_llr, arpeggio/mimo/mimo.py:111-129);
  'anedNA'  NA Chord anneals of the delta-Ising problem (K = 1, around the MMSE decision), their best states + the MMSE
            decision + the single-bit-flip neighbours of the chosen one (mimo.select + mimo.list_llr, mimo.py:421-462);
  'arpdNA'  the same list detector with NA first-accept Arpeggio anneals (no reference counterpart for this pairing:
            the reference's list detectors use Chord, AHC or CIM anneals).

The list LLRs clip at softlib's BIG (1e4); list_llr's own default clip is mimo.list_llr's (50)."""
import torch

from . import _lib
from .detector import DetectPipeline, _f32, _n0_arg, front_end

BIG = 1e4
CHORD_X32 = dict(S=256, t0=0.3, t1=0.2, q=0.4, kappa=0.3)        # fxdetect.SCHED['x32'] (arpeggio/mimo/fxdetect.py)
ARPEGGIO_X32 = dict(passes=64, T0=0.1, T1=0.1, q=0.0)             # abl_far_32_0.1_0.1_0_rand


def mmse_llr(z, mu, M):
    """LLRs (B, D nb) from the front end's LMMSE output z and bias mu (B, D) float32 device tensors."""
    z, mu = _f32(z), _f32(mu)
    B, D = z.shape
    nb = 2 if M == 16 else 1
    llr = torch.empty(B, D * nb, dtype=torch.float32, device=z.device)
    _lib.check(_lib.load().arp_soft_mmse_llr(_lib.ptr(z), _lib.ptr(mu), B, D, int(M), _lib.ptr(llr),
                                             _lib.stream_ptr()), 'soft_mmse_llr')
    return llr


def list_llr(H, y, xg, spins, M, N0, P=2, delta=True, clip=50.0, want_dist=False):
    """List LLRs from the anneal outputs spins (B, R, N) int8 (dDI copies around xg when delta, else RI spins), the
    MMSE levels xg (B, D) and the decision's single-bit neighbours. N0 scalar or (B,). Returns (llr, lhat[, dist])."""
    H, y = _f32(H), _f32(y)
    B, Mr, D = H.shape
    R = spins.shape[1] if spins is not None else 0
    xg = xg.to(torch.int8).contiguous()
    sp = spins.to(torch.int8).contiguous() if spins is not None else None
    nb = 2 if M == 16 else 1
    n0t, n0s = _n0_arg(N0, B)
    llr = torch.empty(B, D * nb, dtype=torch.float32, device=H.device)
    lhat = torch.empty(B, D, dtype=torch.int8, device=H.device)
    dist = torch.empty(B, R + 1, dtype=torch.float32, device=H.device) if want_dist else None
    P_ = _lib.ptr
    _lib.check(_lib.load().arp_soft_list_llr(P_(H), P_(y), P_(xg), P_(sp), B, R, P, D, Mr, int(M), 1 if delta else 0,
                                             P_(n0t), n0s, float(clip), P_(lhat), P_(dist), P_(llr),
                                             _lib.stream_ptr()), 'soft_list_llr')
    return (llr, lhat, dist) if want_dist else (llr, lhat)


def detect_llr(b, method, key=0, cfg=None, inst0=0):
    """The soft detector `method` on a batch b = {H (B, 2Nr, 2Nt), y (B, 2Nr), N0 (scalar or (B,)), M} of device
    tensors. cfg: {'aned': Chord schedule (S, t0, t1, q, kappa), 'arpd': Arpeggio schedule (passes, T0, T1, q)}."""
    H, y, N0, M = _f32(b['H']), _f32(b['y']), b['N0'], int(b['M'])
    B, Mr, D = H.shape
    cfg = cfg or {}
    if method == 'mmse':
        o = front_end(H, y, N0, M, 1, want=('xg', 'z', 'mu'))
        return mmse_llr(o['z'], o['mu'], M), o['xg']
    if method.startswith('aned') or method.startswith('arpd'):
        na = int(method[4:])
        if method.startswith('aned'):
            an = dict(CHORD_X32, **cfg.get('aned', {}))
            pipe = DetectPipeline(B, Mr, D, M, na, policy='chord', cycles=an['S'], T0=an['t0'], T1=an['t1'],
                                  q=an['q'], kappa=an.get('kappa', 0.0), key=key, inst0=inst0)
        else:
            an = dict(ARPEGGIO_X32, **cfg.get('arpd', {}))
            pipe = DetectPipeline(B, Mr, D, M, na, policy='arpeggio', cycles=an['passes'], T0=an['T0'], T1=an['T1'],
                                  q=an['q'], key=key, inst0=inst0)
        pipe.run(H, y, N0)
        llr, lhat = list_llr(H, y, pipe.xg, pipe.spins, M, N0, P=pipe.P, delta=True, clip=BIG)
        return llr, lhat
    raise ValueError('unknown soft method %r (mmse, anedNA, arpdNA)' % method)
