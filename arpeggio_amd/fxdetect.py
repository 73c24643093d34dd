"""Detection with the fixed-point Chord kernel, all on the GPU: the reference's fxdetect.detect
(arpeggio/mimo/fxdetect.py:24-43).

MMSE -> Ising mapping (RI-MIMO, mimo.ising, or dDI K = 1 around the MMSE decision, mimo.ising_delta_at) -> fixed point
(chordfx.to_fixed) -> Chord -> levels -> the lower ||y - Hx||^2 of the candidates and the MMSE decision. For RI the
spins map one-to-one to levels, so the kernel's lowest-energy anneal per instance is the candidate (mode 'best'); for
dDI every anneal's best state goes to the distance-based selection (mode 'chains'), as the reference does.

    detect(b, sc, R, kern=None, seed=0, inst0=0, stats=None) -> levels (B, D) int8

b = {H (B, 2Nr, 2Nt), y (B, 2Nr), N0, M} device tensors; sc a schedule of SCHED; kern a chordfx.Kernel (one is made
when None)."""
import torch

from . import _lib
from .chordfx import QSCALE, Kernel, to_fixed
from .detector import _f32, decode_select, front_end

# the reference's tuned schedules (fxdetect.py:16-21)
SCHED = {
    'q16': dict(kind='ri', Nt=16, Nr=16, M=4, S=128, t0=0.7, t1=0.5, q=0.6, kappa=0.0),
    'q32': dict(kind='ri', Nt=32, Nr=32, M=4, S=128, t0=0.5, t1=0.4, q=0.4, kappa=0.0),
    'x16': dict(kind='ddi', Nt=16, Nr=16, M=16, S=256, t0=0.3, t1=0.3, q=0.4, kappa=0.0),
    'x32': dict(kind='ddi', Nt=32, Nr=32, M=16, S=256, t0=0.3, t1=0.2, q=0.4, kappa=0.3),
}
RI_REG = 0.15   # mimo.ising's regularisation weight r (mimo.py:132)
_DUMMY = dict(J=torch.zeros(1, 2, 2), h=torch.zeros(1, 2), h0=torch.zeros(1, 2))


def ri_fixed(H, y, z, M, S, t0, t1, q, kappa=0.0, r=RI_REG, Q=QSCALE):
    """The RI-MIMO problem in fixed point (arp_ising_ri): the dict chordfx.Kernel.chord takes (J (B, N, N) int16,
    h = h0 + 2 r s_P and h0 (B, N) int32, the schedule T, c, q, S) plus the per-frame scale. z: the LMMSE output
    (B, D) the regulariser's spins are drawn from."""
    H, y, z = _f32(H), _f32(y), _f32(z)
    B, Mr, D = H.shape
    N = 2 * D if M == 16 else D
    dev = H.device
    J = torch.empty(B, N, N, dtype=torch.int16, device=dev)
    h = torch.empty(B, N, dtype=torch.int32, device=dev)
    h0 = torch.empty(B, N, dtype=torch.int32, device=dev)
    scale = torch.empty(B, dtype=torch.float32, device=dev)
    P_ = _lib.ptr
    _lib.check(_lib.load().arp_ising_ri(P_(H), P_(y), P_(z), B, Mr, D, int(M), float(r), float(Q), P_(J), P_(h),
                                        P_(h0), P_(scale), _lib.stream_ptr()), 'ising_ri')
    F = to_fixed(_DUMMY, S, t0, t1, q, kappa, Q=Q)
    return dict(J=J, h=h, h0=h0, T=F['T'], c=F['c'], q=F['q'], S=S, Q=Q, scale=scale)


def detect(b, sc, R, kern=None, seed=0, inst0=0, stats=None):
    """fxdetect.detect on the GPU: returns the decided levels (B, D) int8."""
    H, y, N0, M = _f32(b['H']), _f32(b['y']), b['N0'], int(b['M'])
    kern = kern or Kernel()
    if sc['kind'] == 'ri':
        o = front_end(H, y, N0, M, 1, want=('xg', 'z'))
        F = ri_fixed(H, y, o['z'], M, sc['S'], sc['t0'], sc['t1'], sc['q'], sc.get('kappa', 0.0))
        s8, _ = kern.chord(F, R, seed, P=1, inst0=inst0, mode='best', stats=stats)
        lhat, _, _ = decode_select(H, y, o['xg'], s8[:, None, :].contiguous(), M, P=1, delta=False)
        return lhat
    o = front_end(H, y, N0, M, 1, want=('xg', 'J', 'hN'))
    F = to_fixed(_DUMMY, sc['S'], sc['t0'], sc['t1'], sc['q'], sc.get('kappa', 0.0))
    F.update(J=o['J'], h=o['hN'], h0=o['hN'])
    s8, _ = kern.chord(F, R, seed, P=2, inst0=inst0, mode='chains', stats=stats)
    lhat, _, _ = decode_select(H, y, o['xg'], s8, M, P=2, delta=True)
    return lhat
