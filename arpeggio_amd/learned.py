"""Learned warm start for the first-accept annealer (no reference counterpart).

A small MLP maps per-frame features (matched-filter output H^T y, Gram diagonal, noise variance, and the LMMSE hard
levels and soft residual the dDI spins are relative to; written by the front-end kernel) to per-spin logits; the
logits become per-spin start thresholds, and the annealer starts each anneal from a Philox draw of that distribution,
keeping the plain random start whenever the learned start's Ising energy is higher (arp_anneal_arpeggio_ws).

Network: F -> 256 -> 256 -> P*D, GELU (tanh form), BF16 weights and activations with FP32 accumulation and FP32
master weights (Adam). Forward, backward, loss and optimiser all run in the library's HIP kernels (arp_mlp_*).
Labels: canonical spins of the detection a long anneal makes (1024 passes, 8 anneals per frame): for the chosen levels
l and the LMMSE levels x_g, the correction d = l - x_g in {-2, 0, 2} becomes copy-0 spin +1 iff d >= 0 and copy-1 spin
+1 iff d = 2 (the dDI energy does not depend on which copy carries a +1, so this choice loses nothing)."""
import math
import os

import numpy as np
import torch

from . import _lib
from .detector import arpeggio_schedule

HIDDEN = 256
MASK = 0xFFFFFFFF


def feat_ld(D):
    """Feature row length: 4 D + 1 features padded to a multiple of 32 (the GEMM's K step)."""
    return (4 * D + 1 + 31) // 32 * 32


def _bf16(x):
    return x.to(torch.bfloat16)


class WarmStartMLP:
    def __init__(self, D, P=2, hidden=HIDDEN, seed=0, device='cuda'):
        self.D, self.P, self.H = D, P, hidden
        self.F = feat_ld(D)
        self.N = P * D
        g = torch.Generator(device='cpu')
        g.manual_seed(seed)
        nf = 4 * D + 1

        def init(fan_in, shape, rows=None):
            w = torch.randn(*shape, generator=g) * math.sqrt(2.0 / fan_in)
            if rows is not None:
                w[rows:] = 0.0
            return w.to(device).contiguous()

        self.params = dict(W1=init(nf, (self.F, hidden), rows=nf), b1=torch.zeros(hidden, device=device),
                           W2=init(hidden, (hidden, hidden)), b2=torch.zeros(hidden, device=device),
                           W3=init(hidden, (hidden, self.N)) * 0.5, b3=torch.zeros(self.N, device=device))
        self.moments = {k: (torch.zeros_like(v), torch.zeros_like(v)) for k, v in self.params.items()}
        self.wb = {k: _bf16(self.params[k]) for k in ('W1', 'W2', 'W3')}
        self.step_count = 0
        self.inv_temp = 1.0     # logit scale of the start thresholds, calibrated after training
        self.lib = _lib.load()

    # ---------------------------------------------------------------- persistence
    def state_dict(self):
        return {k: v.detach().cpu().numpy() for k, v in self.params.items()}

    def save(self, path):
        np.savez_compressed(path, D=self.D, P=self.P, H=self.H, inv_temp=self.inv_temp, **self.state_dict())

    @classmethod
    def load(cls, path, device='cuda'):
        z = np.load(path, allow_pickle=False)
        m = cls(int(z['D']), int(z['P']), int(z['H']), device=device)
        for k in m.params:
            m.params[k] = torch.as_tensor(z[k]).to(device).contiguous()
        m.wb = {k: _bf16(m.params[k]) for k in ('W1', 'W2', 'W3')}
        if 'inv_temp' in z.files:
            m.inv_temp = float(z['inv_temp'])
        return m

    # ---------------------------------------------------------------- kernels
    def _gemm(self, A, sam, sak, B, sbk, sbn, M, N, K, bias=None, act=0, Cf=None, Cpre=None, Cb=None, ldc=None,
              stream=None):
        p = _lib.ptr
        _lib.check(self.lib.arp_mlp_gemm(p(A), sam, sak, p(B), sbk, sbn, M, N, K, p(bias), act, p(Cf), p(Cpre), p(Cb),
                                         ldc if ldc is not None else N, _lib.stream_ptr(stream)), 'mlp_gemm')

    def forward(self, X, save=False, stream=None):
        """X (M, F) bf16 features -> logits (M, N) fp32 (and the activations backward needs if save). The GEMMs run on
        `stream` (default: torch's current stream); the outputs are allocated on the current stream's pool."""
        M, H, N, F = X.shape[0], self.H, self.N, self.F
        dev = X.device
        P_ = self.params
        A1 = torch.empty(M, H, device=dev) if save else None
        H1 = torch.empty(M, H, dtype=torch.bfloat16, device=dev)
        self._gemm(X, F, 1, self.wb['W1'], H, 1, M, H, F, P_['b1'], 1, None, A1, H1, stream=stream)
        A2 = torch.empty(M, H, device=dev) if save else None
        H2 = torch.empty(M, H, dtype=torch.bfloat16, device=dev)
        self._gemm(H1, H, 1, self.wb['W2'], H, 1, M, H, H, P_['b2'], 1, None, A2, H2, stream=stream)
        Z = torch.empty(M, N, device=dev)
        self._gemm(H2, H, 1, self.wb['W3'], N, 1, M, N, H, P_['b3'], 0, Z, stream=stream)
        if save:
            return Z, dict(X=X, A1=A1, H1=H1, A2=A2, H2=H2)
        return Z

    def gradients(self, X, Y):
        """Mean BCE of the logits against spins Y (M, N) int8 +-1, and its gradients (fp32, parameter shapes)."""
        M, H, N, F = X.shape[0], self.H, self.N, self.F
        dev = X.device
        Z, S = self.forward(X, save=True)
        L = self.lib
        s = _lib.stream_ptr()
        p = _lib.ptr
        loss = torch.zeros(1, device=dev)
        dZb = torch.empty(M, N, dtype=torch.bfloat16, device=dev)
        dZf = torch.empty(M, N, device=dev)
        _lib.check(L.arp_mlp_bce(p(Z), N, p(Y), M, N, p(dZb), p(dZf), p(loss), s), 'bce')
        g = {}
        g['W3'] = torch.empty(H, N, device=dev)
        self._gemm(S['H2'], 1, H, dZb, N, 1, H, N, M, Cf=g['W3'])           # H2^T dZ
        g['b3'] = torch.empty(N, device=dev)
        _lib.check(L.arp_mlp_colsum(p(dZf), N, M, N, p(g['b3']), s), 'colsum')
        dH2 = torch.empty(M, H, device=dev)
        self._gemm(dZb, N, 1, self.wb['W3'], 1, N, M, H, N, Cf=dH2)         # dZ W3^T
        dA2b = torch.empty(M, H, dtype=torch.bfloat16, device=dev)
        dA2f = torch.empty(M, H, device=dev)
        _lib.check(L.arp_mlp_gelu_bwd(p(dH2), p(S['A2']), M * H, p(dA2b), p(dA2f), s), 'gelu_bwd')
        g['W2'] = torch.empty(H, H, device=dev)
        self._gemm(S['H1'], 1, H, dA2b, H, 1, H, H, M, Cf=g['W2'])           # H1^T dA2
        g['b2'] = torch.empty(H, device=dev)
        _lib.check(L.arp_mlp_colsum(p(dA2f), H, M, H, p(g['b2']), s), 'colsum')
        dH1 = torch.empty(M, H, device=dev)
        self._gemm(dA2b, H, 1, self.wb['W2'], 1, H, M, H, H, Cf=dH1)         # dA2 W2^T
        dA1b = torch.empty(M, H, dtype=torch.bfloat16, device=dev)
        dA1f = torch.empty(M, H, device=dev)
        _lib.check(L.arp_mlp_gelu_bwd(p(dH1), p(S['A1']), M * H, p(dA1b), p(dA1f), s), 'gelu_bwd')
        g['W1'] = torch.empty(F, H, device=dev)
        self._gemm(X, 1, F, dA1b, H, 1, F, H, M, Cf=g['W1'])                 # X^T dA1
        g['b1'] = torch.empty(H, device=dev)
        _lib.check(L.arp_mlp_colsum(p(dA1f), H, M, H, p(g['b1']), s), 'colsum')
        return loss, g

    def adam_step(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        s = _lib.stream_ptr()
        p = _lib.ptr
        for k, W in self.params.items():
            m, v = self.moments[k]
            wb = self.wb.get(k)
            _lib.check(self.lib.arp_mlp_adam(p(W), p(grads[k]), p(m), p(v), W.numel(), lr, b1, b2, eps,
                                             self.step_count, p(wb), s), 'adam')

    def _weights_nmajor(self):
        """bf16 weights transposed to n-major (the fused kernel's operand layout), rebuilt after optimiser steps."""
        if getattr(self, '_wt_step', None) != self.step_count or getattr(self, '_wt_src', None) is not self.wb:
            self._wt = {k: self.wb[k].t().contiguous() for k in ('W1', 'W2', 'W3')}
            self._wt_step, self._wt_src = self.step_count, self.wb
        return self._wt

    def thresholds(self, X, inv_temp=1.0, out=None, fused=True, stream=None):
        """Per-spin start thresholds (M, N) uint32 (as int32 storage) from the features: one fused kernel
        (arp_mlp_start_fused), or the three GEMMs and the threshold kernel (fused=False); both give the same bits.
        The kernels run on `stream` (default: torch's current stream), ordered after the work already queued there."""
        M, N = X.shape[0], self.N
        thr = out if out is not None else torch.empty(M, N, dtype=torch.int32, device=X.device)
        if fused and self.H == 256 and N % 64 == 0 and N <= 256 and self.F <= 1016:
            W = self._weights_nmajor()
            P_ = self.params
            p = _lib.ptr
            _lib.check(self.lib.arp_mlp_start_fused(p(X), M, self.F, p(W['W1']), p(P_['b1']), p(W['W2']), p(P_['b2']),
                                                    p(W['W3']), p(P_['b3']), self.H, N, float(inv_temp), p(thr),
                                                    _lib.stream_ptr(stream)), 'mlp_start_fused')
            return thr
        Z = self.forward(X, stream=stream)
        if stream is not None:
            Z.record_stream(stream)
        _lib.check(self.lib.arp_start_thresholds(_lib.ptr(Z), N, M, N, float(inv_temp), _lib.ptr(thr),
                                                 _lib.stream_ptr(stream)), 'start_thresholds')
        return thr


def features(H, y, N0, M, K=1, Q=4096.0):
    """Front end with the learned-start features: returns dict(xg, Gi, hi, feat (B, feat_ld) bf16)."""
    from .detector import _f32, _n0_arg
    H, y = _f32(H), _f32(y)
    B, Mr, D = H.shape
    F = feat_ld(D)
    dev = H.device
    out = dict(xg=torch.empty(B, D, dtype=torch.int8, device=dev), Gi=torch.empty(B, D, D, dtype=torch.int32, device=dev),
               hi=torch.empty(B, D, dtype=torch.int32, device=dev), scale=torch.empty(B, device=dev),
               feat=torch.empty(B, F, dtype=torch.bfloat16, device=dev))
    n0t, n0s = _n0_arg(N0, B)
    p = _lib.ptr
    _lib.check(_lib.load().arp_front_end(p(H), p(y), p(n0t), n0s, B, Mr, D, int(M), int(K), float(Q), p(out['xg']),
                                         p(out['Gi']), p(out['hi']), p(out['scale']), None, None, None, None, None,
                                         None, None, None, p(out['feat']), F, None, _lib.stream_ptr()), 'front_end')
    return out


def canonical_spins(lhat, xg, D):
    """Labels: d = lhat - x_g in {-2, 0, 2} -> copy 0 +1 iff d >= 0, copy 1 +1 iff d = 2; (B, 2 D) int8."""
    d = lhat.to(torch.int16) - xg.to(torch.int16)
    s0 = torch.where(d >= 0, 1, -1)
    s1 = torch.where(d >= 2, 1, -1)
    return torch.cat([s0, s1], 1).to(torch.int8).contiguous()


def label_batch(H, y, N0, ltrue, M=16, R=8, passes=1024, T0=0.1, T1=0.1, q=0.0, key=0x5EED, inst0=0):
    """Features and labels for one batch of frames: a long first-accept anneal (R anneals, `passes` passes) picks the
    levels; returns (feat, labels, biterr of that detection)."""
    from .detector import decode_select
    fe = features(H, y, N0, M)
    B, D = fe['xg'].shape
    T, qi = arpeggio_schedule(passes // 2, T0, T1, q)
    Tt = torch.tensor(T, dtype=torch.int32, device='cuda')
    inst = ((torch.arange(B, dtype=torch.int64) + inst0) & MASK).to(torch.int32).cuda()
    spins = torch.empty(B, R, 2 * D, dtype=torch.int8, device='cuda')
    de = torch.empty(B, R, dtype=torch.int64, device='cuda')
    st = torch.empty(B, R, 2, dtype=torch.int64, device='cuda')
    p = _lib.ptr
    _lib.check(_lib.load().arp_anneal_arpeggio(p(fe['Gi']), p(fe['hi']), p(Tt), len(T), qi, 2, R, D, B, p(inst),
                                               key & MASK, (key >> 32) & MASK, 0, 0, p(spins), p(de), p(st),
                                               _lib.stream_ptr()), 'arpeggio')
    lhat, be, _ = decode_select(H, y, fe['xg'], spins, M, 2, True, ltrue=ltrue)
    return fe['feat'], canonical_spins(lhat, fe['xg'], D), be


def train(model, feats, labels, epochs=4, batch=8192, lr=2e-3, seed=0, log=None):
    """Mini-batch Adam on (feats (n, F) bf16, labels (n, N) int8); cosine learning-rate decay. Returns losses."""
    n = feats.shape[0]
    g = torch.Generator(device='cuda')
    g.manual_seed(seed)
    steps = epochs * (n // batch)
    losses = []
    k = 0
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g, device='cuda')
        for i in range(n // batch):
            idx = perm[i * batch:(i + 1) * batch]
            X = feats[idx].contiguous()
            Y = labels[idx].contiguous()
            loss, grads = model.gradients(X, Y)
            lr_k = lr * 0.5 * (1 + math.cos(math.pi * k / max(1, steps)))
            model.adam_step(grads, lr=lr_k)
            if i % 32 == 0:
                losses.append(float(loss))
                if log:
                    log('epoch %d step %d loss %.4f' % (ep, i, losses[-1]))
            k += 1
    return losses


DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'weights', 'warmstart_32x32_16qam.npz')
# 64 users x 64 rx, 16-QAM, Kronecker rho = 0.5 at both ends, labels and schedule T 0.08 -> 0.03 (train_warmstart
# --users 64 --rho 0.5 --T0 0.08 --T1 0.03 --snr-lo 24 --snr-hi 32)
WEIGHTS_64 = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'weights', 'warmstart_64x64_16qam_rho05.npz')
