"""Fixed-point Arpeggio on MI355X: drop-in for the reference's arpeggiofx module (arpeggio/sim/arpeggiofx.py).

solve(G, h, P, R, T, q, inst, key, start) keeps the reference's call shape and return values (numpy, after a
synchronisation, as arpeggiofx.py:43-44). solve_device(...) is the asynchronous form: device tensors in and out,
launched on torch's current HIP stream. mode='par' selects the engine's parallel pass
(arpeggio/sim/local_code_a59/arpeggio_ref59.py:72-82) instead of first-accept."""
import numpy as np
import torch

from . import _lib

M32 = 0xFFFFFFFF
INT31 = 2 ** 31


def _range_check(G, h, P, T, q):
    """The kernel keeps fields in int32 (as arpeggio_fx.cu:46 does) and multiplies 4T in 32 bits: refuse inputs for
    which that is not exact. G (B, D, D), h (B, D) tensors on any device; T a host sequence."""
    Tmax = max(int(t) for t in T)
    Tmin = min(int(t) for t in T)
    if Tmin < 0 or Tmax >= 2 ** 29:
        raise ValueError('arpeggio: pass temperatures must lie in [0, 2^29), got [%d, %d]' % (Tmin, Tmax))
    if G.numel() == 0:
        return
    Ga = G.abs().to(torch.int64)
    wmax = int((h.abs().to(torch.int64) + 2 * P * Ga.sum(-1)).max())
    dmax = int(torch.diagonal(Ga, dim1=-2, dim2=-1).max())
    if wmax + 2 * dmax + abs(int(q)) + 2 * Tmax >= INT31:
        raise ValueError('arpeggio: |h| + 2P sum|G| = %d exceeds the int32 field range' % wmax)
    if 6 * Tmax + 2 * dmax + abs(int(q)) >= INT31:
        raise ValueError('arpeggio: 4T + 2|G_jj| + |q| + 2T exceeds the int32 decision range')


def schedule_check(T, q, P, D, Q=4096):
    """The ranges of _range_check for couplings the front end writes (|G_ij|, |h_i| <= Q after its scaling): a check of
    the schedule alone, for pipelines whose integer Gram is produced on the device."""
    Tl = [int(t) for t in T]
    if min(Tl) < 0 or max(Tl) >= 2 ** 29:
        raise ValueError('arpeggio: pass temperatures must lie in [0, 2^29), got [%d, %d]' % (min(Tl), max(Tl)))
    wmax = int(Q) * (1 + 2 * P * D)
    if wmax + 2 * int(Q) + abs(int(q)) + 2 * max(Tl) >= INT31 or 6 * max(Tl) + 2 * int(Q) + abs(int(q)) >= INT31:
        raise ValueError('arpeggio: schedule (max T %d, q %d) exceeds the int32 field range at D = %d' %
                         (max(Tl), q, D))


def solve_device(G, h, P, R, T, q, inst, key, start, mode='arpeggio', transposed=False):
    """Device form. G (B, D, D) integer Gram (G[b, i, j] = G_ij; pass transposed=True if it is already Gt), h (B, D),
    T (L,) integers, inst (B,) integers < 2^32, key 64-bit, start 'rand'/'bal'. Returns device tensors
    (spins int8 (B, R, P D), de int64 (B, R), st int64 (B, R, 2))."""
    G = torch.as_tensor(G)
    h = torch.as_tensor(h)
    assert G.dim() == 3 and G.shape[1] == G.shape[2], 'G must be (B, D, D)'
    B, D, _ = G.shape
    assert h.shape == (B, D), 'h must be (B, D)'
    assert D <= 128 and P in (2, 4, 6) and len(T) >= 1
    assert start in ('rand', 'bal') and mode in ('arpeggio', 'par')
    Tl = [int(x) for x in (T.tolist() if torch.is_tensor(T) else T)]
    _range_check(G, h, P, Tl, q)
    Gt = (G if transposed else G.transpose(1, 2)).to(device='cuda', dtype=torch.int32).contiguous()
    hh = h.to(device='cuda', dtype=torch.int32).contiguous()
    Tt = torch.tensor(Tl, dtype=torch.int32, device='cuda')
    it = torch.as_tensor(np.asarray(torch.as_tensor(inst).cpu().numpy(), dtype=np.int64) & M32).to(torch.int32).cuda()
    out_s = torch.empty(B, R, P * D, dtype=torch.int8, device='cuda')
    out_de = torch.empty(B, R, dtype=torch.int64, device='cuda')
    out_st = torch.empty(B, R, 2, dtype=torch.int64, device='cuda')
    L = _lib.load()
    _lib.check(L.arp_anneal_arpeggio(_lib.ptr(Gt), _lib.ptr(hh), _lib.ptr(Tt), len(Tl), int(q), int(P), int(R), int(D),
                                     int(B), _lib.ptr(it), key & M32, (key >> 32) & M32, 0 if start == 'rand' else 1,
                                     0 if mode == 'arpeggio' else 1, _lib.ptr(out_s), _lib.ptr(out_de),
                                     _lib.ptr(out_st), _lib.stream_ptr()), 'arpeggio')
    return out_s, out_de, out_st


@torch.no_grad()
def solve(G, h, P, R, T, q, inst, key, start, mode='arpeggio'):
    """G (B, D, D), h (B, D): integers from chord_ref.compress; P copies; R anneals per sub-problem; T: L integers, one
    per pass (pass k tests copy k % P); q integer; inst (B,) integers below 2^32; key: 64-bit Philox key (k0 = low
    word, k1 = high word); start 'rand' or 'bal'. Returns best spins (B, R, P*D) int8 (copy-major), best energy change
    (B, R) int64, and (B, R, 2) int64 flips and comparator passes (tests) per anneal."""
    s, de, st = solve_device(G, h, P, R, T, q, inst, key, start, mode=mode)
    torch.cuda.synchronize()
    return s.cpu().numpy(), de.cpu().numpy(), st.cpu().numpy()


def solve_float_device(G, h, P, R, T, q, inst, key, start, want_fields=False, transposed=False):
    """FP32 twin of solve_device: G (B, D, D), h (B, D) float scaled couplings, T (L,) float temperatures per pass,
    q float. Returns device tensors (spins int8, best dE float32 (E0 units), st int64 (B, R, 2), fields or None)."""
    G = torch.as_tensor(G)
    h = torch.as_tensor(h)
    B, D, _ = G.shape
    assert D <= 128 and P in (2, 4, 6) and len(T) >= 1 and start in ('rand', 'bal')
    Gt = (G if transposed else G.transpose(1, 2)).to(device='cuda', dtype=torch.float32).contiguous()
    hh = h.to(device='cuda', dtype=torch.float32).contiguous()
    Tt = torch.as_tensor(T, dtype=torch.float32).to('cuda').contiguous()
    it = torch.as_tensor(np.asarray(torch.as_tensor(inst).cpu().numpy(), dtype=np.int64) & M32).to(torch.int32).cuda()
    out_s = torch.empty(B, R, P * D, dtype=torch.int8, device='cuda')
    out_de = torch.empty(B, R, dtype=torch.float32, device='cuda')
    out_st = torch.empty(B, R, 2, dtype=torch.int64, device='cuda')
    out_w = torch.empty(B, R, D, dtype=torch.float32, device='cuda') if want_fields else None
    L = _lib.load()
    _lib.check(L.arp_anneal_arpeggio_f32(_lib.ptr(Gt), _lib.ptr(hh), _lib.ptr(Tt), Tt.numel(), float(q), int(P),
                                         int(R), int(D), int(B), _lib.ptr(it), key & M32, (key >> 32) & M32,
                                         0 if start == 'rand' else 1, _lib.ptr(out_s), _lib.ptr(out_de),
                                         _lib.ptr(out_st), _lib.ptr(out_w), _lib.stream_ptr()), 'arpeggio_f32')
    return out_s, out_de, out_st, out_w


def chord_ref_device(G, h, P, T, c, q, R, seed, inst0=0, rng='xs', inst=None, transposed=False):
    """The production engine's lane-shared Chord on the GPU (arp_anneal_chord_lane), with chord_ref.chord_ref's
    arguments (arpeggio/sim/x3model/chord_ref.py:73-123): G (B, D, D) / h (B, D) compressed integers, P copies,
    schedule T and Onsager terms c (S each, c may be None), pinning q, R anneals, 64-bit seed, instance words inst0 + b
    (or inst (B,)), rng 'xs' (xorshift32 lane streams, 16-bit draws) or 'philox'. Returns device tensors: best spins
    (B, R, P D) int8, best de (B, R) int64 and (B, R, 2) int64 whose [..., 0] is each anneal's flip count (the
    reference returns flips per step summed over anneals)."""
    G = torch.as_tensor(G)
    h = torch.as_tensor(h)
    assert G.dim() == 3 and G.shape[1] == G.shape[2], 'G must be (B, D, D)'
    B, D, _ = G.shape
    assert h.shape == (B, D), 'h must be (B, D)'
    if rng not in ('xs', 'philox'):
        raise ValueError("rng must be 'xs' or 'philox'")
    Tl = [int(x) for x in (T.tolist() if torch.is_tensor(T) else T)]
    cl = [0] * len(Tl) if c is None else [int(x) for x in (c.tolist() if torch.is_tensor(c) else c)]
    if len(cl) != len(Tl):
        raise ValueError('chord_ref: %d Onsager terms for %d steps' % (len(cl), len(Tl)))
    _range_check(G, h, P, Tl, abs(int(q)) + max(abs(x) for x in cl))
    Gt = (G if transposed else G.transpose(1, 2)).to(device='cuda', dtype=torch.int32).contiguous()
    hh = h.to(device='cuda', dtype=torch.int32).contiguous()
    Tt = torch.tensor(Tl, dtype=torch.int32, device='cuda')
    Ct = torch.tensor(cl, dtype=torch.int32, device='cuda')
    iv = np.arange(B, dtype=np.int64) + int(inst0) if inst is None else np.asarray(
        torch.as_tensor(inst).cpu().numpy(), dtype=np.int64)
    it = torch.as_tensor(iv & M32).to(torch.int32).cuda()
    out_s = torch.empty(B, R, P * D, dtype=torch.int8, device='cuda')
    out_de = torch.empty(B, R, dtype=torch.int64, device='cuda')
    out_st = torch.empty(B, R, 2, dtype=torch.int64, device='cuda')
    _lib.check(_lib.load().arp_anneal_chord_lane(
        _lib.ptr(Gt), _lib.ptr(hh), _lib.ptr(Tt), _lib.ptr(Ct), len(Tl), int(q), int(P), int(R), int(D), int(B),
        _lib.ptr(it), seed & M32, (seed >> 32) & M32, 1 if rng == 'xs' else 0, None, None, _lib.ptr(out_s),
        _lib.ptr(out_de), _lib.ptr(out_st), _lib.stream_ptr()), 'chord_lane')
    return out_s, out_de, out_st
