"""CPU oracle for the delta-encoded Ising detector's kernel path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the benchmark's CPU-baseline leg, as the
checker. The product package (arpeggio_amd) never imports it.

Integer anneals and Philox come from the C restatement in arp_oracle.c (ctypes, built by `make -C oracle`).
The float front end is restated here in numpy float64 from the reference's text (paths relative to the reference):
  geom                     arpeggio/mimo/mimo.py:199-201
  to_fixed                 arpeggio/mimo/chordfx.py:60-76
  compress / expand        arpeggio/sim/x3model/chord_ref.py:47-70
  arpeggio_schedule        arpeggio/sim/local_code_a513/run_a513.py:83-87
  mmse, _mu, mmse_hard     arpeggio/mimo/mimo.py:86-109, 176-182
  ising (RI), ising_delta_at  arpeggio/mimo/mimo.py:132-149, 506-523
  spins_to_levels(_delta), levels_to_bits, dist, select, bit_errors, energy0
                           arpeggio/mimo/mimo.py:30-51, 185-195, 416-431
  boot_ratio, mcnemar      arpeggio/sim/a56an.py:74-106
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle_arp.so')
QSCALE = 4096
MASK32 = 0xFFFFFFFF
PAM = {4: dict(levels=[-1, 1], a=1 / math.sqrt(2), bits=1),
       16: dict(levels=[-3, -1, 1, 3], a=1 / math.sqrt(10), bits=2)}
# Random123 kat_vectors, philox4x32 10: counter, key -> output
KAT = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
       ((MASK32,) * 4, (MASK32, MASK32), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
       ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
        (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]

_lib = None


def build():
    import subprocess
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32, u32, u64, i64 = ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64
        L.orc_philox_kat.argtypes = [P, P, i32]
        L.orc_bal_pat.argtypes = [i32, u32]
        L.orc_arpeggio_fx.argtypes = [P, P, i32, i32, i32, i32, P, i32, i32, P, u64, i32, i32, P, P, P, i32]
        L.orc_arpeggio_trace.argtypes = [P, P, i32, i32, P, i32, i32, u32, u32, u64, i32, i32, P, P, P, P,
                                         ctypes.c_long, P]
        L.orc_arpeggio_ws.argtypes = [P, P, i32, i32, i32, i32, P, i32, i32, P, u64, P, i32, P, P, P, P, i32]
        L.orc_chord_fx.argtypes = [P, P, P, i32, i32, P, P, i32, i32, i32, i32, u64, i64, i32, P, P, P, i32]
        L.orc_chord_lane.argtypes = [P, P, i32, i32, i32, i32, P, P, i32, i32, P, u64, i32, P, P, P, i32]
        L.orc_arpeggio_float_f32.argtypes = [P, P, i32, i32, i32, i32, P, i32, ctypes.c_float, P, u64, i32, P, P, P,
                                             P, i32]
        L.orc_arpeggio_float_f64.argtypes = [P, P, i32, i32, i32, i32, P, i32, ctypes.c_double, P, u64, i32, P, P, P,
                                             P, i32]
        L.orc_chord_float_f32.argtypes = [P, P, P, i32, i32, P, P, i32, ctypes.c_float, i32, i32, u64, i64, P, P, P,
                                          P, i32]
        L.orc_chord_float_f64.argtypes = [P, P, P, i32, i32, P, P, i32, ctypes.c_double, i32, i32, u64, i64, P, P, P,
                                          P, i32]
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _threads(n):
    return int(n) if n else (os.cpu_count() or 1)


# ---------------------------------------------------------------- Philox and the integer anneals
def philox_kat(words):
    w = np.ascontiguousarray(np.asarray(words, dtype=np.uint64) & MASK32, dtype=np.uint32)
    out = np.zeros((w.shape[0], 4), dtype=np.uint32)
    lib().orc_philox_kat(_p(w), _p(out), w.shape[0])
    return out


def bal_pat(P, i):
    return lib().orc_bal_pat(P, i)


def arpeggio_fx(G, h, P, R, T, q, inst, key, start, mode='arpeggio', nthreads=None):
    """arpeggiofx.solve semantics (arpeggiofx.py:23-44) -> spins (B, R, P D) int8, de (B, R) int64, st (B, R, 2)."""
    G = np.ascontiguousarray(G, dtype=np.int32)
    h = np.ascontiguousarray(h, dtype=np.int32)
    B, D = h.shape
    T = np.ascontiguousarray(T, dtype=np.int32)
    inst = np.ascontiguousarray(np.asarray(inst, dtype=np.int64) & MASK32, dtype=np.uint32)
    s = np.zeros((B, R, P * D), dtype=np.int8)
    de = np.zeros((B, R), dtype=np.int64)
    st = np.zeros((B, R, 2), dtype=np.int64)
    rc = lib().orc_arpeggio_fx(_p(G), _p(h), B, D, P, R, _p(T), len(T), int(q), _p(inst), key & 0xFFFFFFFFFFFFFFFF,
                               0 if start == 'rand' else 1, 0 if mode == 'arpeggio' else 1, _p(s), _p(de), _p(st),
                               _threads(nthreads))
    assert rc == 0
    return s, de, st


def arpeggio_ws(G, h, P, R, T, q, inst, key, thr, gate=True, nthreads=None):
    """First-accept anneal from per-spin start thresholds thr (B, P D) uint32, energy-gated against the random start.
    Returns spins, de, st and used (B, R) int8."""
    G = np.ascontiguousarray(G, dtype=np.int32)
    h = np.ascontiguousarray(h, dtype=np.int32)
    B, D = h.shape
    T = np.ascontiguousarray(T, dtype=np.int32)
    inst = np.ascontiguousarray(np.asarray(inst, dtype=np.int64) & MASK32, dtype=np.uint32)
    thr = np.ascontiguousarray(thr, dtype=np.uint32)
    s = np.zeros((B, R, P * D), dtype=np.int8)
    de = np.zeros((B, R), dtype=np.int64)
    st = np.zeros((B, R, 2), dtype=np.int64)
    used = np.zeros((B, R), dtype=np.int8)
    rc = lib().orc_arpeggio_ws(_p(G), _p(h), B, D, P, R, _p(T), len(T), int(q), _p(inst), key & 0xFFFFFFFFFFFFFFFF,
                               _p(thr), 1 if gate else 0, _p(s), _p(de), _p(st), _p(used), _threads(nthreads))
    assert rc == 0
    return s, de, st, used


def arpeggio_trace(G, h, P, T, q, r, inst, key, start, mode='arpeggio'):
    """One anneal with its event record: dict(best, bs (P D) int8, flips, tests, n1, init, acc [(k, j)], tests_k)."""
    G = np.ascontiguousarray(G, dtype=np.int32)
    h = np.ascontiguousarray(h, dtype=np.int32)
    D = h.shape[0]
    T = np.ascontiguousarray(T, dtype=np.int32)
    L = len(T)
    mx = L * D
    ak = np.zeros(mx, dtype=np.int32)
    aj = np.zeros(mx, dtype=np.int32)
    tk = np.zeros(L, dtype=np.int32)
    bs = np.zeros(P * D, dtype=np.int8)
    ob = np.zeros(6, dtype=np.int64)
    rc = lib().orc_arpeggio_trace(_p(G), _p(h), D, P, _p(T), L, int(q), r & MASK32, inst & MASK32,
                                  key & 0xFFFFFFFFFFFFFFFF, 0 if start == 'rand' else 1, 0 if mode == 'arpeggio' else 1,
                                  _p(bs), _p(ob), _p(ak), _p(aj), mx, _p(tk))
    assert rc == 0
    n = int(ob[5])
    return dict(best=int(ob[0]), bs=bs, flips=int(ob[1]), tests=int(ob[2]), n1=int(ob[3]), init=int(ob[4]),
                acc=list(zip(ak[:n].tolist(), aj[:n].tolist())), tests_k=tk.tolist())


def chord_fx(F, R, seed, P=1, inst0=0, mode='chains', nthreads=None, stats=None):
    """chordfx.chord_fx semantics (chordfx.py:80-118) on numpy inputs F = dict(J, h, h0, T, c, q, S)."""
    J = np.ascontiguousarray(F['J'], dtype=np.int16)
    h = np.ascontiguousarray(F['h'], dtype=np.int32)
    h0 = np.ascontiguousarray(F['h0'], dtype=np.int32)
    T = np.ascontiguousarray(F['T'], dtype=np.int32)
    c = np.ascontiguousarray(F['c'], dtype=np.int32)
    B, N = h.shape
    S = int(F['S'])
    best = mode == 'best'
    s = np.zeros((B, N) if best else (B, R, N), dtype=np.int8)
    e = np.zeros((B,) if best else (B, R), dtype=np.int64)
    st = np.zeros(4, dtype=np.int64) if stats is not None else None
    rc = lib().orc_chord_fx(_p(J), _p(h), _p(h0), B, N, _p(T), _p(c), S, int(F['q']), P, R,
                            seed & 0xFFFFFFFFFFFFFFFF, int(inst0), 1 if best else 0, _p(s), _p(e), _p(st),
                            _threads(nthreads))
    assert rc == 0
    if stats is not None:
        stats[:] = st
    return s, e


def chord_lane(G, h, P, R, T, c, q, inst, seed, rng='xs', nthreads=None):
    """chord_ref.chord_ref (arpeggio/sim/x3model/chord_ref.py:73-123): the production engine's lane-shared Chord on
    G (B, D, D), h (B, D) integers, schedule T, c (S), pinning q, instance words inst (B), 64-bit seed; rng 'xs'
    (xorshift32 lane streams) or 'philox'. Returns spins (B, R, P D) int8, best de (B, R), flips (B, R)."""
    G = np.ascontiguousarray(G, dtype=np.int32)
    h = np.ascontiguousarray(h, dtype=np.int32)
    B, D = h.shape
    T = np.ascontiguousarray(T, dtype=np.int32)
    c = np.ascontiguousarray(c, dtype=np.int32)
    inst = np.ascontiguousarray(np.asarray(inst, dtype=np.int64) & MASK32, dtype=np.uint32)
    s = np.zeros((B, R, P * D), dtype=np.int8)
    de = np.zeros((B, R), dtype=np.int64)
    fl = np.zeros((B, R), dtype=np.int64)
    rc = lib().orc_chord_lane(_p(G), _p(h), B, D, P, R, _p(T), _p(c), len(T), int(q), _p(inst),
                              seed & 0xFFFFFFFFFFFFFFFF, 1 if rng == 'xs' else 0, _p(s), _p(de), _p(fl),
                              _threads(nthreads))
    assert rc == 0
    return s, de, fl


REF_ENGINE = os.path.join(HERE, '_ref', 'libape.so')
_eng = None


def ref_engine():
    """The reference's cycle-accurate Chord engine (arpeggio/sim/x3model/engine_model.cpp) built by `make -C oracle
    ref` into oracle/_ref, or None where it is not built (the reference sources are present in the build container
    only; the built library travels with the tree)."""
    global _eng
    if _eng is None and os.path.exists(REF_ENGINE):
        L = ctypes.CDLL(REF_ENGINE)
        L.engine_solve_batch.restype = ctypes.c_int
        _eng = L
    return _eng


def ref_engine_solve(G, h, P, R, T, c, q, inst, seed, philox=0, A=4, SEEDL=None, nthreads=8):
    """engine_solve_batch (engine_model.cpp:336-380) as the reference's engine_py.solve_batch calls it
    (engine_py.py:116-133): spins (B, R, P D) int8, de (B, R), cycles (B,), stats (B, 4) = D-ops, F-ops, I-ops,
    R-port busy cycles."""
    L = ref_engine()
    G = np.ascontiguousarray(G, dtype=np.int16)
    h = np.ascontiguousarray(h, dtype=np.int16)
    B, D, _ = G.shape
    SEEDL = SEEDL or max(1, D // 32)
    T = np.ascontiguousarray(T, dtype=np.int16)
    c = np.ascontiguousarray(c, dtype=np.int16)
    inst = np.ascontiguousarray(np.asarray(inst, dtype=np.int64) & MASK32, dtype=np.uint32)
    spins = np.zeros((B, R, P * D), np.int8)
    de = np.zeros((B, R), np.int64)
    cyc = np.zeros(B, np.int64)
    st = np.zeros((B, 4), np.int64)
    L.engine_solve_batch(D, A, SEEDL, philox, B, _p(G), _p(h), P, R, len(T), _p(T), _p(c), int(q), _p(inst),
                         ctypes.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), _p(spins), _p(de), _p(cyc), _p(st), nthreads)
    return spins, de, cyc, st


def arpeggio_float(G, h, P, R, T, q, inst, key, start, dtype=np.float32, nthreads=None):
    """Float twin of arpeggio_fx: G, h, T, q real-valued (scaled). Returns spins, best dE (E0 units), st, fields w."""
    G = np.ascontiguousarray(G, dtype=dtype)
    h = np.ascontiguousarray(h, dtype=dtype)
    T = np.ascontiguousarray(T, dtype=dtype)
    B, D = h.shape
    inst = np.ascontiguousarray(np.asarray(inst, dtype=np.int64) & MASK32, dtype=np.uint32)
    s = np.zeros((B, R, P * D), dtype=np.int8)
    de = np.zeros((B, R), dtype=dtype)
    st = np.zeros((B, R, 2), dtype=np.int64)
    w = np.zeros((B, R, D), dtype=dtype)
    fn = lib().orc_arpeggio_float_f32 if dtype == np.float32 else lib().orc_arpeggio_float_f64
    rc = fn(_p(G), _p(h), B, D, P, R, _p(T), len(T), float(q), _p(inst), key & 0xFFFFFFFFFFFFFFFF,
            0 if start == 'rand' else 1, _p(s), _p(de), _p(st), _p(w), _threads(nthreads))
    assert rc == 0
    return s, de, st, w


def chord_float(J, h, h0, T, c, q, R, seed, P=1, inst0=0, dtype=np.float32, nthreads=None):
    """Float twin of chord_fx: returns best spins (B, R, N), best E0 (B, R), final fields u (B, R, N), flips (B, R)."""
    J = np.ascontiguousarray(J, dtype=dtype)
    h = np.ascontiguousarray(h, dtype=dtype)
    h0 = np.ascontiguousarray(h0, dtype=dtype)
    T = np.ascontiguousarray(T, dtype=dtype)
    c = np.ascontiguousarray(c, dtype=dtype)
    B, N = h.shape
    s = np.zeros((B, R, N), dtype=np.int8)
    e = np.zeros((B, R), dtype=dtype)
    u = np.zeros((B, R, N), dtype=dtype)
    fl = np.zeros((B, R), dtype=np.int64)
    fn = lib().orc_chord_float_f32 if dtype == np.float32 else lib().orc_chord_float_f64
    rc = fn(_p(J), _p(h), _p(h0), B, N, _p(T), _p(c), len(T), float(q), P, R, seed & 0xFFFFFFFFFFFFFFFF,
            int(inst0), _p(s), _p(e), _p(u), _p(fl), _threads(nthreads))
    assert rc == 0
    return s, e, u, fl


# ---------------------------------------------------------------- schedules and fixed point
def geom(t0, t1, S):
    """mimo.geom: computed in float64, returned as float32 (mimo.py:199-201)."""
    k = np.arange(S, dtype=np.float64)
    return (t0 * (t1 / t0) ** (k / max(1, S - 1))).astype(np.float32)


def to_fixed(J, h, h0, S, t0, t1, q, kappa=0.0, ramp=True, Q=QSCALE):
    """chordfx.to_fixed on float32 arrays; rounding is half-to-even (numpy.rint, as torch.round)."""
    J = np.asarray(J, dtype=np.float32)
    Js = (J + np.swapaxes(J, 1, 2)) / np.float32(2)
    Ji = np.rint(Js * np.float32(Q)).astype(np.int16)
    hi = np.rint(np.asarray(h, dtype=np.float32) * np.float32(Q)).astype(np.int32)
    h0i = np.rint(np.asarray(h0, dtype=np.float32) * np.float32(Q)).astype(np.int32)
    T = geom(t0, t1, S).astype(np.float64)
    rho = np.ones(S)
    if ramp:
        t = np.arange(S, dtype=np.float64)
        rho = np.where(t >= 0.7 * S, (S - 1 - t) / max(1.0, S - 1 - 0.7 * S), 1.0)
    c = np.zeros(S)
    c[1:] = kappa * T[:-1] * rho[1:]
    return dict(J=Ji, h=hi, h0=h0i, T=np.rint(T * Q).astype(np.int32), c=np.rint(c * Q).astype(np.int32),
                q=int(round(q * Q)), S=S, Q=Q)


def compress(J, h, h0, K):
    """chord_ref.compress: J (B, N, N) int16 dDI replication -> G (B, D, D), h (B, D) int64 (asserts structure)."""
    Ji = np.asarray(J).astype(np.int64)
    B, N = np.asarray(h).shape
    D = N // (2 * K)
    G = -Ji[:, :D, D:2 * D]
    E = -np.tile(G, (1, 2 * K, 2 * K))
    E[:, np.arange(N), np.arange(N)] = 0
    assert (E == Ji).all(), 'coupling matrix is not the dDI replication'
    hh = np.asarray(h).astype(np.int64)
    assert (hh.reshape(B, 2 * K, D) == hh[:, None, :D]).all()
    assert (np.asarray(h) == np.asarray(h0)).all()
    return np.ascontiguousarray(G), np.ascontiguousarray(hh[:, :D])


def expand(G, h, K):
    B, D, _ = G.shape
    N = 2 * K * D
    J = -np.tile(np.asarray(G, dtype=np.int64), (1, 2 * K, 2 * K))
    J[:, np.arange(N), np.arange(N)] = 0
    return J.astype(np.int16), np.tile(np.asarray(h), (1, 2 * K)).astype(np.int32)


def arpeggio_schedule(S, T0, T1, q, K):
    """run_a513.arpeggio_schedule: one integer T per pass (S sweeps x 2K passes), q in Q units."""
    Tsw = [int(round(v * QSCALE)) for v in geom(T0, T1, S).astype(np.float64).tolist()]
    return [t for t in Tsw for _ in range(2 * K)], int(round(q * QSCALE))


# ---------------------------------------------------------------- MIMO front end and decisions (float64)
def quantize_levels(z, M):
    lv = np.asarray(PAM[M]['levels'], dtype=np.float64)
    return lv[np.abs(np.asarray(z)[..., None] - lv).argmin(-1)]


def mmse(H, y, N0, M):
    """(z, Ainv, el2): z = (H^T H + (N0/2)/E[l^2] I)^-1 H^T y (mimo.py:98-109)."""
    H = np.asarray(H, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    el2 = float(np.mean(np.asarray(PAM[M]['levels'], dtype=np.float64) ** 2))
    G = np.swapaxes(H, 1, 2) @ H
    D = G.shape[-1]
    A = G + (N0 / 2) / el2 * np.eye(D)
    Ainv = np.linalg.inv(A)
    z = np.einsum('bij,bj->bi', Ainv @ np.swapaxes(H, 1, 2), y)
    return z, Ainv, el2


def mmse_hard(H, y, N0, M, z=None, Ainv=None):
    """x_g = nearest level of z / mu, mu = diag(A^-1 G) (mimo.py:91-95)."""
    if z is None:
        z, Ainv, _ = mmse(H, y, N0, M)
    H = np.asarray(H, dtype=np.float64)
    G = np.swapaxes(H, 1, 2) @ H
    mu = np.diagonal(Ainv @ G, axis1=-2, axis2=-1)
    return quantize_levels(z / mu, M)


def ising_delta_at(H, y, xg, K=1):
    """dDI around xg: h = 2 T^T H^T ybar, J = -offdiag(T^T H^T H T), scaled into [-1, 1] (mimo.py:506-523).
    Returns dict(h, h0, J (B, N, N), scale, G (B, D, D) scaled Gram incl. diagonal, hD (B, D) scaled)."""
    H = np.asarray(H, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xg = np.asarray(xg, dtype=np.float64)
    B, _, D = H.shape
    yb = y - np.einsum('bij,bj->bi', H, xg)
    hD = 2 * np.einsum('bij,bi->bj', H, yb)
    G = np.swapaxes(H, 1, 2) @ H
    scale = np.maximum(np.abs(hD).max(-1), np.abs(G).reshape(B, -1).max(-1) if K >= 1 else 0)
    scale = np.maximum(scale, 1e-12)
    P = 2 * K
    N = P * D
    J = -np.tile(G, (1, P, P))
    J[:, np.arange(N), np.arange(N)] = 0
    h = np.tile(hD, (1, P))
    return dict(h=h / scale[:, None], h0=h / scale[:, None], J=J / scale[:, None, None], scale=scale,
                G=G / scale[:, None, None], hD=hD / scale[:, None], xq=xg, K=K)


def ising_ri(H, y, z, M, r=0.15):
    """RI-MIMO mapping (mimo.py:132-149): spins [MSB; LSB] for 16-QAM, h_reg = h0 + 2 r s_P."""
    H = np.asarray(H, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    B, _, D = H.shape
    T = np.eye(D) if M == 4 else np.concatenate([2 * np.eye(D), np.eye(D)], 1)
    HT = H @ T
    h = 2 * np.einsum('bij,bi->bj', HT, y)
    Q = np.swapaxes(HT, 1, 2) @ HT
    J = -(Q - np.einsum('bii->bi', Q)[:, :, None] * np.eye(Q.shape[-1]))
    scale = np.maximum(np.maximum(np.abs(h).max(-1), np.abs(J).reshape(B, -1).max(-1)), 1e-12)
    h0 = h / scale[:, None]
    J = J / scale[:, None, None]
    lq = quantize_levels(z, M)
    sP = lq if M == 4 else np.concatenate([np.where(lq > 0, 1.0, -1.0), np.where((lq == 3) | (lq == -1), 1.0, -1.0)], -1)
    return dict(h=h0 + 2 * r * sP, h0=h0, J=J, sP=sP, scale=scale)


def spins_to_levels_delta(S, xq, K, M):
    """levels = clip(x_q + sum over the 2K copies, constellation range) (mimo.py:185-190). S (B, R, 2K D)."""
    S = np.asarray(S)
    D = np.asarray(xq).shape[-1]
    d = S.reshape(*S.shape[:-1], 2 * K, D).sum(-2)
    lv = PAM[M]['levels']
    return np.clip(np.asarray(xq)[:, None, :] + d, lv[0], lv[-1])


def spins_to_levels(S, M, D):
    S = np.asarray(S)
    return S if M == 4 else 2 * S[..., :D] + S[..., D:]


def dist(H, y, L):
    """||y - H l||^2 for candidates L (B, K, D) (mimo.py:416-418)."""
    H = np.asarray(H, dtype=np.float64)
    return ((np.asarray(y, dtype=np.float64)[:, None, :] - np.einsum('bij,bkj->bki', H, np.asarray(L, np.float64))) ** 2).sum(-1)


def select(H, y, cand, mmse_levels):
    """First argmin of ||y - Hx||^2 over candidates then the MMSE estimate (mimo.py:421-425)."""
    allc = np.concatenate([np.asarray(cand, np.float64), np.asarray(mmse_levels, np.float64)[:, None, :]], 1)
    d = dist(H, y, allc)
    return allc[np.arange(allc.shape[0]), d.argmin(1)], allc, d


def levels_to_bits(l, M):
    l = np.asarray(l)
    if M == 4:
        return (l > 0).astype(np.int8)
    return np.stack([(l > 0).astype(np.int8), (np.abs(l) == 1).astype(np.int8)], -1).reshape(*l.shape[:-1], -1)


def biterr_per_use(l_true, l_hat, M):
    return (levels_to_bits(l_true, M) != levels_to_bits(l_hat, M)).sum(-1).astype(np.int32)


def energy0(s, h0, J):
    s = np.asarray(s, np.float64)
    return -(s * np.asarray(h0)[:, None, :]).sum(-1) - (s * np.einsum('bij,brj->bri', np.asarray(J, np.float64), s)).sum(-1)


# ---------------------------------------------------------------- statistics (a56an.py:74-106)
def _ratio(num, den):
    if den > 0:
        return num / den
    return float('inf') if num > 0 else 1.0


def boot_ratio(a, b, B=4000, seed=12345):
    """Paired cluster bootstrap of sum(a) / sum(b) over uses (a56an.py:74-95). Only uses with an error in either
    vector can change a sum, so each replicate draws how many of them it holds, Binomial(U, n_err / U), and then that
    many uniform picks among them. The generator is consumed in that order, replicate by replicate, which is what
    makes the result reproducible against the reference's published outputs. Returns (ratio, lo, hi): the 2.5 % and
    97.5 % order statistics of the replicates."""
    ea, eb = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    U = ea.shape[0]
    if U == 0:
        return None, None, None
    err = np.flatnonzero((ea > 0) | (eb > 0))
    whole = _ratio(ea.sum(), eb.sum())
    if err.size == 0:
        return 1.0, 1.0, 1.0
    gen = np.random.default_rng(seed)
    ca, cb = ea[err], eb[err]
    reps = []
    for _ in range(B):
        k = gen.binomial(U, err.size / U)
        cnt = np.bincount(gen.integers(0, err.size, k), minlength=err.size)
        reps.append(_ratio(float((cnt * ca).sum()), float((cnt * cb).sum())))
    reps.sort()
    return float(whole), float(reps[int(0.025 * (B - 1))]), float(reps[int(math.ceil(0.975 * (B - 1)))])


def mcnemar(a, b):
    """Exact two-sided McNemar test on per-use error indicators (a56an.py:98-106): (p, only-a, only-b)."""
    fa, fb = np.asarray(a) > 0, np.asarray(b) > 0
    only_a, only_b = int((fa & ~fb).sum()), int((fb & ~fa).sum())
    n = only_a + only_b
    if n == 0:
        return 1.0, only_a, only_b
    tail = sum(math.comb(n, i) for i in range(min(only_a, only_b) + 1))
    return min(1.0, 2 * tail / 2 ** n), only_a, only_b
