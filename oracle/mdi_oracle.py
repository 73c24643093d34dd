"""CPU oracle of the multi-stage delta-Ising (MDI) wrapper, the soft outputs and the production xorshift streams.

TEST INFRASTRUCTURE ONLY: imported by tests/ and the benchmark's CPU-baseline leg as the checker; the product package
(arpeggio_amd) never imports it. numpy float64 restatements of the reference's text (paths relative to the reference):
  mmse_sic                 arpeggio/mimo/mimo.py:466-503 (ordered MMSE-SIC, per real stream or per user)
  mdi_radius               arpeggio/mimo/mimo.py:526-543 (search radius, mean / max / min eigenvalue)
  ddi_fixed                mimo.ising_delta_at (mimo.py:506-523) -> chordfx.to_fixed (chordfx.py:60-76) ->
                           chord_ref.compress (arpeggio/sim/x3model/chord_ref.py:47-59), in the compressed form
  mdi                      arpeggio/mimo/mimo.py:546-602 and arpeggio/sim/local_code_a513/run_a513.py:43-80
  ArpeggioMDISolver        arpeggio/sim/local_code_a513/run_a513.py:90-130 (per-call Philox keys, fixed point)
  ChordMDISolver           arpeggio/sim/x3_prod.py:80-112 (the production Chord engine in MDI, via chord_lane)
  mmse_soft_llr, list_llr  arpeggio/mimo/mimo.py:111-129, 434-462
  fmix32, xs32, lane_seeds arpeggio/sim/x3model/chord_ref.py:17-45
"""
import math

import numpy as np

import oracle as orc

M64 = 0xFFFFFFFFFFFFFFFF
MASK32 = 0xFFFFFFFF
KEY_MUL = 0x9E3779B97F4A7C15     # seed base: draw seed times this (run_a513.py:99)
KEY_STEP = 0xD1B54A32D192ED03    # per solver call (run_a513.py:28)
KEY_TAG = 0xA513                 # first-accept stream tag (run_a513.py:29)
# MDI configurations of Singh et al. 2024 Table I as the reference runs them (arpeggio/mimo/mdi_solvers.py:16-20)
MDI_CFG = {(16, 16): dict(Rmax=3, S=8, guesses=('mmse', 'sic')), (32, 16): dict(Rmax=3, S=8, guesses=('mmse', 'sic')),
           (64, 16): dict(Rmax=3, S=8, guesses=('mmse', 'sic')), (32, 4): dict(Rmax=1, S=1, guesses=('mmse',))}
# the Chord schedule whose to_fixed call also rounds the couplings of the Arpeggio sub-problems (mdi_solvers.py:15)
MDICHORD = dict(Sk=128, t0=0.3, t1=0.3, q=0.4, kappa=0.0)


def _el2(M):
    return float(np.mean(np.asarray(orc.PAM[M]['levels'], dtype=np.float64) ** 2))


def gram(H):
    H = np.asarray(H, dtype=np.float64)
    return np.swapaxes(H, 1, 2) @ H


def mmse_guess(H, y, N0, M):
    """The MMSE guess of the wrapper: nearest level of z / mu with mu = diag(A^-1 G) (mimo._mu by solve)."""
    z, Ainv, _ = orc.mmse(H, y, N0, M)
    mu = np.diagonal(Ainv @ gram(H), axis1=-2, axis2=-1)
    return orc.quantize_levels(z / mu, M)


def _pick(v, tie):
    """argmin over the last axis; tie 'first' / 'last': among values within a relative 1e-9 of the minimum, the
    lowest / highest index (None: the exact first minimum)."""
    if tie is None:
        return v.argmin(1)
    near = v <= v.min(1, keepdims=True) * (1 + 1e-9)
    if tie == 'first':
        return near.argmax(1)
    return v.shape[1] - 1 - near[:, ::-1].argmax(1)


def mmse_sic(H, y, N0, M, per='real', tie=None):
    """Ordered MMSE-SIC: each stage detects the stream (or user) with the smallest MMSE error variance among those left,
    quantises its bias-corrected estimate and cancels it. Returns levels (B, D).
    In the real lift of a complex channel the two real streams of a user start with equal error variances (the Gram
    is [[C, -E], [E, C]]), so the order between them is decided by rounding; tie selects a rule for such near-ties."""
    H = np.asarray(H, dtype=np.float64)
    y = np.array(y, dtype=np.float64)
    B, _, D = H.shape
    Nt = D // 2
    s2 = (N0 / 2) / _el2(M)
    act = np.ones((B, D), dtype=bool)
    x = np.zeros((B, D))
    ar = np.arange(B)
    for _ in range(D if per == 'real' else Nt):
        Hm = H * act[:, None, :]
        G = np.swapaxes(Hm, 1, 2) @ Hm
        P = np.linalg.inv(G + s2 * np.eye(D))
        mse = np.diagonal(P, axis1=-2, axis2=-1).copy()
        if per == 'complex':
            mc = mse[:, :Nt] + mse[:, Nt:]
            mc = np.where(act[:, :Nt], mc, np.inf)
            us = _pick(mc, tie)
            js = [us, us + Nt]
        else:
            js = [_pick(np.where(act, mse, np.inf), tie)]
        z = np.einsum('bij,bj->bi', P @ np.swapaxes(Hm, 1, 2), y)
        mu = np.diagonal(P @ G, axis1=-2, axis2=-1)
        for j in js:
            xj = orc.quantize_levels(z[ar, j] / mu[ar, j], M)
            x[ar, j] = xj
            y = y - H[ar, :, j] * xj[:, None]
            act[ar, j] = False
    return x


def mdi_radius(H, y, xg, M, Rmax, heuristic='mean'):
    """Gamma = 2 ||H^T (y - H x_g)|| / lambda; R = omega if 2 omega <= Gamma else floor(Gamma / 2), clipped to
    [0, Rmax]; 2 omega = largest level difference (2 QPSK, 6 16-QAM). Returns (R int64, Gamma)."""
    H = np.asarray(H, dtype=np.float64)
    lv = orc.PAM[M]['levels']
    omega = (lv[-1] - lv[0]) // 2
    yb = np.asarray(y, np.float64) - np.einsum('bij,bj->bi', H, np.asarray(xg, np.float64))
    g = np.linalg.norm(np.einsum('bij,bi->bj', H, yb), axis=-1)
    G = gram(H)
    if heuristic == 'mean':
        lam = np.diagonal(G, axis1=-2, axis2=-1).mean(-1)
    else:
        ev = np.linalg.eigvalsh(G)
        lam = ev[:, -1] if heuristic == 'max' else np.maximum(ev[:, 0], 1e-12)
    Gam = 2 * g / lam
    R = np.where(2 * omega <= Gam, float(omega), np.floor(Gam / 2))
    return np.clip(R, 0, Rmax).astype(np.int64), Gam


def ddi_fixed(H, y, xg, K, Q=orc.QSCALE):
    """Fixed-point dDI sub-problem around levels xg in the compressed form the first-accept kernel takes:
    G = rne(G_ij / scale * Q), h = rne(2 (H^T (y - H xg))_j / scale * Q) with scale = max(max|h|, max|G|) (for
    2K >= 2 copies the expanded J holds every entry of G, its diagonal included). The rounding goes through float32
    as the reference's (to_fixed on float32 tensors)."""
    H = np.asarray(H, dtype=np.float64)
    B, _, D = H.shape
    yb = np.asarray(y, np.float64) - np.einsum('bij,bj->bi', H, np.asarray(xg, np.float64))
    hD = 2 * np.einsum('bij,bi->bj', H, yb)
    G = gram(H)
    scale = np.maximum(np.maximum(np.abs(hD).max(-1), np.abs(G).reshape(B, -1).max(-1)), 1e-12)
    Gs = (G / scale[:, None, None]).astype(np.float32)
    Gs = (Gs + np.swapaxes(Gs, 1, 2)) / np.float32(2)
    Gi = np.rint(Gs * np.float32(Q)).astype(np.int64)
    hi = np.rint((hD / scale[:, None]).astype(np.float32) * np.float32(Q)).astype(np.int64)
    return Gi, hi, scale


def mdi(H, y, N0, M, solver, Ns, Rmax, S, guesses=('mmse', 'sic'), heuristic='mean', sic_per='real',
        na_round='ceil', start=None):
    """The MDI-MIMO wrapper. solver(H_sub, y_sub, xg_sub, K, na, idx, stage, guess) -> spins (n, na, 2K D).
    Per stage and guess the radius R_g; sub-problems (g, K = 1..R_g) grouped by the total radius t = sum_g R_g, each
    group's frames solved with ceil(Ns / t) anneals; each guess becomes its best candidate by ||y - Hx||^2 (strictly
    better than the guess); after S stages (or when every radius is 0) the better guess (first on ties).
    start: optional dict guess -> starting levels (B, D) in place of computing them.
    Returns (levels (B, D), anneals per frame)."""
    H = np.asarray(H, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    B, _, D = H.shape
    cur = {}
    for g in guesses:
        if start is not None and g in start:
            cur[g] = np.asarray(start[g], dtype=np.float64)
        else:
            cur[g] = mmse_guess(H, y, N0, M) if g == 'mmse' else mmse_sic(H, y, N0, M, per=sic_per)
    ann = np.zeros(B)
    for st in range(S):
        R = {g: mdi_radius(H, y, cur[g], M, Rmax, heuristic)[0] for g in guesses}
        tot = sum(R[g] for g in guesses)
        if int(tot.max()) == 0:
            break
        nxt = {}
        for g in guesses:
            bl = cur[g].copy()
            bd = orc.dist(H, y, cur[g][:, None, :])[:, 0]
            for K in range(1, Rmax + 1):
                need = R[g] >= K
                if not need.any():
                    continue
                for t in np.unique(tot[need]).tolist():
                    idx = np.flatnonzero(need & (tot == t))
                    na = math.ceil(Ns / t) if na_round == 'ceil' else max(1, Ns // t)
                    Ssp = solver(H[idx], y[idx], cur[g][idx], K, na, idx, st, g)
                    Lc = orc.spins_to_levels_delta(Ssp, cur[g][idx], K, M)
                    dd = orc.dist(H[idx], y[idx], Lc)
                    im = dd.argmin(1)
                    dm = dd[np.arange(len(idx)), im]
                    better = dm < bd[idx]
                    bd[idx] = np.where(better, dm, bd[idx])
                    bl[idx] = np.where(better[:, None], Lc[np.arange(len(idx)), im], bl[idx])
                    ann[idx] += na
            nxt[g] = bl
        cur = nxt
    allc = np.stack([cur[g] for g in guesses], 1)
    d = orc.dist(H, y, allc)
    return allc[np.arange(B), d.argmin(1)], ann


def mdi_seed_base(draw_seed):
    return (draw_seed * KEY_MUL) & M64


class ArpeggioMDISolver:
    """The fixed-point first-accept solver of the reference's MDI runs: call ci (in call order) anneals its frames with
    key (seed_base + ci KEY_STEP + KEY_TAG) mod 2^64, instance word = frame index in the batch, schedule S sweeps x 2K
    passes at geometric T (T0 -> T1), q, start 'rand' / 'bal'. Keeps per-K counters [flips, tests, anneals]."""

    def __init__(self, draw_seed, S, T0, T1, q, start, nthreads=None):
        self.base = mdi_seed_base(draw_seed)
        self.S, self.T0, self.T1, self.q, self.start = S, T0, T1, q, start
        self.calls, self.nthreads = 0, nthreads
        self.cnt = {}

    def key(self, ci):
        return (self.base + ci * KEY_STEP + KEY_TAG) & M64

    def __call__(self, H, y, xg, K, na, idx, stage, guess):
        ci = self.calls
        self.calls += 1
        G, h, _ = ddi_fixed(H, y, xg, K)
        T, qi = orc.arpeggio_schedule(self.S, self.T0, self.T1, self.q, K)
        s, de, st = orc.arpeggio_fx(G, h, 2 * K, na, T, qi, idx, self.key(ci), self.start, nthreads=self.nthreads)
        v = self.cnt.setdefault(K, [0, 0, 0])
        v[0] += int(st[..., 0].sum())
        v[1] += int(st[..., 1].sum())
        v[2] += int(na * len(idx))
        return s.astype(np.float64)


class ChordMDISolver:
    """The production Chord engine as the MDI solver (x3_prod.py:80-112; run_a513.py:97-122 'prod'): call ci anneals
    its frames with the lane-shared Chord (chord_lane, xorshift32 lane streams) under key (seed_base + ci KEY_STEP)
    mod 2^64 (no tag), instance word = frame index, 2K copies, the MDICHORD schedule of Sk 2K steps (T, c, q from
    to_fixed on an empty problem, as the reference tabulates them). Keeps per-K counters [steps, flips, anneals], the
    order the reference logs the engine's D-ops and F-ops."""

    def __init__(self, draw_seed, nthreads=None, cfg=MDICHORD):
        self.base = mdi_seed_base(draw_seed)
        self.cfg = cfg
        self.calls, self.nthreads = 0, nthreads
        self.cnt = {}

    def key(self, ci):
        return (self.base + ci * KEY_STEP) & M64

    def __call__(self, H, y, xg, K, na, idx, stage, guess):
        ci = self.calls
        self.calls += 1
        G, h, _ = ddi_fixed(H, y, xg, K)
        c = self.cfg
        F = orc.to_fixed(np.zeros((1, 2, 2)), np.zeros((1, 2)), np.zeros((1, 2)), c['Sk'] * 2 * K, c['t0'], c['t1'],
                         c['q'], c['kappa'])
        s, de, fl = orc.chord_lane(G, h, 2 * K, na, F['T'], F['c'], F['q'], idx, self.key(ci), 'xs',
                                   nthreads=self.nthreads)
        v = self.cnt.setdefault(K, [0, 0, 0])
        v[0] += int(len(F['T']) * na * len(idx))
        v[1] += int(fl.sum())
        v[2] += int(na * len(idx))
        return s.astype(np.float64)


# ---------------------------------------------------------------- soft output (max-log LLRs)
def level_bits(M):
    lv = orc.PAM[M]['levels']
    return orc.levels_to_bits(np.asarray(lv)[:, None], M)      # (L, bits)


def mmse_soft_llr(H, y, N0, M):
    """Max-log LLRs per bit from the unbiased MMSE output z_k = mu_k l_k + e_k, Var(e_k) = E[l^2] mu_k (1 - mu_k):
    LLR(b) = (min over levels with b = 0 - min over levels with b = 1) of (z - mu l)^2 / var, halved."""
    z, Ainv, el2 = orc.mmse(H, y, N0, M)
    mu = np.diagonal(Ainv @ gram(H), axis1=-2, axis2=-1)
    var = el2 * mu * (1 - mu)
    lv = np.asarray(orc.PAM[M]['levels'], dtype=np.float64)
    d2 = (z[..., None] - mu[..., None] * lv) ** 2 / np.maximum(var[..., None], 1e-12)
    bits = level_bits(M).astype(bool)
    out = []
    for k in range(bits.shape[1]):
        m1 = np.where(bits[:, k], d2, np.inf).min(-1)
        m0 = np.where(~bits[:, k], d2, np.inf).min(-1)
        out.append((m0 - m1) / 2)
    return np.stack(out, -1).reshape(z.shape[0], -1)


def bit_neighbours(best, M):
    """The single-Gray-bit neighbours of each candidate: (B, D * bits, D), neighbour (dim, k) flips bit k of dim."""
    best = np.asarray(best)
    B, D = best.shape
    lv = np.asarray(orc.PAM[M]['levels'])
    gray = level_bits(M)
    nb = gray.shape[1]
    out = np.repeat(best[:, None, :], D * nb, 1).astype(np.float64)
    idx = np.abs(best[..., None] - lv).argmin(-1)               # (B, D) level index
    for dim in range(D):
        for k in range(nb):
            want = gray[idx[:, dim]].copy()
            want[:, k] ^= 1
            new = (gray[None, :, :] == want[:, None, :]).all(-1).argmax(1)
            out[:, dim * nb + k, dim] = lv[new]
    return out


def list_llr(H, y, N0, M, allc, d, best, clip=50.0):
    """Max-log LLRs from a candidate list (allc (B, K, D) with distances d) plus the single-bit neighbours of best:
    LLR(b) = (min d over candidates with b = 0 - min d with b = 1) / N0, clipped."""
    nbr = bit_neighbours(best, M)
    allc2 = np.concatenate([np.asarray(allc, np.float64), nbr], 1)
    d2 = np.concatenate([np.asarray(d, np.float64), orc.dist(H, y, nbr)], 1)
    bits = orc.levels_to_bits(allc2, M).astype(bool)
    m1 = np.where(bits, d2[..., None], np.inf).min(1)
    m0 = np.where(~bits, d2[..., None], np.inf).min(1)
    return np.clip((m0 - m1) / N0, -clip, clip)


# ---------------------------------------------------------------- production xorshift streams
PHI = 0x9E3779B9


def fmix32(h):
    h = np.asarray(h, dtype=np.uint64) & MASK32
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & MASK32
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & MASK32
    return h ^ (h >> 16)


def xs32(x):
    x = np.asarray(x, dtype=np.uint64) & MASK32
    x ^= (x << 13) & MASK32
    x ^= x >> 17
    return x ^ ((x << 5) & MASK32)


def lane_seeds(seed, inst, r, D):
    """Lane j of anneal r of instance inst starts at fmix32(fmix32(fmix32(fmix32(lo ^ inst) ^ hi) ^ r) + (j + 1) PHI),
    with 0 replaced by PHI (a xorshift state must be nonzero)."""
    lo, hi = seed & MASK32, (seed >> 32) & MASK32
    b = fmix32(fmix32(fmix32(np.uint64(lo) ^ np.uint64(inst)) ^ np.uint64(hi)) ^ np.uint64(r))
    j = np.arange(1, D + 1, dtype=np.uint64)
    st = fmix32((b + j * PHI) & MASK32)
    return np.where(st == 0, PHI, st).astype(np.uint64)
