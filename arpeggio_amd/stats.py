"""Paired statistics for comparing two detectors on the same frames.

boot_ratio follows the paired cluster bootstrap of arpeggio/sim/a56an.py:74-95 (ratio of summed per-frame bit
errors, resampling frames jointly for both detectors), vectorised over replicates: only frames where either detector
errs can move a sum, so a replicate draws its number of such frames from Binomial(U, n_err / U) and then picks them
uniformly. The replicate order statistics at 2.5 % and 97.5 % bound the ratio."""
import math

import numpy as np


def boot_ratio(a, b, reps=2000, seed=12345, chunk=256):
    """(ratio, lo, hi) of sum(a) / sum(b) for per-frame error counts a, b of equal length."""
    ea, eb = np.asarray(a, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()
    if ea.shape != eb.shape:
        raise ValueError('boot_ratio needs paired vectors of one length')
    U = ea.size
    if U == 0:
        return None, None, None

    def ratio(num, den):
        with np.errstate(divide='ignore', invalid='ignore'):
            r = np.where(den > 0, num / np.where(den > 0, den, 1), np.where(num > 0, np.inf, 1.0))
        return r

    err = np.flatnonzero((ea > 0) | (eb > 0))
    whole = float(ratio(np.float64(ea.sum()), np.float64(eb.sum())))
    if err.size == 0:
        return 1.0, 1.0, 1.0
    gen = np.random.default_rng(seed)
    ca, cb = ea[err], eb[err]
    k = gen.binomial(U, err.size / U, size=reps)
    out = np.empty(reps)
    for s in range(0, reps, chunk):
        kk = k[s:s + chunk]
        kmax = int(kk.max())
        pick = gen.integers(0, err.size, size=(kk.size, kmax))
        live = np.arange(kmax)[None, :] < kk[:, None]
        out[s:s + chunk] = ratio((ca[pick] * live).sum(1), (cb[pick] * live).sum(1))
    out.sort()
    return whole, float(out[int(0.025 * (reps - 1))]), float(out[int(math.ceil(0.975 * (reps - 1)))])
