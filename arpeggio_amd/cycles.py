"""Modeled-cycle accounting of the annealing engines, as the reference prices Chord against Arpeggio.

The cycle model charges one cycle per decision step and one per flip that updates all fields from a coupling row
(README.md "The price of parallel flips"). Per anneal:
  Chord (the production engine, lane-shared):  L + F, L = Sk 2K steps, F = flips      (a56an.py:43-55)
  Arpeggio (first-accept):                     tests + F, tests = comparator passes   (analyze_a513.py:48-53)
Counters are kept per K (the MDI sub-problem's radius: 2K copies) as [a, b, anneals] with a + b the modeled cycles:
[flips, tests, anneals] for Arpeggio and [steps, flips, anneals] for Chord, the order the reference's runs log them
(run_a513.py:119-126). x_F is the mean over K of Arpeggio's cost per anneal over Chord's; "within default Chord's
modeled budget" means x_F <= 1."""
import math

import numpy as np


def add_counts(acc, cnt):
    """acc[K] += cnt[K] element-wise ([a, b, anneals] per K; keys may be int or str)."""
    for K, v in cnt.items():
        a = acc.setdefault(str(K), [0, 0, 0])
        for i in range(3):
            a[i] += v[i]
    return acc


def cost_per_anneal(cnt):
    """{K: (a + b) / anneals} over the K that have anneals."""
    return {str(K): (v[0] + v[1]) / v[2] for K, v in cnt.items() if v[2]}


def x_f(cnt_arpeggio, cnt_chord):
    """(x_F, {K: ratio}, Arpeggio cost per anneal, Chord cost per anneal): x_F = mean over the K both ran of
    Arpeggio's (tests + flips) per anneal over Chord's (L + F) per anneal (analyze_a513.py:48-53)."""
    kf, kp = cost_per_anneal(cnt_arpeggio), cost_per_anneal(cnt_chord)
    xk = {K: kf[K] / kp[K] for K in kf if K in kp}
    return (float(np.mean(list(xk.values()))) if xk else None), xk, kf, kp


def chord_cycle_table(cnt, Rmax, Sk=128):
    """Chord's cycles per anneal for K = 1..Rmax: L_K + F_K with L_K = Sk 2K and F_K the mean flips per anneal measured
    in cnt ([steps, flips, anneals] per K); a K without anneals costs L_K (a56an.py:43-55)."""
    out = []
    for K in range(1, Rmax + 1):
        v = cnt.get(str(K), cnt.get(K))
        F = v[1] / v[2] if v and v[2] else 0.0
        out.append(Sk * 2 * K + F)
    return np.array(out)


def per_use_cycles(anneals_by_K, table):
    """Modeled cycles per use: anneals_by_K (U, Rmax) anneals each use spent at each K, times the per-anneal cost
    table (Rmax,)."""
    return np.asarray(anneals_by_K, dtype=np.float64) @ np.asarray(table, dtype=np.float64)


def within_budget(x):
    """True when Arpeggio's mean cost per anneal is at most Chord's (x_F <= 1)."""
    return x is not None and not math.isnan(x) and x <= 1.0
