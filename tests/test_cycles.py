"""Modeled-cycle accounting (arpeggio_amd.cycles) against the reference's A5.13 analysis (tests/golden/a513_cycles.json,
from data/results/a513/analysis.json): the first-accept cost per anneal from its counters, x_F from both solvers' costs,
Chord's L_K + F_K table; and the counter meaning pinned to the reference's engine model (D-ops = steps, F-ops = flips
of the lane-shared Chord)."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT

GOLD = json.load(open(os.path.join(ROOT, 'tests', 'golden', 'a513_cycles.json')))
SK = 128


def chord_counts(cost, n=10 ** 6):
    """Chord counters [steps, flips, anneals] per K with the given cost per anneal (L_K = Sk 2K steps)."""
    return {K: [SK * 2 * int(K) * n, (c - SK * 2 * int(K)) * n, n] for K, c in cost.items()}


@pytest.mark.parametrize('pt', ['P1', 'P2', 'P4'])
def test_cost_and_x_f_match_reference_analysis(pt):
    from arpeggio_amd import cycles
    g = GOLD[pt]
    kf = cycles.cost_per_anneal(g['counters_fx'])
    for K, v in g['cost_per_anneal']['fx'].items():
        assert kf[K] == pytest.approx(v, rel=1e-12)
    x, xk, kf2, kp = cycles.x_f(g['counters_fx'], chord_counts(g['cost_per_anneal']['prod']))
    assert x == pytest.approx(g['x_F'], rel=1e-9)
    for K, v in g['x_F_per_K'].items():
        assert xk[K] == pytest.approx(v, rel=1e-9)
    assert cycles.within_budget(x)            # the reference's claim: Arpeggio stays within Chord's budget
    tab = cycles.chord_cycle_table(chord_counts(g['cost_per_anneal']['prod']), 3, SK)
    np.testing.assert_allclose(tab, [g['cost_per_anneal']['prod'][K] for K in ('1', '2', '3')], rtol=1e-12)


def test_cycle_table_and_per_use_closed_forms():
    from arpeggio_amd import cycles
    cnt = {1: [0, 300, 10], 3: [0, 90, 3]}                 # K = 2 has no anneals: L_2 only
    tab = cycles.chord_cycle_table(cnt, 3, 128)
    np.testing.assert_allclose(tab, [256 + 30, 512, 768 + 30])
    ann = np.array([[8, 0, 0], [4, 4, 0], [0, 0, 2]])
    np.testing.assert_allclose(cycles.per_use_cycles(ann, tab), [8 * 286, 4 * 286 + 4 * 512, 2 * 798])
    assert cycles.x_f({}, cnt)[0] is None and not cycles.within_budget(None)
    acc = cycles.add_counts({}, {1: [1, 2, 3]})
    cycles.add_counts(acc, {'1': [1, 1, 1], 2: [5, 5, 5]})
    assert acc == {'1': [2, 3, 4], '2': [5, 5, 5]}


def test_engine_counters_are_steps_and_flips(orc):
    """The reference engine's stats columns 0 / 1 (D-ops, F-ops: what the production runs log as the Chord cost) are
    L x R and the lane-shared Chord's flip count."""
    if orc.ref_engine() is None:
        pytest.skip('oracle/_ref/libape.so is not built (make -C oracle ref, where the reference is mounted)')
    rs = np.random.default_rng(3)
    D, R, S = 32, 5, 256
    A = rs.normal(size=(4, D + 4, D))
    G = np.einsum('bri,brj->bij', A, A)
    m = np.abs(G).reshape(4, -1).max(-1)
    Gi = np.rint(G / m[:, None, None] * 4096).astype(np.int32)
    hi = np.rint(rs.normal(size=(4, D)) * 900).astype(np.int32)
    F = orc.to_fixed(np.zeros((1, 2, 2)), np.zeros((1, 2)), np.zeros((1, 2)), S, 0.3, 0.3, 0.4, 0.0)
    _, _, _, st = orc.ref_engine_solve(Gi, hi, 2, R, F['T'], F['c'], F['q'], [0, 1, 2, 3], 99)
    _, _, fl = orc.chord_lane(Gi, hi, 2, R, F['T'], F['c'], F['q'], [0, 1, 2, 3], 99, 'xs')
    assert (st[:, 0] == S * R).all() and (st[:, 1] == fl.sum(1)).all()


def test_oracle_chord_mdi_counters(orc):
    """The oracle's production-Chord MDI solver logs [steps, flips, anneals] with steps = Sk 2K per anneal."""
    import mdi_oracle as mdo
    from arpeggio_amd import frames
    b = frames.draw(102, 22.0, 0, 12, 16, 16, dev='cpu')
    sol = mdo.ChordMDISolver(frames.draw_seed(102, 22.0, 0), nthreads=4)
    lo, ann = mdo.mdi(b['H'].numpy(), b['y'].numpy(), b['N0'], 16, sol, 8, 3, 8)
    assert sol.cnt and ann.sum() == sum(v[2] for v in sol.cnt.values())
    for K, v in sol.cnt.items():
        assert v[0] == SK * 2 * K * v[2] and 0 <= v[1] <= v[0] * 2 * K * 32
