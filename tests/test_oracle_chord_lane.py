"""CPU pins of the lane-shared Chord oracle (chord_ref.chord_ref, arpeggio/sim/x3model/chord_ref.py:73-123, with the
production xorshift32 lane streams or Philox): bit-equal to the reference's own cycle-accurate engine model
(arpeggio/sim/x3model/engine_model.cpp, compiled into oracle/_ref from the reference's source in the build
container), and equal to the first-accept oracle's parallel pass ('par') when c = 0 under Philox."""
import numpy as np
import pytest


def problem(rs, B, D):
    A = rs.normal(size=(B, D + 4, D))
    G = np.einsum('bri,brj->bij', A, A)
    hv = rs.normal(size=(B, D)) * 3
    m = np.maximum(np.abs(G).reshape(B, -1).max(-1), np.abs(hv).max(-1))
    return (np.rint(G / m[:, None, None] * 4096).astype(np.int32),
            np.rint(hv / m[:, None] * 4096).astype(np.int32))


CASES = [(16, 2, 64, 'xs', 0.3), (32, 4, 48, 'xs', 0.0), (16, 2, 40, 'philox', 0.3), (24, 6, 36, 'xs', 0.3),
         (64, 2, 32, 'philox', 0.0), (64, 2, 256, 'xs', 0.0)]


@pytest.mark.parametrize('D,P,S,rng,kappa', CASES)
def test_chord_lane_equals_reference_engine(orc, D, P, S, rng, kappa):
    if orc.ref_engine() is None:
        pytest.skip('oracle/_ref/libape.so is not built (make -C oracle ref, where the reference is mounted)')
    rs = np.random.default_rng(D * 100 + S)
    B, R = 3, 5
    G, h = problem(rs, B, D)
    F = orc.to_fixed(np.zeros((1, 2, 2)), np.zeros((1, 2)), np.zeros((1, 2)), S, 0.3, 0.3, 0.4, kappa)
    inst = [7, 100, 2 ** 32 - 3]
    seed = int(rs.integers(0, 2 ** 63))
    s1, de1, fl1 = orc.chord_lane(G, h, P, R, F['T'], F['c'], F['q'], inst, seed, rng)
    s2, de2, cyc, st = orc.ref_engine_solve(G, h, P, R, F['T'], F['c'], F['q'], inst, seed,
                                            philox=1 if rng == 'philox' else 0)
    assert (s1 == s2).all() and (de1 == de2).all()
    assert (fl1.sum(1) == st[:, 1]).all()          # flips = the engine's F-ops


def test_chord_lane_philox_is_the_parallel_pass(orc):
    """With Philox streams and c = 0, the lane-shared Chord step is the first-accept oracle's 'par' pass."""
    rs = np.random.default_rng(5)
    for D, P in ((20, 2), (33, 4), (64, 6)):
        G, h = problem(rs, 4, D)
        T = rs.integers(100, 3000, 24).tolist()
        inst = [3, 4, 5, 6]
        s1, de1, fl1 = orc.chord_lane(G, h, P, 6, T, np.zeros(24, np.int32), 250, inst, 0xBEEF, 'philox')
        s2, de2, st2 = orc.arpeggio_fx(G, h, P, 6, T, 250, inst, 0xBEEF, 'rand', mode='par')
        assert (s1 == s2).all() and (de1 == de2).all() and (fl1 == st2[..., 0]).all()
