"""GPU parity of the production engine's lane-shared Chord (arp_anneal_chord_lane) against the C oracle of
chord_ref.chord_ref (arpeggio/sim/x3model/chord_ref.py:73-123), itself bit-equal to the reference's cycle-accurate
engine model on the CPU: spins, best energies and flip counts exact, for the xorshift32 lane streams and Philox, with
and without the Onsager term, over full and partial blocks of 4P steps, D up to 128 and P = 2, 4, 6."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print('[case]', *a, flush=True)


def problem(rs, B, D):
    A = rs.normal(size=(B, D + 4, D))
    G = np.einsum('bri,brj->bij', A, A)
    hv = rs.normal(size=(B, D)) * 3
    m = np.maximum(np.abs(G).reshape(B, -1).max(-1), np.abs(hv).max(-1))
    return (np.rint(G / m[:, None, None] * 4096).astype(np.int32),
            np.rint(hv / m[:, None] * 4096).astype(np.int32))


CASES = [(16, 2, 64, 'xs', 0.3), (32, 4, 48, 'xs', 0.0), (16, 2, 40, 'philox', 0.3), (24, 6, 36, 'xs', 0.3),
         (64, 2, 32, 'philox', 0.0), (64, 2, 256, 'xs', 0.3), (100, 2, 70, 'xs', 0.3), (128, 4, 96, 'philox', 0.3),
         (128, 2, 256, 'xs', 0.0), (37, 2, 30, 'xs', 0.3)]


@pytest.mark.parametrize('D,P,S,rng,kappa', CASES)
def test_chord_lane_vs_oracle(orc, D, P, S, rng, kappa):
    from arpeggio_amd.arpeggiofx import chord_ref_device
    import torch
    say('chord_lane', D, P, S, rng, kappa)
    rs = np.random.default_rng(D * 1000 + S)
    B, R = 6, 9
    G, h = problem(rs, B, D)
    F = orc.to_fixed(np.zeros((1, 2, 2)), np.zeros((1, 2)), np.zeros((1, 2)), S, 0.3, 0.2, 0.4, kappa)
    inst = [7, 100, 2 ** 32 - 3, 0, 5, 12345]
    seed = int(rs.integers(0, 2 ** 63))
    s, de, st = chord_ref_device(G, h, P, F['T'], F['c'], F['q'], R, seed, inst=inst, rng=rng)
    torch.cuda.synchronize()
    so, deo, flo = orc.chord_lane(G, h, P, R, F['T'], F['c'], F['q'], inst, seed, rng)
    s, de, st = s.cpu().numpy(), de.cpu().numpy(), st.cpu().numpy()
    say('spins equal', (s == so).all(-1).mean(), 'de equal', (de == deo).mean())
    assert (s == so).all() and (de == deo).all() and (st[..., 0] == flo).all()


def test_chord_lane_philox_no_onsager_is_parallel_pass(orc):
    """rng philox with c = 0 on the GPU equals arp_anneal_arpeggio's parallel pass (mode 1)."""
    from arpeggio_amd.arpeggiofx import chord_ref_device, solve
    import torch
    rs = np.random.default_rng(11)
    G, h = problem(rs, 5, 48)
    T = rs.integers(100, 3000, 40).tolist()
    s1, de1, st1 = chord_ref_device(G, h, 4, T, None, 250, 7, 0xBEEF, inst=[3, 4, 5, 6, 7], rng='philox')
    torch.cuda.synchronize()
    s2, de2, st2 = solve(G, h, 4, 7, T, 250, [3, 4, 5, 6, 7], 0xBEEF, 'rand', mode='par')
    assert (s1.cpu().numpy() == s2).all() and (de1.cpu().numpy() == de2).all()
    assert (st1.cpu().numpy() == st2).all()
