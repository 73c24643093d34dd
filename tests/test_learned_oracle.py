"""CPU checks of the learned-start oracles: the MLP's backward pass against central finite differences of its own
float64 loss, and the warm-start anneal's defining properties (no reference counterpart exists)."""
import numpy as np
import pytest


def small_params(rs, F, H, N):
    return dict(W1=rs.normal(size=(F, H)) / np.sqrt(F), b1=rs.normal(size=H) * 0.1,
                W2=rs.normal(size=(H, H)) / np.sqrt(H), b2=rs.normal(size=H) * 0.1,
                W3=rs.normal(size=(H, N)) / np.sqrt(H), b3=rs.normal(size=N) * 0.1)


def test_mlp_oracle_finite_differences():
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
    import mlp_oracle as mo
    rs = np.random.default_rng(0)
    F, H, N, M = 12, 7, 6, 5
    p = small_params(rs, F, H, N)
    X = rs.normal(size=(M, F))
    Y = np.where(rs.random((M, N)) < 0.5, -1, 1)
    loss, g = mo.loss_and_grads(p, X, Y, round_bf16=False)
    eps = 1e-6
    for k in p:
        flat = p[k].reshape(-1)
        for idx in rs.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp, _ = mo.loss_and_grads(p, X, Y, round_bf16=False)
            flat[idx] = old - eps
            lm, _ = mo.loss_and_grads(p, X, Y, round_bf16=False)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[k].reshape(-1)[idx]) <= 1e-6 * max(1.0, abs(fd)), (k, idx)


def test_bf16_rounding_nearest_even():
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
    import mlp_oracle as mo
    # 1 + 2^-8 is exactly halfway between bf16 1 and 1 + 2^-7: ties to even -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.1415926], dtype=np.float32)
    r = mo.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-7


def test_ws_half_threshold_is_random_start(orc):
    """thr = 2^31 for every spin reproduces the plain random start exactly (gated or not)."""
    rs = np.random.default_rng(3)
    for D, P in ((7, 2), (33, 4)):
        A = rs.integers(-2000, 2000, (3, D, D))
        G = (A + A.transpose(0, 2, 1)) // 2
        h = rs.integers(-3000, 3000, (3, D))
        T = rs.integers(100, 2000, 24).tolist()
        key = int(rs.integers(0, 2 ** 63))
        s0, de0, st0 = orc.arpeggio_fx(G, h, P, 4, T, 50, [1, 2, 3], key, 'rand')
        thr = np.full((3, P * D), 2 ** 31, dtype=np.uint32)
        for gate in (False, True):
            s, de, st, used = orc.arpeggio_ws(G, h, P, 4, T, 50, [1, 2, 3], key, thr, gate=gate)
            assert (s == s0).all() and (de == de0).all() and (st == st0).all() and (used == 1).all()


def test_ws_gate_takes_the_lower_energy_start(orc):
    """With gating, the learned start is kept iff its energy is not above the random start's (brute force)."""
    from test_oracle_golden import energy2, start_spins
    rs = np.random.default_rng(4)
    D, P = 6, 2
    A = rs.integers(-2000, 2000, (2, D, D))
    G = (A + A.transpose(0, 2, 1)) // 2
    h = rs.integers(-3000, 3000, (2, D))
    key = int(rs.integers(0, 2 ** 63))
    thr = rs.integers(1, 2 ** 32 - 1, (2, P * D), dtype=np.uint64).astype(np.uint32)
    s, de, st, used = orc.arpeggio_ws(G, h, P, 6, [1000], 0, [8, 9], key, thr, gate=True)
    assert 0 < used.mean() < 1
    k0, k1 = key & 0xFFFFFFFF, key >> 32
    for b in range(2):
        for r in range(6):
            x = np.array([orc.philox_kat([[0xFFFFFFFF, i, r, 8 + b, k0, k1]])[0, 0] for i in range(P * D)],
                         dtype=np.uint64)
            sl = np.where(x + thr[b] >= 2 ** 32, 1, -1)
            sr = start_spins(orc, D, P, r, 8 + b, key, 'rand')
            el, er = energy2(G[b], h[b], sl, P), energy2(G[b], h[b], sr, P)
            assert used[b, r] == (1 if el <= er else 0)


def test_canonical_labels():
    import torch
    from arpeggio_amd.learned import canonical_spins
    xg = torch.tensor([[-3, -1, 1, 3, 1]], dtype=torch.int8)
    lh = torch.tensor([[-1, -3, 1, 3, 3]], dtype=torch.int8)
    s = canonical_spins(lh, xg, 5)
    # d = 2, -2, 0, 0, 2 -> copy0 = d >= 0, copy1 = d == 2
    assert s.tolist() == [[1, -1, 1, 1, 1, 1, -1, -1, -1, 1]]
    d = s[:, :5].int() + s[:, 5:].int()
    assert (xg.int() + d).tolist() == lh.int().tolist()
