"""The CPU oracle against the reference's golden vectors and known answers (no GPU).

Golden vectors (tests/golden/golden.npz, written by tests/golden/make_golden.py from the reference's own functions):
arpeggio_ref59.anneal (arpeggio/sim/local_code_a59/arpeggio_ref59.py:41-85), chordfx.chord_fx / to_fixed / philox
(arpeggio/mimo/chordfx.py:28-118), mimo front end (arpeggio/mimo/mimo.py), a56an.boot_ratio / mcnemar. Known answers:
Random123 Philox vectors; SHA-256 of the B200 Chord kernel's outputs (data/ber/equivalence/kernel_b200_v2.json)."""
import hashlib

import numpy as np
import pytest


def key_of(rec):
    return int(rec['key'], 16)


# ---------------------------------------------------------------- Philox
def test_philox_kat(golden, orc):
    z, idx = golden
    out = orc.philox_kat(z['philox/in'])
    assert (out == z['philox/out']).all()
    for c, k, w in orc.KAT:
        assert tuple(orc.philox_kat([list(c) + list(k)])[0]) == w


def test_bal_pat(orc):
    for P in (2, 4, 6):
        words = [v for v in range(1 << P) if bin(v).count('1') == P // 2]
        assert [orc.bal_pat(P, i) for i in range(len(words))] == words


# ---------------------------------------------------------------- Arpeggio
def arp_inputs(z, idx, name):
    rec = idx['arp'][name]
    G = z['arpprob/%s/G' % rec['prob']]
    h = z['arpprob/%s/h' % rec['prob']]
    return rec, G, h, z['arp/%s/T' % name], z['arp/%s/inst' % name]


def test_arpeggio_batch_golden(golden, orc):
    z, idx = golden
    n = 0
    for name, rec in idx['arp'].items():
        rec, G, h, T, inst = arp_inputs(z, idx, name)
        s, de, st = orc.arpeggio_fx(G, h, rec['P'], rec['R'], T, rec['q'], inst, key_of(rec), rec['start'],
                                    mode=rec['mode'])
        assert (de == z['arp/%s/best' % name]).all(), name
        assert (s == z['arp/%s/spins' % name]).all(), name
        assert (st[..., 0] == z['arp/%s/flips' % name]).all(), name
        assert (st[..., 1] == z['arp/%s/tests' % name]).all(), name
        n += de.size
    assert n > 300


def test_arpeggio_trace_golden(golden, orc):
    """Accepted (pass, lane) sequence, tests per pass, n1 and init per anneal, as the reference's anneal returns."""
    z, idx = golden
    for name, rec in idx['arp'].items():
        rec, G, h, T, inst = arp_inputs(z, idx, name)
        acc = z['arp/%s/acc' % name]
        nacc = z['arp/%s/nacc' % name]
        off = 0
        for b in range(rec['B']):
            for r in range(rec['R']):
                t = orc.arpeggio_trace(G[b], h[b], rec['P'], T, rec['q'], r, int(inst[b]), key_of(rec), rec['start'],
                                       rec['mode'])
                want = [tuple(x) for x in acc[off:off + nacc[b * rec['R'] + r]].tolist()]
                off += nacc[b * rec['R'] + r]
                assert t['acc'] == want, (name, b, r)
                assert t['tests_k'] == z['arp/%s/tests_k' % name][b, r].tolist(), (name, b, r)
                assert t['n1'] == z['arp/%s/n1' % name][b, r] and t['init'] == z['arp/%s/init' % name][b, r]


def energy2(G, h, s, P):
    """2 E0 (Q units) of copy-major spins s (P D) for the dDI expansion of (G, h): -2 h.s - 2 s^T J s,
    J = -offdiag(T^T G T) (mimo.py:506-523), evaluated by brute force."""
    D = len(h)
    N = P * D
    J = -np.tile(G.astype(np.int64), (P, P))
    J[np.arange(N), np.arange(N)] = 0
    hh = np.tile(h.astype(np.int64), P)
    s = s.astype(np.int64)
    return int(-2 * hh @ s - 2 * s @ J @ s)


def start_spins(orc, D, P, r, inst, key, start):
    k0, k1 = key & 0xFFFFFFFF, key >> 32
    s = np.zeros(P * D, dtype=np.int64)
    for j in range(D):
        if start == 'rand':
            for p in range(P):
                x = orc.philox_kat([[0xFFFFFFFF, p * D + j, r, inst, k0, k1]])[0, 0]
                s[p * D + j] = 1 if x >= 2 ** 31 else -1
        else:
            x = int(orc.philox_kat([[0xFFFFFFFF, j, r, inst, k0, k1]])[0, 0])
            pat = orc.bal_pat(P, (x * {2: 2, 4: 6, 6: 20}[P]) >> 32)
            for p in range(P):
                s[p * D + j] = 1 if (pat >> p) & 1 else -1
    return s


@pytest.mark.parametrize('mode', ['arpeggio', 'par'])
def test_arpeggio_energy_bookkeeping(orc, mode):
    """best de equals 2 E0(best state) - 2 E0(start state), both by brute force, on tiny random problems."""
    rs = np.random.default_rng(11)
    for D, P in ((3, 2), (4, 4), (2, 6), (6, 2)):
        A = rs.integers(-2000, 2000, (D, D))
        G = (A + A.T) // 2
        h = rs.integers(-3000, 3000, D)
        T = rs.integers(100, 2500, 40).tolist()
        for start in ('rand', 'bal'):
            key = int(rs.integers(0, 2 ** 63))
            s, de, st = orc.arpeggio_fx(G[None], h[None], P, 3, T, 150, [9], key, start, mode=mode)
            for r in range(3):
                s0 = start_spins(orc, D, P, r, 9, key, start)
                assert de[0, r] == energy2(G, h, s[0, r], P) - energy2(G, h, s0, P)


def test_par_equals_chord_commit(orc):
    """Mode 'par' on (G, h) is Chord's snapshot rule on the expanded problem with c = 0 (chord_ref.py:1-9):
    same spins, and best de = best 2E0 - 2E0(start)."""
    rs = np.random.default_rng(5)
    for D, P in ((8, 2), (5, 4), (4, 6)):
        A = rs.integers(-1500, 1500, (2, D, D))
        G = (A + A.transpose(0, 2, 1)) // 2
        h = rs.integers(-3000, 3000, (2, D))
        T = rs.integers(200, 2000, 30).astype(np.int32)
        key = int(rs.integers(0, 2 ** 63))
        s, de, _ = orc.arpeggio_fx(G, h, P, 4, T, 100, [3, 4], key, 'rand', mode='par')
        J, hh = orc.expand(G, h, P // 2)
        F = dict(J=J, h=hh, h0=hh, T=T, c=np.zeros_like(T), q=100, S=len(T))
        cs, ce = orc.chord_fx(F, 4, key, P=P, inst0=3)
        assert (cs == s).all()
        for b in range(2):
            for r in range(4):
                s0 = start_spins(orc, D, P, r, 3 + b, key, 'rand')
                e0 = energy2(G[b], h[b], s0, P)
                assert de[b, r] == ce[b, r] - e0


# ---------------------------------------------------------------- Chord
def chord_F(z, name, rec):
    return dict(J=z['chord/%s/J' % name], h=z['chord/%s/h' % name], h0=z['chord/%s/h0' % name],
                T=z['chord/%s/T' % name], c=z['chord/%s/c' % name], q=rec['q'], S=rec['S'])


def test_chord_golden(golden, orc):
    z, idx = golden
    for name, rec in idx['chord'].items():
        F = chord_F(z, name, rec)
        s, e = orc.chord_fx(F, rec['R'], rec['seed'], P=rec['P'], inst0=rec['inst0'])
        assert (e == z['chord/%s/chains_e' % name]).all(), name
        assert (s == z['chord/%s/chains_s' % name]).all(), name
        s, e = orc.chord_fx(F, rec['R'], rec['seed'], P=rec['P'], inst0=rec['inst0'], mode='best')
        assert (e == z['chord/%s/best_e' % name]).all(), name
        assert (s == z['chord/%s/best_s' % name]).all(), name


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_chord_b200_hashes(golden, orc):
    """Cases whose rebuilt integer inputs reproduce the B200 kernel's outputs under chordfx.chord_fx: the oracle must
    give the B200 kernel's SHA-256 digests for spins (int8) and energies (int64)."""
    z, idx = golden
    pinned = 0
    for key, r in idx['b200'].items():
        G, h = z['b200/%s/G' % key].astype(np.int64), z['b200/%s/h' % key]
        J, hh = orc.expand(G, h, r['K'])
        F = dict(J=J, h=hh, h0=hh, T=z['b200/%s/T' % key], c=z['b200/%s/c' % key], q=r['q'], S=r['S'])
        for cname, c in r['cases'].items():
            s, e = orc.chord_fx(F, c['R'], c['seed'], P=c['P'])
            got = [sha(s), sha(e.astype(np.int64))]
            assert got == c['chord_fx'], cname          # the oracle equals the reference's chord_fx here
            if c['chord_fx_equals_b200']:
                assert got == c['b200'], cname
                pinned += 1
    assert pinned >= 2


def test_chord_stats_invariants(golden, orc):
    z, idx = golden
    rec = idx['chord']['x16_K1']
    F = chord_F(z, 'x16_K1', rec)
    st = np.zeros(4, dtype=np.int64)
    orc.chord_fx(F, rec['R'], rec['seed'], P=rec['P'], stats=st)
    B, N = F['h'].shape
    S, P, R = rec['S'], rec['P'], rec['R']
    assert st[3] == S * B * R
    # one block per active spin every 4th update of its copy
    nblk = sum(1 for k in range(S) if (k // P) % 4 == 0)
    assert st[2] == nblk * (N // P) * B * R
    assert st[0] > 0 and 0 < st[1] <= S * B * R


# ---------------------------------------------------------------- float twins reproduce the integer paths
def test_float64_twins_match_integer_paths(golden, orc):
    z, idx = golden
    rec = idx['arp']['x32_K1_hot_k0_rand']
    _, G, h, T, inst = arp_inputs(z, idx, 'x32_K1_hot_k0_rand')
    Q = 4096.0
    s, de, st = orc.arpeggio_fx(G, h, rec['P'], 3, T, rec['q'], inst, key_of(rec), 'rand')
    sf, def_, stf, _ = orc.arpeggio_float(G / Q, h / Q, rec['P'], 3, np.asarray(T) / Q, rec['q'] / Q, inst,
                                          key_of(rec), 'rand', dtype=np.float64)
    assert (sf == s).all() and (stf == st).all()
    assert np.array_equal(def_ * 2 * Q, de.astype(np.float64))
    rec = idx['chord']['x32_K1_kappa']
    F = chord_F(z, 'x32_K1_kappa', rec)
    s, e = orc.chord_fx(F, rec['R'], rec['seed'], P=rec['P'], inst0=rec['inst0'])
    sf, ef, _, _ = orc.chord_float(F['J'] / Q, F['h'] / Q, F['h0'] / Q, F['T'] / Q, F['c'] / Q, F['q'] / Q, rec['R'],
                                   rec['seed'], P=rec['P'], inst0=rec['inst0'], dtype=np.float64)
    assert (sf == s).all()
    assert np.array_equal(ef * 2 * Q, e.astype(np.float64))


# ---------------------------------------------------------------- fixed point, schedules, front end
def test_to_fixed_and_geom(golden, orc):
    z, idx = golden
    t = idx['front']['tofixed']
    Jin, hin = z['tofixed/Jin'], z['tofixed/hin']
    F = orc.to_fixed(Jin, hin, hin, t['S'], t['t0'], t['t1'], t['qf'], t['kappa'])
    for k in ('J', 'h', 'T', 'c'):
        assert (F[k] == z['tofixed/' + k]).all(), k
    assert F['q'] == t['q']
    assert np.array_equal(orc.geom(0.3, 0.2, 256), z['geom/a'])
    assert np.array_equal(orc.geom(1.2, 0.1, 7), z['geom/b'])


@pytest.mark.parametrize('tag', ['x16', 'x32', 'q16', 'x16K2'])
def test_front_end(golden, orc, tag):
    """Float64 restatement of the front end against the reference's float32 outputs (tolerance: float32 rounding),
    then the integer outputs where no rounding boundary is hit."""
    z, idx = golden
    f = idx['front'][tag]
    p = 'front/%s/' % tag
    H, y, M, K = z[p + 'H'], z[p + 'y'], f['M'], f['K']
    zz, Ainv, el2 = orc.mmse(H, y, f['N0'], M)
    assert np.allclose(zz, z[p + 'z'], rtol=2e-4, atol=2e-4)
    assert np.allclose(np.diagonal(Ainv, axis1=1, axis2=2), z[p + 'Ainvdiag'], rtol=1e-4)
    xg = orc.mmse_hard(H, y, f['N0'], M, zz, Ainv)
    assert (xg == z[p + 'lm_hard']).mean() > 0.99
    I = orc.ising_delta_at(H, y, z[p + 'lm_hard'], K)
    assert np.allclose(I['scale'], z[p + 'scale'], rtol=1e-5)
    assert np.allclose(I['h0'], z[p + 'h0'], rtol=1e-4, atol=1e-5)
    assert np.allclose(I['J'][:, 0, :], z[p + 'Jrow0'], atol=1e-5)
    Gi = np.rint(I['G'] * 4096).astype(np.int64)
    assert np.abs(Gi - z[p + 'G']).max() <= 1 and (Gi == z[p + 'G']).mean() > 0.99
    assert np.abs(np.rint(I['h'] * 4096) - z[p + 'hint']).max() <= 1
    Lc = orc.spins_to_levels_delta(z[p + 'S'], z[p + 'lm_hard'], K, M)
    assert (Lc == z[p + 'Lc']).all()
    lh, allc, d = orc.select(H, y, Lc, z[p + 'lm_hard'])
    assert np.allclose(d, z[p + 'd'], rtol=1e-4, atol=1e-4)
    assert (lh == z[p + 'lh']).all()
    be = orc.biterr_per_use(z[p + 'l'], lh, M)
    assert int(be.sum()) == f['bit_errors']
    E0 = orc.energy0(z[p + 'S'], z[p + 'h0'], I['J'])
    assert np.allclose(E0, z[p + 'E0'], rtol=1e-4, atol=1e-4)


def test_boot_ratio_mcnemar(golden, orc):
    z, idx = golden
    for i, r in enumerate(idx['stats']):
        a, b = z['stats/%d/a' % i], z['stats/%d/b' % i]
        assert list(orc.boot_ratio(a, b, B=2000)) == r['boot']
        assert list(orc.mcnemar(a, b)) == r['mcnemar']


def test_draws_match_reference(golden):
    """frames.draw (the product's frame source) reproduces the reference's arpx.draw bytes: H, y, true levels."""
    import torch
    from arpeggio_amd import frames
    z, idx = golden
    for tag, f in idx['front'].items():
        if tag == 'tofixed':
            continue
        p = 'front/%s/' % tag
        assert frames.draw_seed(f['seed'], f['snr'], f['bi']) == f['draw_seed']
        b = frames.draw(f['seed'], f['snr'], f['bi'], z[p + 'H'].shape[0], f['nt'], f['M'], dev='cpu')
        assert np.array_equal(b['H'].numpy(), z[p + 'H']) and np.array_equal(b['y'].numpy(), z[p + 'y'])
        assert np.array_equal(b['l'].numpy(), z[p + 'l']) and b['N0'] == f['N0']
        lt = torch.as_tensor(z[p + 'l'])
        bits = frames.levels_to_bits(lt, f['M'])
        assert bits.shape[-1] == lt.shape[-1] * (2 if f['M'] == 16 else 1)


def test_product_to_fixed_and_geom_golden(golden):
    """The product's chordfx.to_fixed and geom (the drop-in module) equal the reference's on its golden inputs,
    including .5 ties, the Onsager ramp and the schedules of the front-end cases."""
    import torch
    from arpeggio_amd import chordfx
    z, idx = golden
    t = idx['front']['tofixed']
    I = dict(J=torch.as_tensor(z['tofixed/Jin']), h=torch.as_tensor(z['tofixed/hin']),
             h0=torch.as_tensor(z['tofixed/hin']))
    F = chordfx.to_fixed(I, t['S'], t['t0'], t['t1'], t['qf'], t['kappa'])
    for k in ('J', 'h', 'T', 'c'):
        assert (F[k].numpy() == z['tofixed/' + k]).all(), k
    assert F['q'] == t['q']
    assert np.array_equal(chordfx.geom(0.3, 0.2, 256).numpy(), z['geom/a'])
    assert np.array_equal(chordfx.geom(1.2, 0.1, 7).numpy(), z['geom/b'])
    for tag, f in idx['front'].items():
        if tag == 'tofixed':
            continue
        p = 'front/%s/' % tag
        F = chordfx.to_fixed(dict(J=torch.zeros(1, 2, 2), h=torch.zeros(1, 2), h0=torch.zeros(1, 2)), 64, 0.3, 0.2,
                             0.4, 0.3)
        assert (F['T'].numpy() == z[p + 'T']).all() and (F['c'].numpy() == z[p + 'c']).all(), tag
        assert F['q'] == f['q']
