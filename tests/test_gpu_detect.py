"""GPU parity of the front end (LMMSE + dDI Ising build + fixed point), decode/select, and the whole detection chain.

Float stages are compared with the float64 oracle within float32 tolerances; the integer hand-off (G_int, h_int,
x_g) must equal the oracle's rounding except where the oracle's value sits within 1e-3 of a rounding boundary; the
anneal and decode stages are bit-exact on the integers the GPU front end produced."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print('[case]', *a, flush=True)


def frames(nt, M, snr, B, seed=4242, bi=0, **kw):
    from arpeggio_amd import frames as fr
    return fr.draw(seed, snr, bi, B, nt, M, dev='cuda', **kw)


def near_half(x, tol=1e-3):
    return np.abs(np.abs(x - np.floor(x)) - 0.5) < tol


@pytest.fixture
def fe_form():
    """Selects a front-end form for one test and restores the default afterwards."""
    from arpeggio_amd import _lib
    L = _lib.load()
    yield lambda f: _lib.check(L.arp_set_front_end_form(f), 'set_front_end_form')
    L.arp_set_front_end_form(0)


@pytest.mark.parametrize('form', [0, 1, 2])
@pytest.mark.parametrize('nt,M,snr,K', [(16, 16, 16.0, 1), (32, 16, 20.0, 1), (16, 4, 8.0, 1), (64, 16, 22.0, 1),
                                        (16, 16, 14.0, 2), (8, 16, 12.0, 3), (31, 16, 20.0, 1), (24, 4, 10.0, 1),
                                        (32, 4, 10.0, 2)])
def test_front_end_vs_oracle(orc, fe_form, nt, M, snr, K, form):
    """The front-end forms (0: matrix-core block sweep for D, Mr <= 64; 1: register sweep; 2: the complex-lift form
    for 16 and 32 users, the others as 0) against the oracle; the int16 Gram equals the int32 one; the expanded Chord
    inputs (register-sweep kernels) expand the same integers."""
    from arpeggio_amd.detector import front_end
    say('front_end', nt, M, snr, K, 'form', form)
    fe_form(form)
    b = frames(nt, M, snr, 64)
    o = front_end(b['H'], b['y'], b['N0'], M, K, want=('xg', 'Gi', 'hi', 'scale', 'Gf', 'hf', 'z', 'mu', 'Gh'))
    torch.cuda.synchronize()
    assert (o['Gh'].cpu().numpy().astype(np.int32) == o['Gi'].cpu().numpy()).all()
    H, y = b['H'].cpu().numpy(), b['y'].cpu().numpy()
    z, Ainv, _ = orc.mmse(H, y, b['N0'], M)
    assert np.allclose(o['z'].cpu().numpy(), z, rtol=1e-3, atol=1e-3)
    Gm = np.swapaxes(H, 1, 2).astype(np.float64) @ H
    mu = np.diagonal(Ainv @ Gm, axis1=1, axis2=2)
    assert np.allclose(o['mu'].cpu().numpy(), mu, rtol=1e-4, atol=1e-5)
    xg = orc.quantize_levels(z / mu, M)
    v = z / mu
    edge = np.abs(v - np.round(v)) < 1e-3       # z/mu on a decision boundary (an even integer)
    got_xg = o['xg'].cpu().numpy()
    assert ((got_xg == xg) | edge).all()
    # Ising build on the GPU's own x_g (the integer hand-off point)
    I = orc.ising_delta_at(H, y, got_xg.astype(np.float64), K)
    assert np.allclose(o['scale'].cpu().numpy(), I['scale'], rtol=1e-5)
    Gf, hf = o['Gf'].cpu().numpy(), o['hf'].cpu().numpy()
    assert np.allclose(Gf, I['G'], atol=1e-5)
    assert np.allclose(hf, I['hD'], atol=1e-5)
    Gi, hi = o['Gi'].cpu().numpy(), o['hi'].cpu().numpy()
    # the fixed point is the half-even rounding of the GPU's own scaled floats (x 2^12 is exact in float32)
    assert (Gi == np.rint(Gf * np.float32(4096))).all() and (hi == np.rint(hf * np.float32(4096))).all()
    assert (Gi == np.swapaxes(Gi, 1, 2)).all()
    Gq, hq = I['G'] * 4096, I['hD'] * 4096
    assert ((Gi == np.rint(Gq)) | near_half(Gq, 0.05)).all()
    assert ((hi == np.rint(hq)) | near_half(hq, 0.05)).all()
    if 2 * K * 2 * nt <= 256:
        oj = front_end(b['H'], b['y'], b['N0'], M, K, want=('Gi', 'hi', 'J', 'hN'))
        torch.cuda.synchronize()
        J, hN = orc.expand(oj['Gi'].cpu().numpy().astype(np.int64), oj['hi'].cpu().numpy(), K)
        assert (oj['J'].cpu().numpy() == J).all() and (oj['hN'].cpu().numpy() == hN).all()


@pytest.mark.parametrize('nt,M,snr', [(32, 16, 23.7), (16, 16, 16.0), (32, 4, 10.0), (64, 16, 28.0), (64, 4, 12.0)])
def test_complex_lift_form_against_general(fe_form, nt, M, snr):
    """Form 2 (Hermitian Gram and complex sweep from half of H) against form 0 on the same real-lift frames, learned
    features included: decisions equal but for float near-ties, integers within one unit."""
    from arpeggio_amd.detector import front_end
    from arpeggio_amd.learned import feat_ld, features
    say('complex lift vs general', nt, M, snr)
    b = frames(nt, M, snr, 4096 if nt < 64 else 1024, seed=77)
    out = []
    for form in (0, 2):
        fe_form(form)
        o = front_end(b['H'], b['y'], b['N0'], M, 1, want=('xg', 'Gi', 'hi', 'scale', 'z', 'mu'))
        o['feat'] = features(b['H'], b['y'], b['N0'], M)['feat']
        torch.cuda.synchronize()
        out.append({k: v.cpu().float().numpy() if v.dtype == torch.bfloat16 else v.cpu().numpy() for k, v in o.items()})
    a, c = out
    assert a['feat'].shape[1] == feat_ld(2 * nt)
    assert (a['xg'] == c['xg']).mean() > 0.9999
    same = (a['xg'] == c['xg']).all(1)
    assert np.abs(a['Gi'][same] - c['Gi'][same]).max() <= 1 and (a['Gi'][same] == c['Gi'][same]).mean() > 0.999
    assert np.abs(a['hi'][same] - c['hi'][same]).max() <= 2
    np.testing.assert_allclose(c['z'], a['z'], rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(c['mu'], a['mu'], rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(c['scale'][same], a['scale'][same], rtol=1e-4)
    np.testing.assert_allclose(c['feat'][same], a['feat'][same], rtol=2e-2, atol=2e-2)


@pytest.mark.parametrize('nt,M', [(32, 16), (16, 16), (32, 4), (64, 16)])
def test_decode_lift_form_equals_general(fe_form, nt, M):
    """With form 2 the two-copy decodes (one wave at 16 and 32 users, two at 64) read only the first D/2 columns of
    the real-lift H and rebuild each row: the distances, decisions and bit errors equal the full-H decode's bit for
    bit."""
    from arpeggio_amd.detector import decode_select, front_end
    say('decode lift', nt, M)
    b = frames(nt, M, 14.0, 2048, seed=5)
    o = front_end(b['H'], b['y'], b['N0'], M, 1, want=('xg',))
    g = torch.Generator(device='cuda').manual_seed(3)
    spins = (torch.randint(0, 2, (2048, 9, 4 * nt), generator=g, device='cuda') * 2 - 1).to(torch.int8)
    res = []
    for form in (0, 2):
        fe_form(form)
        lhat, be, dist = decode_select(b['H'], b['y'], o['xg'], spins, M, P=2, delta=True, ltrue=b['l'],
                                       want_dist=True)
        torch.cuda.synchronize()
        res.append((lhat.clone(), be.clone(), dist.clone()))
    assert all(torch.equal(x, z) for x, z in zip(*res))


def test_front_end_golden(golden):
    """The reference's own float32 front end (mimo.mmse / mmse_hard / ising_delta_at / to_fixed + compress) on its
    seeded frames: the GPU's x_g and integer Gram agree except at rounding boundaries."""
    from arpeggio_amd.detector import front_end
    z, idx = golden
    for tag, f in idx['front'].items():
        if tag == 'tofixed':
            continue
        say('front golden', tag)
        p = 'front/%s/' % tag
        H, y = torch.as_tensor(z[p + 'H']).cuda(), torch.as_tensor(z[p + 'y']).cuda()
        o = front_end(H, y, f['N0'], f['M'], f['K'], want=('xg', 'Gi', 'hi'))
        torch.cuda.synchronize()
        xg = o['xg'].cpu().numpy()
        assert (xg == z[p + 'lm_hard']).mean() > 0.99
        same = (xg == z[p + 'lm_hard']).all(1)
        Gi, Gr = o['Gi'].cpu().numpy()[same], z[p + 'G'][same]
        assert np.abs(Gi - Gr).max() <= 1 and (Gi == Gr).mean() > 0.999


@pytest.mark.parametrize('R,delta,nt', [(16, True, 32), (1, True, 32), (33, True, 32), (8, True, 64), (8, True, 16)])
def test_decode_select_vs_oracle(orc, R, delta, nt):
    from arpeggio_amd.detector import decode_select, front_end
    say('decode', R, delta, nt)
    M, B = 16, 48
    b = frames(nt, M, 18.0, B, seed=77)
    o = front_end(b['H'], b['y'], b['N0'], M, 1, want=('xg',))
    g = torch.Generator(device='cuda')
    g.manual_seed(3)
    spins = torch.where(torch.rand(B, R, 4 * nt, device='cuda', generator=g) < 0.5, -1, 1).to(torch.int8)
    lt = b['l'].to(torch.int8)
    lhat, be, dist = decode_select(b['H'], b['y'], o['xg'], spins, M, 2, delta, ltrue=lt, want_dist=True)
    torch.cuda.synchronize()
    H, y = b['H'].cpu().numpy(), b['y'].cpu().numpy()
    xg = o['xg'].cpu().numpy()
    Lc = orc.spins_to_levels_delta(spins.cpu().numpy(), xg, 1, M)
    lh, allc, d = orc.select(H, y, Lc, xg)
    dg = dist.cpu().numpy()
    assert np.allclose(dg, d, rtol=1e-4, atol=1e-3)
    pick = dg.argmin(1)                                   # numpy argmin: the first minimum, as the kernel's scan
    lg = lhat.cpu().numpy()
    assert (lg == allc[np.arange(B), pick]).all()
    # the GPU's choice is the oracle's, or a near-tie of distinct candidates within float32 rounding
    for i in range(B):
        if not (lg[i] == lh[i]).all():
            assert d[i, pick[i]] <= d[i].min() * (1 + 1e-5) + 1e-6
    be_o = orc.biterr_per_use(b['l'].cpu().numpy(), lhat.cpu().numpy(), M)
    assert (be.cpu().numpy() == be_o).all()


@pytest.mark.parametrize('policy,nt,rho,R', [('arpeggio', 32, 0.0, 16), ('chord', 32, 0.0, 16),
                                             ('arpeggio', 64, 0.5, 32), ('arpeggio', 16, 0.0, 8)])
def test_pipeline_chain_bit_exact(orc, policy, nt, rho, R):
    """The whole chain (front end -> anneal -> decode) equals the oracle chain run on the GPU front end's integers;
    the 64-user case is the Kronecker-correlated channel (rho = 0.5 at both ends) of the 64 x 64 configuration."""
    from arpeggio_amd.detector import DetectPipeline
    M, B = 16, 40 if nt <= 32 else 24
    say('pipeline', policy, nt, rho, R)
    b = frames(nt, M, 20.0 if nt <= 32 else 24.0, B, seed=11, rho_rx=rho, rho_tx=rho)
    cyc = 64 if policy == 'arpeggio' else 128
    pipe = DetectPipeline(B, 2 * nt, 2 * nt, M, R, policy=policy, cycles=cyc, T0=0.1 if policy == 'arpeggio' else 0.3,
                          T1=0.1 if policy == 'arpeggio' else 0.2, q=0.0 if policy == 'arpeggio' else 0.4,
                          kappa=0.0 if policy == 'arpeggio' else 0.3, key=0xDEADBEEF12345, inst0=5)
    lt = b['l'].to(torch.int8)
    lhat, be = pipe.run(b['H'], b['y'], b['N0'], lt)
    torch.cuda.synchronize()
    xg = pipe.xg.cpu().numpy()
    if policy == 'arpeggio':
        G, h = pipe.Gi.cpu().numpy(), pipe.hi.cpu().numpy()
        s, de, st = orc.arpeggio_fx(G, h, 2, R, pipe.T.cpu().numpy(), pipe.q, np.arange(B) + 5, 0xDEADBEEF12345, 'rand')
        assert (pipe.spins.cpu().numpy() == s).all() and (pipe.de.cpu().numpy() == de).all()
        assert (pipe.st.cpu().numpy() == st).all()
    else:
        F = dict(J=pipe.J.cpu().numpy(), h=pipe.hN.cpu().numpy(), h0=pipe.hN.cpu().numpy(), T=pipe.T.cpu().numpy(),
                 c=pipe.c.cpu().numpy(), q=pipe.q, S=pipe.S)
        s, e = orc.chord_fx(F, R, 0xDEADBEEF12345, P=2, inst0=5)
        assert (pipe.spins.cpu().numpy() == s).all() and (pipe.e.cpu().numpy() == e).all()
    H, y = b['H'].cpu().numpy(), b['y'].cpu().numpy()
    Lc = orc.spins_to_levels_delta(s, xg, 1, M)
    lh, allc, d = orc.select(H, y, Lc, xg)
    lg = lhat.cpu().numpy()
    # the GPU's choice is the oracle's, or a distinct candidate within float32 rounding of the oracle's minimum
    for i in range(B):
        if not (lg[i] == lh[i]).all():
            same = (allc[i] == lg[i]).all(1)
            assert same.any(), i
            assert d[i][same].min() <= d[i].min() * (1 + 1e-5) + 1e-6, i
    be_o = orc.biterr_per_use(b['l'].cpu().numpy(), lhat.cpu().numpy(), M)
    assert (be.cpu().numpy() == be_o).all()
