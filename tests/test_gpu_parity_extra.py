"""GPU parity cases beyond the main suites: the front end with a per-frame noise density, the QPSK and RI decodes,
the FP32 Chord whose couplings exceed the LDS (column-streamed from L2), the first-accept kernels' variants for a
pass count that is not a whole number of Philox blocks, the mixed-size launch at such a pass count, and the learned
start pipeline on a side stream."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print('[case]', *a, flush=True)


def near_half(x, tol):
    return np.abs(np.abs(x - np.floor(x)) - 0.5) < tol


@pytest.fixture
def fe_form():
    from arpeggio_amd import _lib
    L = _lib.load()
    yield lambda f: _lib.check(L.arp_set_front_end_form(f), 'set_front_end_form')
    L.arp_set_front_end_form(0)


@pytest.mark.parametrize('form', [0, 2])
@pytest.mark.parametrize('nt,M', [(16, 16), (32, 16), (64, 16), (32, 4)])
def test_front_end_per_frame_n0(orc, fe_form, nt, M, form):
    """Per-frame N0 (a (B,) tensor of distinct values, 8-24 dB) against the oracle run frame by frame (form 2: the
    complex-lift front end where it applies)."""
    from arpeggio_amd import frames
    from arpeggio_amd.detector import front_end
    say('front end per-frame N0', nt, M, 'form', form)
    fe_form(form)
    B = 24
    snrs = np.linspace(8.0, 24.0, B)
    b = frames.draw_mixed_snr(91, snrs, 0, nt, M)
    o = front_end(b['H'], b['y'], b['N0'], M, 1, want=('xg', 'Gi', 'hi', 'scale', 'Gf', 'hf', 'z', 'mu'))
    torch.cuda.synchronize()
    H, y = b['H'].cpu().numpy(), b['y'].cpu().numpy()
    N0 = b['N0'].cpu().numpy().astype(np.float64)
    assert len(set(N0.tolist())) == B
    got = {k: v.cpu().numpy() for k, v in o.items()}
    for i in range(B):
        z, Ainv, _ = orc.mmse(H[i:i + 1], y[i:i + 1], float(N0[i]), M)
        Gm = H[i].T.astype(np.float64) @ H[i]
        mu = np.diagonal(Ainv[0] @ Gm)
        np.testing.assert_allclose(got['z'][i], z[0], rtol=1e-3, atol=1e-3)
        np.testing.assert_allclose(got['mu'][i], mu, rtol=1e-4, atol=1e-5)
        v = z[0] / mu
        edge = np.abs(v - np.round(v)) < 1e-3
        assert ((got['xg'][i] == orc.quantize_levels(v, M)) | edge).all(), i
        I = orc.ising_delta_at(H[i:i + 1], y[i:i + 1], got['xg'][i:i + 1].astype(np.float64), 1)
        np.testing.assert_allclose(got['scale'][i], I['scale'][0], rtol=1e-5)
        Gq, hq = I['G'][0] * 4096, I['hD'][0] * 4096
        assert ((got['Gi'][i] == np.rint(Gq)) | near_half(Gq, 0.05)).all(), i
        assert ((got['hi'][i] == np.rint(hq)) | near_half(hq, 0.05)).all(), i


def _check_choice(orc, H, y, allc, lhat, dist, be, l, M):
    """The GPU's choice is the first minimum of its own distances, and the oracle's choice or a near-tie of it."""
    B = H.shape[0]
    d = orc.dist(H, y, allc)
    np.testing.assert_allclose(dist, d, rtol=1e-4, atol=1e-3)
    pick = dist.argmin(1)
    assert (lhat == allc[np.arange(B), pick]).all()
    best = d.argmin(1)
    for i in range(B):
        if not (lhat[i] == allc[i, best[i]]).all():
            assert d[i, pick[i]] <= d[i].min() * (1 + 1e-5) + 1e-6
    assert (be == orc.biterr_per_use(l, lhat, M)).all()


@pytest.mark.parametrize('nt,R', [(16, 8), (32, 16), (64, 8)])
def test_decode_qpsk_delta(orc, nt, R):
    from arpeggio_amd import frames
    from arpeggio_amd.detector import decode_select, front_end
    say('decode QPSK dDI', nt, R)
    M, B = 4, 40
    b = frames.draw(78, 8.0, 0, B, nt, M, dev='cuda')
    o = front_end(b['H'], b['y'], b['N0'], M, 1, want=('xg',))
    g = torch.Generator(device='cuda')
    g.manual_seed(5)
    spins = torch.where(torch.rand(B, R, 4 * nt, device='cuda', generator=g) < 0.5, -1, 1).to(torch.int8)
    lhat, be, dist = decode_select(b['H'], b['y'], o['xg'], spins, M, 2, True, ltrue=b['l'].to(torch.int8),
                                   want_dist=True)
    torch.cuda.synchronize()
    H, y, xg = b['H'].cpu().numpy(), b['y'].cpu().numpy(), o['xg'].cpu().numpy()
    Lc = orc.spins_to_levels_delta(spins.cpu().numpy(), xg, 1, M)
    allc = np.concatenate([Lc, xg[:, None, :].astype(np.float64)], 1)
    _check_choice(orc, H, y, allc, lhat.cpu().numpy(), dist.cpu().numpy(), be.cpu().numpy(), b['l'].cpu().numpy(), M)


@pytest.mark.parametrize('nt,M,R', [(16, 4, 8), (16, 16, 8), (32, 16, 5), (64, 4, 4)])
def test_decode_ri(orc, nt, M, R):
    """RI spins (delta = 0): levels s (QPSK) or 2 s_MSB + s_LSB (16-QAM), the reference's spins_to_levels."""
    from arpeggio_amd import frames
    from arpeggio_amd.detector import decode_select, front_end
    say('decode RI', nt, M, R)
    B = 32
    b = frames.draw(79, 12.0, 0, B, nt, M, dev='cuda')
    o = front_end(b['H'], b['y'], b['N0'], M, 1, want=('xg',))
    D = 2 * nt
    N = D if M == 4 else 2 * D
    g = torch.Generator(device='cuda')
    g.manual_seed(6)
    spins = torch.where(torch.rand(B, R, N, device='cuda', generator=g) < 0.5, -1, 1).to(torch.int8)
    lhat, be, dist = decode_select(b['H'], b['y'], o['xg'], spins, M, 1, False, ltrue=b['l'].to(torch.int8),
                                   want_dist=True)
    torch.cuda.synchronize()
    H, y, xg = b['H'].cpu().numpy(), b['y'].cpu().numpy(), o['xg'].cpu().numpy()
    Lc = orc.spins_to_levels(spins.cpu().numpy().astype(np.float64), M, D)
    allc = np.concatenate([Lc, xg[:, None, :].astype(np.float64)], 1)
    _check_choice(orc, H, y, allc, lhat.cpu().numpy(), dist.cpu().numpy(), be.cpu().numpy(), b['l'].cpu().numpy(), M)


def test_chord_f32_couplings_streamed_from_l2(orc):
    """N = 256 (64 users, dDI K = 1): the float couplings (256 KiB) exceed the LDS, so the kernel reads J's rows from
    L2; bit-exact against the oracle's FP32 twin, energies within 1e-5 relative of float64."""
    from arpeggio_amd import chordfx, frames
    from arpeggio_amd.detector import front_end
    say('chord f32 N=256')
    B, nt = 6, 64
    b = frames.draw(31, 22.0, 0, B, nt, 16, dev='cuda')
    o = front_end(b['H'], b['y'], b['N0'], 16, 1, want=('xg', 'Jf', 'hNf'))
    torch.cuda.synchronize()
    J, hN = o['Jf'], o['hNf']
    assert J.shape[1] == 256 and J.numel() * 4 > 160 * 1024
    T, c = chordfx.float_schedule(64, 0.3, 0.2, 0.3)
    R, seed, q = 3, 2 ** 33 + 5, 0.4
    s, e, u, fl = chordfx.Kernel().chord_float(J, hN, hN, T, c, q, R, seed, P=2, inst0=4, want_fields=True)
    torch.cuda.synchronize()
    Jn, hn = J.cpu().numpy(), hN.cpu().numpy()
    os_, oe, ou, ofl = orc.chord_float(Jn, hn, hn, T.numpy(), c.numpy(), q, R, seed, P=2, inst0=4, dtype=np.float32)
    assert (s.cpu().numpy() == os_).all() and (fl.cpu().numpy() == ofl).all()
    assert np.array_equal(e.cpu().numpy(), oe) and np.array_equal(u.cpu().numpy(), ou)
    sb = s.cpu().numpy().astype(np.float64)
    E0 = -(sb * hn[:, None]).sum(-1) - np.einsum('brn,bnm,brm->br', sb, Jn.astype(np.float64), sb)
    assert np.abs(e.cpu().numpy() - E0).max() <= 1e-5 * np.abs(E0).max()


@pytest.mark.parametrize('D,L,start', [(32, 30, 'rand'), (64, 20, 'rand'), (128, 12, 'bal'), (64, 13, 'bal')])
def test_static_d_partial_philox_blocks(orc, D, L, start):
    """The compile-time-D first-accept kernels (D = 32, 64, 128, two copies) when L is not a whole number of 4P-pass
    Philox blocks (the bounded pass loop)."""
    from arpeggio_amd import arpeggiofx
    say('static D partial blocks', D, L, start)
    rs = np.random.default_rng(D + L)
    B, R = 5, 6
    A = rs.normal(size=(B, D + 4, D))
    G = np.einsum('bri,brj->bij', A, A)
    hv = rs.normal(size=(B, D)) * 3
    m = np.maximum(np.abs(G).reshape(B, -1).max(-1), np.abs(hv).max(-1))
    G = np.rint(G / m[:, None, None] * 4096).astype(np.int32)
    h = np.rint(hv / m[:, None] * 4096).astype(np.int32)
    T = rs.integers(50, 3000, L).tolist()
    inst = [3 + i for i in range(B)]
    s, de, st = arpeggiofx.solve(G, h, 2, R, T, 0, inst, 0xFEED, start)
    os_, ode, ost = orc.arpeggio_fx(G, h, 2, R, T, 0, inst, 0xFEED, start)
    assert (de == ode).all() and (st == ost).all() and (s == os_).all()


def test_batched_partial_philox_blocks(orc):
    """The persistent mixed-size launch with 30 passes (not a multiple of 8): per class equal to the oracle."""
    from arpeggio_amd.batch import BatchedDetector, draw_mixed, prepare
    say('batched 30 passes')
    mixed = prepare(draw_mixed(6, 120, snr_lo=10.0, snr_hi=22.0))
    det = BatchedDetector([dict(nt=f['nt'], M=f['M'], B=f['l'].shape[0], idx=f['idx']) for f in mixed], R=3,
                          passes=30, key=0x77AA, shape=(5, 4, 66048))
    det.run(mixed)
    torch.cuda.synchronize()
    for c in det.cls:
        s, de, st = orc.arpeggio_fx(c['Gi'].cpu().numpy(), c['hi'].cpu().numpy(), 2, det.R, det.T.cpu().numpy(),
                                    det.q, c['inst'].cpu().numpy().astype(np.int64) & 0xFFFFFFFF, det.key, 'rand')
        assert (c['spins'].cpu().numpy() == s).all() and (c['de'].cpu().numpy() == de).all(), (c['nt'], c['M'])
        assert (c['st'].cpu().numpy() == st).all()


def test_learned_pipeline_on_side_stream():
    """DetectPipeline.run(stream=s) with the learned start runs the MLP on s too: the result equals the default-stream
    run bit for bit (the side stream starts behind the inputs' producer and the host reads after a full sync)."""
    from arpeggio_amd import frames
    from arpeggio_amd.detector import DetectPipeline
    from arpeggio_amd.learned import DEFAULT_WEIGHTS, WarmStartMLP
    say('learned pipeline side stream')
    model = WarmStartMLP.load(DEFAULT_WEIGHTS)
    B = 256
    b = frames.draw(12, 23.0, 0, B, 32, 16, dev='cuda')
    lt = b['l'].to(torch.int8).contiguous()
    outs = []
    for side in (False, True):
        pipe = DetectPipeline(B, 64, 64, 16, 8, cycles=32, start='learned', model=model, key=5)
        if side:
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            pipe.run(b['H'], b['y'], b['N0'], lt, stream=s)
            torch.cuda.synchronize()
        else:
            pipe.run(b['H'], b['y'], b['N0'], lt)
            torch.cuda.synchronize()
        outs.append((pipe.spins.clone(), pipe.thr.clone(), pipe.lhat.clone(), pipe.biterr.clone()))
    assert all(torch.equal(a, c) for a, c in zip(*outs))


@pytest.mark.parametrize('start,learned', [('rand', False), ('bal', False), ('rand', True)])
def test_list_anneal_int16_gram_equals_int32(start, learned):
    """arp_anneal_arpeggio_list on the front end's int16 Gram (Gh) gives exactly the anneals of the int32 Gram (Gt),
    with and without a frame list and the learned-start thresholds."""
    from arpeggio_amd import _lib, frames
    from arpeggio_amd.detector import arpeggio_schedule, front_end
    say('int16 gram', start, learned)
    B, nt, R = 96, 32, 6
    D = 2 * nt
    b = frames.draw(77, 20.0, 0, B, nt, 16, dev='cuda')
    o = front_end(b['H'], b['y'], b['N0'], 16, 1, want=('Gi', 'hi', 'Gh'))
    T, q = arpeggio_schedule(12, 0.1, 0.1, 0.0, 1)
    Tt = torch.tensor(T, dtype=torch.int32, device='cuda')
    inst = torch.arange(B, dtype=torch.int32, device='cuda')
    thr = torch.randint(0, 2 ** 31, (B, 2 * D), dtype=torch.int64, device='cuda').to(torch.int32) if learned else None
    lst = torch.tensor(sorted(np.random.default_rng(3).choice(B, 40, replace=False)), dtype=torch.int32,
                       device='cuda')
    cnt = torch.tensor([40], dtype=torch.int32, device='cuda')
    L = _lib.load()
    outs = []
    for gt, gh in ((o['Gi'], None), (None, o['Gh'])):
        for ls, cn in ((None, None), (lst, cnt)):
            s = torch.empty(B, R, 2 * D, dtype=torch.int8, device='cuda')
            de = torch.empty(B, R, dtype=torch.int64, device='cuda')
            st = torch.empty(B, R, 2, dtype=torch.int64, device='cuda')
            used = torch.empty(B, R, dtype=torch.int8, device='cuda')
            _lib.check(L.arp_anneal_arpeggio_list(_lib.ptr(gt), _lib.ptr(gh), _lib.ptr(o['hi']), _lib.ptr(Tt),
                                                  len(T), q, 2, R, D, B, _lib.ptr(inst), 0x1234, 0x5678,
                                                  0 if start == 'rand' else 1, _lib.ptr(thr), 1, 0, _lib.ptr(ls),
                                                  _lib.ptr(cn), _lib.ptr(s), _lib.ptr(de), _lib.ptr(st),
                                                  _lib.ptr(used) if learned else None, _lib.stream_ptr()), 'list')
            torch.cuda.synchronize()
            n = B if ls is None else 40
            outs.append((s[:n].cpu().numpy(), de[:n].cpu().numpy(), st[:n].cpu().numpy()))
    for a, c in ((outs[0], outs[2]), (outs[1], outs[3])):
        assert all((x == y).all() for x, y in zip(a, c))
