"""Soft output on the GPU against the oracle: MMSE max-log LLRs (mimo.mmse_soft_llr) and list LLRs (mimo.select +
mimo.list_llr) for dDI anneal candidates, 16/32/64 users, QPSK and 16-QAM, scalar and per-frame N0; the soft
detector entry (softlib.detect_llr's methods) end to end."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print('[case]', *a, flush=True)


@pytest.fixture(scope='module')
def mdo(orc):
    import mdi_oracle
    return mdi_oracle


@pytest.mark.parametrize('nt,M,snr', [(16, 16, 14.0), (32, 16, 22.0), (32, 4, 8.0), (64, 16, 20.0)])
def test_mmse_llr_vs_oracle(mdo, nt, M, snr):
    """LLRs from the GPU front end's z and mu against the oracle's float64 MMSE LLRs on the same frames."""
    from arpeggio_amd import frames
    from arpeggio_amd.detector import front_end
    from arpeggio_amd.soft import mmse_llr
    say('mmse llr', nt, M, snr)
    b = frames.draw(31, snr, 0, 40, nt, M, dev='cuda')
    o = front_end(b['H'], b['y'], b['N0'], M, 1, want=('z', 'mu'))
    llr = mmse_llr(o['z'], o['mu'], M).cpu().numpy()
    ref = mdo.mmse_soft_llr(b['H'].cpu().numpy(), b['y'].cpu().numpy(), b['N0'], M)
    assert llr.shape == ref.shape
    np.testing.assert_allclose(llr, ref, rtol=2e-3, atol=2e-3)


def _oracle_list(orc, mdo, H, y, xg, spins, M, N0, lhat):
    """The oracle's list LLRs with the GPU's decision as the best candidate (near-ties of the choice aside)."""
    Lc = orc.spins_to_levels_delta(spins, xg, 1, M)
    lh, allc, d = orc.select(H, y, Lc, xg)
    N0 = np.asarray(N0, np.float64)
    n0 = N0[:, None] if N0.ndim else N0
    ll = mdo.list_llr(H, y, n0, M, allc, d, lhat.astype(np.float64), clip=1e4)
    return ll, lh, allc, d


@pytest.mark.parametrize('nt,M,R,snr,per_frame', [(16, 16, 8, 16.0, False), (32, 16, 16, 22.0, False),
                                                   (32, 4, 4, 6.0, False), (32, 16, 0, 20.0, False),
                                                   (64, 16, 8, 22.0, False), (32, 16, 8, None, True)])
def test_list_llr_vs_oracle(orc, mdo, nt, M, R, snr, per_frame):
    from arpeggio_amd import frames
    from arpeggio_amd.detector import DetectPipeline
    from arpeggio_amd.soft import list_llr
    say('list llr', nt, M, R, snr, per_frame)
    B = 48
    if per_frame:
        b = frames.draw_mixed_snr(44, np.linspace(10.0, 24.0, B), 0, nt, M)
    else:
        b = frames.draw(45, snr, 0, B, nt, M, dev='cuda')
    Rr = max(R, 1)
    pipe = DetectPipeline(B, 2 * nt, 2 * nt, M, Rr, policy='arpeggio', cycles=16, T0=0.1, T1=0.1, q=0.0, key=9)
    pipe.run(b['H'], b['y'], b['N0'])
    spins = pipe.spins[:, :R].contiguous() if R > 0 else None
    llr, lhat, dist = list_llr(b['H'], b['y'], pipe.xg, spins, M, b['N0'], clip=1e4, want_dist=True)
    torch.cuda.synchronize()
    H, y = b['H'].cpu().numpy(), b['y'].cpu().numpy()
    xg = pipe.xg.cpu().numpy()
    sp = spins.cpu().numpy() if spins is not None else np.zeros((B, 0, 4 * nt), np.int8)
    N0 = b['N0'].cpu().numpy() if torch.is_tensor(b['N0']) else b['N0']
    lhat, llr, dist = lhat.cpu().numpy(), llr.cpu().numpy(), dist.cpu().numpy()
    ll, lh, allc, d = _oracle_list(orc, mdo, H, y, xg, sp, M, N0, lhat)
    np.testing.assert_allclose(dist, d, rtol=1e-4, atol=1e-3)
    # the decision is the first minimum of the GPU's own distances, and the oracle's up to float near-ties
    pick = dist.argmin(1)
    assert (lhat == allc[np.arange(B), pick]).all()
    tie = np.array([d[i, pick[i]] <= d[i].min() * (1 + 1e-5) + 1e-6 for i in range(B)])
    assert ((lhat == lh).all(1) | tie).all()
    # LLRs: distances differ by float32 rounding (relative 1e-6 of ||y||^2), divided by N0
    n0 = np.asarray(N0, np.float64)
    n0 = n0[:, None] if n0.ndim else n0
    tol = 1e-5 * (np.abs(y.astype(np.float64)) ** 2).sum(1)[:, None] / n0 + 1e-3
    assert (np.abs(llr - ll) <= tol + 1e-4 * np.abs(ll)).all(), np.abs(llr - ll).max()
    assert np.isfinite(llr).all()
    # the sign of every LLR is the decision's bit unless a neighbour beat the decision
    bits = orc.levels_to_bits(lhat, M).astype(bool)
    agree = (llr > 0) == bits
    assert agree.mean() > 0.97


def test_list_llr_closed_form():
    """One coordinate, two receive rows, QPSK, R = 0 (the MMSE level alone): the list is {x_g} plus its one
    neighbour, so the LLR is (d(-1) - d(+1)) / N0 exactly as computed on the host."""
    from arpeggio_amd.soft import list_llr
    H = torch.tensor([[[1.0], [0.5]]], device='cuda')
    y = torch.tensor([[0.3, -0.2]], device='cuda')
    xg = torch.tensor([[1]], dtype=torch.int8, device='cuda')
    llr, lhat = list_llr(H, y, xg, None, 4, 0.7)
    Hn, yn = H.cpu().double().numpy()[0], y.cpu().double().numpy()[0]
    dm = ((yn - Hn[:, 0] * -1.0) ** 2).sum()
    dp = ((yn - Hn[:, 0] * 1.0) ** 2).sum()
    assert int(lhat.item()) == 1
    assert abs(float(llr.item()) - (dm - dp) / 0.7) < 1e-5


@pytest.mark.parametrize('method', ['mmse', 'arpd8', 'aned4'])
def test_detect_llr_methods(orc, method):
    """The soft detectors end to end: the hard decision is the LLR sign's decision source (the list decision, or
    the MMSE levels), and the LLRs carry the decision's bits on nearly every bit."""
    from arpeggio_amd import frames
    from arpeggio_amd.soft import detect_llr
    say('detect_llr', method)
    b = frames.draw(52, 20.0, 0, 64, 16, 16, dev='cuda')
    llr, lhat = detect_llr(b, method, key=5)
    torch.cuda.synchronize()
    assert llr.shape == (64, 64) and lhat.shape == (64, 32)
    bits = orc.levels_to_bits(lhat.cpu().numpy(), 16).astype(bool)
    assert ((llr.cpu().numpy() > 0) == bits).mean() > 0.95
    be = (orc.levels_to_bits(b['l'].cpu().numpy(), 16) != bits).mean()
    assert be < (0.1 if method == 'mmse' else 0.05)
