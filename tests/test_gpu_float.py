"""GPU parity of the FP32 annealers: bit-exact against the oracle's FP32 twins (same operation order, no fused
multiply-add), energies within 1e-5 relative of a float64 recomputation, and BER of the FP32 path inside the paired
bootstrap CI of the fixed-point path (arpeggio/sim/a56an.py:74-95)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print('[case]', *a, flush=True)


def problem(nt, snr, B, seed=31, K=1):
    from arpeggio_amd import frames
    from arpeggio_amd.detector import front_end
    b = frames.draw(seed, snr, 0, B, nt, 16, dev='cuda')
    want = ('xg', 'Gf', 'hf', 'Gi', 'hi') + (('Jf', 'hNf') if 4 * K * nt <= 256 else ())
    o = front_end(b['H'], b['y'], b['N0'], 16, K, want=want)
    torch.cuda.synchronize()
    return b, o


def e0_expanded(G, h, s, P):
    """E0 = -h.s - s^T J s of copy-major spins for the dDI expansion of float (G, h), in float64."""
    B, D, _ = G.shape
    N = P * D
    J = -np.tile(G.astype(np.float64), (1, P, P))
    J[:, np.arange(N), np.arange(N)] = 0
    hh = np.tile(h.astype(np.float64), (1, P))
    s = s.astype(np.float64)
    return -(s * hh[:, None]).sum(-1) - np.einsum('brn,bnm,brm->br', s, J, s)


@pytest.mark.parametrize('nt,P,start', [(16, 2, 'rand'), (32, 2, 'rand'), (32, 4, 'bal'), (64, 2, 'rand'),
                                        (24, 6, 'rand')])
def test_arpeggio_f32_vs_oracle(orc, nt, P, start):
    from arpeggio_amd import arpeggiofx
    say('arpeggio f32', nt, P, start)
    b, o = problem(nt, 20.0, 24, K=P // 2)
    G, h = o['Gf'].cpu().numpy(), o['hf'].cpu().numpy()
    L = 48 * P // 2
    T = np.repeat(np.geomspace(0.3, 0.05, L // P).astype(np.float32), P)
    R, key, q = 7, 0x1234567890ABCDEF, np.float32(0.05)
    inst = np.arange(24) + 100
    s, de, st, w = arpeggiofx.solve_float_device(torch.as_tensor(G), torch.as_tensor(h), P, R, T, q, inst, key, start,
                                                 want_fields=True)
    torch.cuda.synchronize()
    os_, ode, ost, ow = orc.arpeggio_float(G, h, P, R, T, q, inst, key, start, dtype=np.float32)
    assert (s.cpu().numpy() == os_).all() and (st.cpu().numpy() == ost).all()
    assert np.array_equal(de.cpu().numpy(), ode) and np.array_equal(w.cpu().numpy(), ow)
    # energy bookkeeping in FP32 against float64: best dE = E0(best) - E0(start)
    from test_oracle_golden import start_spins
    s0 = np.stack([np.stack([start_spins(orc, G.shape[1], P, r, int(inst[bb]), key, start) for r in range(R)])
                   for bb in range(24)])
    e_best, e_start = e0_expanded(G, h, s.cpu().numpy(), P), e0_expanded(G, h, s0, P)
    ref = e_best - e_start
    scale = np.maximum(np.abs(e_start), np.abs(e_best)).max()
    assert np.abs(de.cpu().numpy() - ref).max() <= 1e-5 * scale


@pytest.mark.parametrize('nt,P,kappa', [(16, 2, 0.0), (32, 2, 0.3), (16, 1, 0.3), (48, 6, 0.0)])
def test_chord_f32_vs_oracle(orc, nt, P, kappa):
    from arpeggio_amd import chordfx
    say('chord f32', nt, P, kappa)
    K = 3 if P == 6 else 1
    b, o = problem(nt if P != 6 else 16, 20.0, 12, K=K)
    J, hN = o['Jf'], o['hNf']
    S = 96
    T, c = chordfx.float_schedule(S, 0.3, 0.2, kappa)
    R, seed, q = 5, 2 ** 40 + 17, 0.4
    kern = chordfx.Kernel()
    s, e, u, fl = kern.chord_float(J, hN, hN, T, c, q, R, seed, P=P, inst0=9, want_fields=True)
    torch.cuda.synchronize()
    Jn, hn = J.cpu().numpy(), hN.cpu().numpy()
    os_, oe, ou, ofl = orc.chord_float(Jn, hn, hn, T.numpy(), c.numpy(), q, R, seed, P=P, inst0=9, dtype=np.float32)
    assert (s.cpu().numpy() == os_).all() and (fl.cpu().numpy() == ofl).all()
    assert np.array_equal(e.cpu().numpy(), oe) and np.array_equal(u.cpu().numpy(), ou)
    sb = s.cpu().numpy().astype(np.float64)
    E0 = -(sb * hn[:, None]).sum(-1) - np.einsum('brn,bnm,brm->br', sb, Jn.astype(np.float64), sb)
    assert np.abs(e.cpu().numpy() - E0).max() <= 1e-5 * np.abs(E0).max()


@pytest.mark.parametrize('nt,cycles,snrs', [
    (32, 64, (16.0, 20.0, 24.0)),
    # 16 users x 16 rx, 16-QAM, 8 copies, 256 cycles, 4,096 frames per point, 10-20 dB in 2 dB steps
    (16, 256, (10.0, 12.0, 14.0, 16.0, 18.0, 20.0)),
])
def test_float_path_ber_in_bootstrap_ci(orc, nt, cycles, snrs):
    """Per-use bit errors of the FP32 and the fixed-point Arpeggio chains on identical frames: the paired bootstrap
    CI of their BER ratio contains 1 at every SNR point."""
    from arpeggio_amd import arpeggiofx, frames
    from arpeggio_amd.detector import DetectPipeline, decode_select
    B, R = 4096, 8
    for snr in snrs:
        say('ber ci', nt, snr)
        b = frames.draw(7, snr, 0, B, nt, 16, dev='cuda')
        lt = b['l'].to(torch.int8)
        pipe = DetectPipeline(B, 2 * nt, 2 * nt, 16, R, policy='arpeggio', cycles=cycles, T0=0.1, T1=0.1, q=0.0,
                              key=5)
        _, be_fx = pipe.run(b['H'], b['y'], b['N0'], lt)
        from arpeggio_amd.detector import front_end
        o = front_end(b['H'], b['y'], b['N0'], 16, 1, want=('xg', 'Gf', 'hf'))
        T = np.full(cycles, np.float32(0.1), dtype=np.float32)
        s, _, _, _ = arpeggiofx.solve_float_device(o['Gf'], o['hf'], 2, R, T, 0.0, np.arange(B), 5, 'rand')
        _, be_f, _ = decode_select(b['H'], b['y'], o['xg'], s, 16, 2, True, ltrue=lt)
        torch.cuda.synchronize()
        a, c = be_f.cpu().numpy(), be_fx.cpu().numpy()
        r, lo, hi = orc.boot_ratio(a, c, B=2000)
        say('ber', nt, snr, a.sum() / (B * 4 * nt), c.sum() / (B * 4 * nt), r, lo, hi)
        assert lo <= 1.0 <= hi
