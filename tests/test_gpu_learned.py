"""GPU parity of the learned warm start: BF16 MFMA GEMM (all stride forms), forward/backward gradients against the
numpy oracle, Adam, start thresholds, the warm-start anneal bit-exact against the C oracle, and a training smoke run."""
import os
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))


def say(*a):
    print('[case]', *a, flush=True)


def bf(x):
    import mlp_oracle as mo
    return mo.bf16(np.asarray(x, dtype=np.float32))


@pytest.mark.parametrize('M,N,K,ta,tb', [(64, 64, 32, 0, 0), (100, 70, 96, 0, 0), (33, 128, 288, 0, 1),
                                         (288, 256, 200, 1, 0), (129, 17, 64, 1, 1)])
def test_gemm_bf16(M, N, K, ta, tb):
    from arpeggio_amd.learned import WarmStartMLP
    say('gemm', M, N, K, ta, tb)
    rs = np.random.default_rng(M + N + K)
    A = bf(rs.normal(size=(K, M) if ta else (M, K)))
    B = bf(rs.normal(size=(N, K) if tb else (K, N)))
    bias = rs.normal(size=N).astype(np.float32)
    At = torch.as_tensor(A).to(torch.bfloat16).cuda()
    Bt = torch.as_tensor(B).to(torch.bfloat16).cuda()
    sam, sak = (1, M) if ta else (K, 1)
    sbk, sbn = (1, K) if tb else (N, 1)
    m = WarmStartMLP(8)
    Cf = torch.empty(M, N, device='cuda')
    Cpre = torch.empty(M, N, device='cuda')
    Cb = torch.empty(M, N, dtype=torch.bfloat16, device='cuda')
    m._gemm(At, sam, sak, Bt, sbk, sbn, M, N, K, torch.as_tensor(bias).cuda(), 1, Cf, Cpre, Cb)
    torch.cuda.synchronize()
    Am = (A.T if ta else A).astype(np.float64)
    Bm = (B.T if tb else B).astype(np.float64)
    ref = Am @ Bm + bias
    tol = 1e-5 * (np.abs(Am) @ np.abs(Bm)) + 1e-6
    assert (np.abs(Cpre.cpu().numpy() - ref) <= tol).all()
    import mlp_oracle as mo
    assert np.allclose(Cf.cpu().numpy(), mo.gelu(ref), atol=1e-4, rtol=1e-4)
    assert np.allclose(Cb.float().cpu().numpy(), mo.bf16(mo.gelu(ref).astype(np.float32)), atol=2e-2, rtol=1e-2)


def test_gradients_vs_oracle():
    import mlp_oracle as mo
    from arpeggio_amd.learned import WarmStartMLP
    say('gradients')
    rs = np.random.default_rng(11)
    m = WarmStartMLP(16, seed=3)
    for k in m.params:
        m.params[k] += torch.as_tensor(rs.normal(size=tuple(m.params[k].shape)) * 0.05, dtype=torch.float32).cuda()
    m.wb = {k: m.params[k].to(torch.bfloat16) for k in ('W1', 'W2', 'W3')}
    M = 200
    X = bf(rs.normal(size=(M, m.F)))
    Y = np.where(rs.random((M, m.N)) < 0.4, 1, -1).astype(np.int8)
    loss, g = m.gradients(torch.as_tensor(X).to(torch.bfloat16).cuda(), torch.as_tensor(Y).cuda())
    torch.cuda.synchronize()
    p = {k: v.cpu().numpy() for k, v in m.params.items()}
    lo, go = mo.loss_and_grads(p, X, Y)
    assert abs(float(loss) - lo) <= 1e-3 * abs(lo)
    for k in p:
        gg, oo = g[k].cpu().numpy(), go[k]
        err = np.linalg.norm(gg - oo) / np.linalg.norm(oo)
        say('grad', k, err)
        assert err < 2e-2, k


def test_adam_and_thresholds():
    import mlp_oracle as mo
    from arpeggio_amd import _lib
    from arpeggio_amd.learned import WarmStartMLP
    say('adam')
    rs = np.random.default_rng(2)
    m = WarmStartMLP(8, seed=1)
    w0 = {k: v.cpu().numpy().astype(np.float64) for k, v in m.params.items()}
    g = {k: torch.as_tensor(rs.normal(size=tuple(v.shape)), dtype=torch.float32).cuda() for k, v in m.params.items()}
    m.adam_step(g, lr=1e-2)
    m.adam_step(g, lr=1e-2)
    torch.cuda.synchronize()
    for k in w0:
        w, mm, vv = w0[k], 0.0, 0.0
        for step in (1, 2):
            w, mm, vv = mo.adam(w, g[k].cpu().numpy().astype(np.float64), mm, vv, step, 1e-2)
        assert np.allclose(m.params[k].cpu().numpy(), w, atol=1e-6, rtol=1e-5), k
        assert (m.wb.get(k, m.params[k].to(torch.bfloat16)).float().cpu().numpy()
                == mo.bf16(m.params[k].cpu().numpy())).all()
    Z = torch.as_tensor(rs.normal(size=(9, 13)) * 3, dtype=torch.float32).cuda()
    thr = torch.empty(9, 13, dtype=torch.int32, device='cuda')
    _lib.check(_lib.load().arp_start_thresholds(_lib.ptr(Z), 13, 9, 13, 1.0, _lib.ptr(thr), _lib.stream_ptr()), 't')
    torch.cuda.synchronize()
    want = np.clip(np.rint(1 / (1 + np.exp(-Z.cpu().numpy().astype(np.float64))) * 2.0 ** 32), 1, 2 ** 32 - 1)
    got = thr.cpu().numpy().view(np.uint32).astype(np.float64)
    assert np.abs(got - want).max() <= 1


@pytest.mark.parametrize('D,P,gate', [(64, 2, True), (64, 2, False), (32, 4, True), (100, 2, True), (128, 2, True),
                                      (17, 6, True)])
def test_arpeggio_ws_vs_oracle(orc, D, P, gate):
    from arpeggio_amd import _lib
    from test_gpu_anneal import random_problem
    say('arpeggio ws', D, P, gate)
    rs = np.random.default_rng(D + P)
    B, R, L = 5, 7, 40
    G, h = random_problem(rs, B, D)
    T = rs.integers(100, 2500, L).astype(np.int32)
    thr = rs.integers(1, 2 ** 32 - 1, (B, P * D), dtype=np.uint64).astype(np.uint32)
    thr[0] = 2 ** 31
    key, inst = 0xC0FFEE1234, np.arange(B) + 77
    Gt = torch.as_tensor(G.transpose(0, 2, 1).copy()).cuda()
    ht = torch.as_tensor(h).cuda()
    Tt = torch.as_tensor(T).cuda()
    it = torch.as_tensor(inst.astype(np.int32)).cuda()
    tt = torch.as_tensor(thr.view(np.int32)).cuda()
    s = torch.empty(B, R, P * D, dtype=torch.int8, device='cuda')
    de = torch.empty(B, R, dtype=torch.int64, device='cuda')
    st = torch.empty(B, R, 2, dtype=torch.int64, device='cuda')
    used = torch.empty(B, R, dtype=torch.int8, device='cuda')
    p = _lib.ptr
    _lib.check(_lib.load().arp_anneal_arpeggio_ws(p(Gt), p(ht), p(Tt), L, 30, P, R, D, B, p(it), key & 0xFFFFFFFF,
                                                  key >> 32, p(tt), 1 if gate else 0, p(s), p(de), p(st), p(used),
                                                  _lib.stream_ptr()), 'ws')
    torch.cuda.synchronize()
    os_, ode, ost, oused = orc.arpeggio_ws(G, h, P, R, T, 30, inst, key, thr, gate=gate)
    assert (s.cpu().numpy() == os_).all() and (de.cpu().numpy() == ode).all() and (st.cpu().numpy() == ost).all()
    assert (used.cpu().numpy() == oused).all()
    s0, de0, _ = orc.arpeggio_fx(G, h, P, R, T, 30, inst, key, 'rand')
    assert (os_[0] == s0[0]).all() and (ode[0] == de0[0]).all()      # thr = 2^31 is the random start


def test_features_and_training_smoke():
    from arpeggio_amd import frames
    from arpeggio_amd.learned import WarmStartMLP, feat_ld, label_batch, train
    say('features/train')
    b = frames.draw(21, 24.0, 0, 2048, 32, 16, dev='cuda')
    feat, lab, be = label_batch(b['H'], b['y'], b['N0'], b['l'].to(torch.int8), passes=64, R=4)
    torch.cuda.synchronize()
    assert feat.shape == (2048, feat_ld(64)) and lab.shape == (2048, 128)
    f = feat.float().cpu().numpy()
    assert np.isfinite(f).all() and (f[:, 4 * 64 + 1:] == 0).all()
    levels = bf(np.array([-1, -1 / 3, 1 / 3, 1], dtype=np.float32))                     # bf16(x_g / 3)
    assert np.isin(f[:, 2 * 64:3 * 64], levels).all()
    m = WarmStartMLP(64, seed=0)
    losses = train(m, feat, lab, epochs=6, batch=512, lr=3e-3)
    say('losses', losses)
    assert losses[-1] < losses[0] * 0.8


def test_trained_start_gains_over_random_start():
    """What the trained network buys: on held-out frames (the reference's draw recipe, seed 7) the learned start
    reaches a lower BER than the random start at a short pass budget, with fewer flips, and the energy gate keeps it
    for almost every anneal."""
    from arpeggio_amd import frames
    from arpeggio_amd.detector import DetectPipeline
    from arpeggio_amd.learned import DEFAULT_WEIGHTS, WarmStartMLP
    from arpeggio_amd.stats import boot_ratio
    model = WarmStartMLP.load(DEFAULT_WEIGHTS)
    B = 16384
    b = frames.draw(7, 23.7, 5, B, 32, 16, dev='cuda')
    lt = b['l'].to(torch.int8).contiguous()
    out = {}
    for start in ('rand', 'learned'):
        pipe = DetectPipeline(B, 64, 64, 16, 16, cycles=8, start=start, key=31,
                              model=model if start == 'learned' else None, inv_temp=1.0)
        _, be = pipe.run(b['H'], b['y'], b['N0'], lt)
        torch.cuda.synchronize()
        out[start] = (be.cpu().numpy().astype(np.int64), float(pipe.st[..., 0].double().mean()),
                      float(pipe.used.double().mean()) if start == 'learned' else 1.0)
    r, lo, hi = boot_ratio(out['learned'][0], out['rand'][0])
    say('BER learned %.3e rand %.3e ratio %.3f [%.3f, %.3f], flips %.1f vs %.1f, kept %.4f' % (
        out['learned'][0].sum() / (B * 128), out['rand'][0].sum() / (B * 128), r, lo, hi, out['learned'][1],
        out['rand'][1], out['learned'][2]))
    assert hi < 1.0
    assert out['learned'][1] < 0.6 * out['rand'][1]
    assert out['learned'][2] > 0.9


@pytest.mark.parametrize('M', [1, 63, 64, 1000])
@pytest.mark.parametrize('net', ['32u', '64u'])
def test_fused_inference_equals_gemm_path(M, net):
    """The fused features -> thresholds kernel gives the same bits as the three GEMMs + threshold kernel: the 32-user
    network (F = 288, N = 128) and the 64-user one (F = 544, N = 256: the tile's activations past 64 KiB of LDS)."""
    from arpeggio_amd.learned import DEFAULT_WEIGHTS, WEIGHTS_64, WarmStartMLP
    say('fused', M, net)
    model = WarmStartMLP.load(DEFAULT_WEIGHTS if net == '32u' else WEIGHTS_64)
    rs = np.random.default_rng(M)
    X = torch.as_tensor(bf(rs.normal(size=(M, model.F)) * 0.7)).to(torch.bfloat16).cuda()
    X[:, 4 * model.D + 1:] = 0
    a = model.thresholds(X, 0.35, fused=True)
    b = model.thresholds(X, 0.35, fused=False)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_trained_start_gains_at_64_users():
    """The 64-user network (Kronecker rho = 0.5, D = 128): on held-out frames at 28 dB the
    learned start at 64 passes beats the random start at 128 passes (fewer than half the passes), and at 32 passes its
    BER is less than half the random start's."""
    from arpeggio_amd import frames
    from arpeggio_amd.detector import DetectPipeline
    from arpeggio_amd.learned import WEIGHTS_64, WarmStartMLP
    from arpeggio_amd.stats import boot_ratio
    model = WarmStartMLP.load(WEIGHTS_64)
    assert model.D == 128
    B = 8192
    b = frames.draw_device(7042, 28.0, 28.0, B, 64, 16, rho_rx=0.5, rho_tx=0.5)
    lt = b['l'].to(torch.int8).contiguous()
    be = {}
    for start, L in (('rand', 32), ('rand', 128), ('learned', 32), ('learned', 64)):
        pipe = DetectPipeline(B, 128, 128, 16, 32, cycles=L, T0=0.08, T1=0.03, start=start, key=41,
                              model=model if start == 'learned' else None)
        _, e = pipe.run(b['H'], b['y'], b['N0'], lt)
        torch.cuda.synchronize()
        be[(start, L)] = e.cpu().numpy().astype(np.int64)
        del pipe
    r32, _, hi32 = boot_ratio(be[('learned', 32)], be[('rand', 32)])
    r64, _, hi64 = boot_ratio(be[('learned', 64)], be[('rand', 128)])
    say('64 users: learned/rand at 32 passes %.3f (hi %.3f); learned 64 / rand 128 %.3f (hi %.3f)' % (r32, hi32, r64,
                                                                                                     hi64))
    assert hi32 < 0.5
    assert hi64 < 1.0
