"""Two anneals per wave (arp_set_anneals_per_wave(2), frames of D <= 32): every anneal equals the one-anneal-per-wave
launch and the oracle bit for bit (spins, best energy change, flips, tests) - random and balanced starts, odd R (the
last pair's upper half idle), D below 32, partial Philox blocks, the learned start with and without the energy gate,
frame lists with a stream offset, the adaptive chain and the mixed-size persistent launch."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print('[case]', *a, flush=True)


@pytest.fixture
def apw():
    from arpeggio_amd import _lib
    L = _lib.load()

    def set_(n):
        _lib.check(L.arp_set_anneals_per_wave(n), 'set_anneals_per_wave')
    yield set_
    set_(0)


def random_problem(rs, B, D, scale=4096):
    A = rs.normal(size=(B, D + 4, D))
    G = np.einsum('bri,brj->bij', A, A)
    hv = rs.normal(size=(B, D)) * 3
    m = np.maximum(np.abs(G).reshape(B, -1).max(-1), np.abs(hv).max(-1))
    return (np.rint(G / m[:, None, None] * scale).astype(np.int32),
            np.rint(hv / m[:, None] * scale).astype(np.int32))


CASES = [  # D, R, B, L, start, q, key, inst0
    (32, 16, 40, 64, 'rand', 0, 0x1234567890ABCDEF, 0),
    (32, 5, 9, 30, 'bal', 300, 77, 2 ** 32 - 3),
    (20, 7, 11, 64, 'rand', 100, 5, 7),
    (1, 3, 4, 9, 'rand', 0, 1, 0),
    (31, 1, 5, 17, 'bal', 0, 2, 0),
    (17, 4, 6, 24, 'rand', 0, 3, 100),
]


@pytest.mark.parametrize('case', CASES, ids=lambda c: 'D%d_R%d_L%d_%s' % (c[0], c[1], c[3], c[4]))
def test_pair_equals_single_and_oracle(orc, apw, case):
    from arpeggio_amd import arpeggiofx
    D, R, B, L, start, q, key, inst0 = case
    say('pack', case)
    rs = np.random.default_rng(D * 100 + R)
    G, h = random_problem(rs, B, D)
    T = rs.integers(50, 3000, L).tolist()
    inst = [(inst0 + b) % 2 ** 32 for b in range(B)]
    apw(1)
    one = arpeggiofx.solve(G, h, 2, R, T, q, inst, key, start)
    apw(2)
    two = arpeggiofx.solve(G, h, 2, R, T, q, inst, key, start)
    for a, b in zip(one, two):
        assert (a == b).all()
    n = min(B, 6)
    os_, ode, ost = orc.arpeggio_fx(G[:n], h[:n], 2, R, T, q, inst[:n], key, start)
    assert (two[0][:n] == os_).all() and (two[1][:n] == ode).all() and (two[2][:n] == ost).all()


@pytest.mark.parametrize('gate', [True, False])
def test_pair_learned_start(orc, apw, gate):
    """arp_anneal_arpeggio_ws at D = 32 with per-spin thresholds: packed equals one per wave and the oracle."""
    from arpeggio_amd import _lib
    say('pack learned, gate', gate)
    D, R, B, L = 32, 6, 12, 40
    rs = np.random.default_rng(5)
    G, h = random_problem(rs, B, D)
    T = rs.integers(50, 3000, L).astype(np.int32)
    thr = rs.integers(0, 2 ** 32, (B, 2 * D), dtype=np.uint64).astype(np.uint32)
    thr[:, :5] = 2 ** 31
    key, inst = 0xBEEF12345, np.arange(B, dtype=np.int64) + 40
    Gt = torch.as_tensor(np.ascontiguousarray(G.transpose(0, 2, 1))).cuda()
    ht, Tt = torch.as_tensor(h).cuda(), torch.as_tensor(T).cuda()
    it = torch.as_tensor(inst.astype(np.int32)).cuda()
    tt = torch.as_tensor(thr.view(np.int32)).cuda()
    p = _lib.ptr
    outs = []
    for a in (1, 2):
        apw(a)
        s = torch.empty(B, R, 2 * D, dtype=torch.int8, device='cuda')
        de = torch.empty(B, R, dtype=torch.int64, device='cuda')
        st = torch.empty(B, R, 2, dtype=torch.int64, device='cuda')
        used = torch.empty(B, R, dtype=torch.int8, device='cuda')
        _lib.check(_lib.load().arp_anneal_arpeggio_ws(p(Gt), p(ht), p(Tt), L, 30, 2, R, D, B, p(it), key & 0xFFFFFFFF,
                                                      key >> 32, p(tt), 1 if gate else 0, p(s), p(de), p(st), p(used),
                                                      _lib.stream_ptr()), 'ws')
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in (s, de, st, used)])
    for a, b in zip(*outs):
        assert (a == b).all()
    os_, ode, ost, oused = orc.arpeggio_ws(G, h, 2, R, T, 30, inst, key, thr, gate=gate)
    s, de, st, used = outs[1]
    assert (s == os_).all() and (de == ode).all() and (st == ost).all() and (used == oused).all()


def test_pair_frame_list(apw):
    """arp_anneal_arpeggio_list (int16 Gram, a frame list, stream offset r0): packed equals one per wave."""
    from arpeggio_amd import _lib
    say('pack list')
    D, R, B, L, r0 = 32, 9, 30, 32, 4
    rs = np.random.default_rng(9)
    G, h = random_problem(rs, B, D)
    Gh = torch.as_tensor(G.astype(np.int16)).cuda()   # symmetric: its own transpose
    ht = torch.as_tensor(h).cuda()
    Tt = torch.as_tensor(rs.integers(50, 3000, L).astype(np.int32)).cuda()
    it = torch.arange(B, dtype=torch.int32, device='cuda')
    lst = torch.tensor([3, 17, 0, 29, 11, 5, 22], dtype=torch.int32, device='cuda')
    cnt = torch.tensor([5], dtype=torch.int32, device='cuda')   # the last two slots exit
    p = _lib.ptr
    outs = []
    for a in (1, 2):
        apw(a)
        s = torch.zeros(B, R, 2 * D, dtype=torch.int8, device='cuda')
        de = torch.zeros(B, R, dtype=torch.int64, device='cuda')
        st = torch.zeros(B, R, 2, dtype=torch.int64, device='cuda')
        _lib.check(_lib.load().arp_anneal_arpeggio_list(None, p(Gh), p(ht), p(Tt), L, 0, 2, R, D, B, p(it), 11, 12, 0,
                                                        None, 0, r0, p(lst), p(cnt), p(s), p(de), p(st), None,
                                                        _lib.stream_ptr()), 'list')
        torch.cuda.synchronize()
        outs.append((s, de, st))
    for a, b in zip(*outs):
        assert torch.equal(a, b)
    assert bool((outs[1][1][:5] != 0).any())


def test_pair_adaptive_chain_16_users(apw):
    """arp_detect_adaptive at 16 users (D = 32): decisions and bit errors equal with and without packing."""
    from arpeggio_amd import frames
    from arpeggio_amd.detector import detect_chain
    say('pack adaptive chain')
    b = frames.draw(71, 14.0, 0, 2000, 16, 16, dev='cuda')
    res = []
    for a in (1, 2):
        apw(a)
        lhat, be = detect_chain(b['H'], b['y'], b['N0'], 16, c=6.0, key=3, ltrue=b['l'])
        torch.cuda.synchronize()
        res.append((lhat.clone(), be.clone()))
    assert torch.equal(res[0][0], res[1][0]) and torch.equal(res[0][1], res[1][1])


def test_pair_mixed_batch(apw):
    """The persistent mixed-size launch with the 16-user classes packed: every anneal equals the unpacked launch."""
    from arpeggio_amd.batch import BatchedDetector, draw_mixed, prepare
    say('pack mixed batch')
    fr = prepare(draw_mixed(13, 600, snr_lo=10.0, snr_hi=22.0))
    spec = [dict(nt=f['nt'], M=f['M'], B=f['l'].shape[0], idx=f['idx']) for f in fr]
    res = []
    for a in (1, 2):
        apw(a)
        det = BatchedDetector(spec, R=7, passes=30, key=0x77)
        det.run(fr)
        torch.cuda.synchronize()
        res.append([(c['spins'].clone(), c['de'].clone(), c['st'].clone(), c['lhat'].clone()) for c in det.cls])
    for c1, c2 in zip(*res):
        assert all(torch.equal(x, y) for x, y in zip(c1, c2))
