"""GPU parity of the mixed-size persistent launch: every anneal equals the per-class arp_anneal_arpeggio launch (and
the C oracle) bit for bit, for several launch shapes and the autotuned one; the captured HIP graph of the whole step
replays to the eager results."""
import ctypes
import os
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))


def say(*a):
    print('[case]', *a, flush=True)


def per_class(det, c):
    from arpeggio_amd import _lib
    B, D, R = c['B'], c['D'], det.R
    s = torch.empty(B, R, 2 * D, dtype=torch.int8, device='cuda')
    de = torch.empty(B, R, dtype=torch.int64, device='cuda')
    st = torch.empty(B, R, 2, dtype=torch.int64, device='cuda')
    p = _lib.ptr
    _lib.check(_lib.load().arp_anneal_arpeggio(p(c['Gi']), p(c['hi']), p(det.T), det.L, det.q, 2, R, D, B, p(c['inst']),
                                               det.key & 0xFFFFFFFF, det.key >> 32, 0, 0, p(s), p(de), p(st),
                                               _lib.stream_ptr()), 'arpeggio')
    torch.cuda.synchronize()
    return s, de, st


@pytest.fixture(scope='module')
def mixed():
    from arpeggio_amd.batch import draw_mixed, prepare
    say('draw mixed')
    return prepare(draw_mixed(5, 300, snr_lo=10.0, snr_hi=22.0))


@pytest.mark.parametrize('shape', [(3, 4, 66048), (7, 16, 163840), (1, 1, 66048), (64, 8, 81920)])
def test_batched_equals_per_class(mixed, shape):
    from arpeggio_amd.batch import BatchedDetector
    say('batched', shape, [(f['nt'], f['M'], f['l'].shape[0]) for f in mixed])
    det = BatchedDetector([dict(nt=f['nt'], M=f['M'], B=f['l'].shape[0], idx=f['idx']) for f in mixed], R=6,
                          passes=32, key=0x1234ABCD5678, shape=shape)
    det.run(mixed)
    torch.cuda.synchronize()
    for c in det.cls:
        s, de, st = per_class(det, c)
        assert torch.equal(s, c['spins']) and torch.equal(de, c['de']) and torch.equal(st, c['st']), (c['nt'], c['M'])


def test_batched_vs_oracle(mixed, orc):
    from arpeggio_amd.batch import BatchedDetector
    say('batched vs oracle')
    det = BatchedDetector([dict(nt=f['nt'], M=f['M'], B=f['l'].shape[0], idx=f['idx']) for f in mixed], R=4,
                          passes=16, key=77)
    det.run(mixed)
    torch.cuda.synchronize()
    T = det.T.cpu().numpy()
    for c in det.cls:
        n = min(4, c['B'])
        G = c['Gi'][:n].cpu().numpy()
        h = c['hi'][:n].cpu().numpy()
        inst = c['inst'][:n].cpu().numpy().astype(np.int64) & 0xFFFFFFFF
        s, de, st = orc.arpeggio_fx(G, h, 2, 4, T, det.q, inst, 77, 'rand')
        assert (c['spins'][:n].cpu().numpy() == s).all() and (c['de'][:n].cpu().numpy() == de).all()
        assert (c['st'][:n].cpu().numpy() == st).all()


def test_graph_replay_and_autotune(mixed):
    from arpeggio_amd.autotune import tune_batched
    from arpeggio_amd.batch import BatchedDetector
    say('graph + autotune')
    det = BatchedDetector([dict(nt=f['nt'], M=f['M'], B=f['l'].shape[0], idx=f['idx']) for f in mixed], R=8,
                          passes=32, key=9)
    best_ms, shape, tried = tune_batched(det, reps=1)
    say('autotuned', shape, best_ms, len(tried))
    assert shape in [t[1] for t in tried]
    det.run(mixed)
    torch.cuda.synchronize()
    ref = [(c['spins'].clone(), c['de'].clone(), c['st'].clone(), c['lhat'].clone(), c['biterr'].clone())
           for c in det.cls]
    g = det.capture(mixed)
    for k in range(3):
        for c in det.cls:
            for t in ('spins', 'de', 'st', 'lhat', 'biterr'):
                c[t].zero_()
        g.replay()
        torch.cuda.synchronize()
        for c, r in zip(det.cls, ref):
            got = (c['spins'], c['de'], c['st'], c['lhat'], c['biterr'])
            assert all(torch.equal(a, b) for a, b in zip(got, r)), ('replay', k, c['nt'], c['M'])
        det.run(mixed)             # eager work between replays
        torch.cuda.synchronize()
    errs = det.bit_errors()
    say('bit errors', errs)
    assert errs >= 0


def test_batched_learned_start(mixed, orc):
    """Classes with a learned start (the trained 32- and 64-user networks) in the persistent launch: every anneal
    equals arp_anneal_arpeggio_ws on the class's own integers and thresholds, and the oracle's arpeggio_ws on a few
    frames; the random-start classes of the same launch still equal arp_anneal_arpeggio."""
    from arpeggio_amd import _lib
    from arpeggio_amd.batch import BatchedDetector
    from arpeggio_amd.learned import DEFAULT_WEIGHTS, WEIGHTS_64, WarmStartMLP
    models = {32: WarmStartMLP.load(DEFAULT_WEIGHTS), 64: WarmStartMLP.load(WEIGHTS_64)}
    for gate in (True, False):
        say('batched learned, gate', gate)
        det = BatchedDetector([dict(nt=f['nt'], M=f['M'], B=f['l'].shape[0], idx=f['idx']) for f in mixed], R=4,
                              passes=32, key=0xABCDEF12345, models=models, gate=gate)
        det.run(mixed)
        torch.cuda.synchronize()
        T = det.T.cpu().numpy()
        p = _lib.ptr
        for c in det.cls:
            if 'thr' not in c:
                s, de, st = per_class(det, c)
                assert torch.equal(s, c['spins']) and torch.equal(de, c['de']) and torch.equal(st, c['st'])
                continue
            B, D, R = c['B'], c['D'], det.R
            s = torch.empty(B, R, 2 * D, dtype=torch.int8, device='cuda')
            de = torch.empty(B, R, dtype=torch.int64, device='cuda')
            st = torch.empty(B, R, 2, dtype=torch.int64, device='cuda')
            used = torch.empty(B, R, dtype=torch.int8, device='cuda')
            _lib.check(_lib.load().arp_anneal_arpeggio_ws(p(c['Gi']), p(c['hi']), p(det.T), det.L, det.q, 2, R, D, B,
                                                          p(c['inst']), det.key & 0xFFFFFFFF, det.key >> 32,
                                                          p(c['thr']), 1 if gate else 0, p(s), p(de), p(st), p(used),
                                                          _lib.stream_ptr()), 'ws')
            torch.cuda.synchronize()
            assert torch.equal(s, c['spins']) and torch.equal(de, c['de']) and torch.equal(st, c['st']), c['nt']
            assert torch.equal(used, c['used']), c['nt']
            assert int(c['used'].sum()) > 0 if gate else bool((c['used'] == 1).all())
            n = min(3, B)
            os_, ode, ost, oused = orc.arpeggio_ws(c['Gi'][:n].cpu().numpy(), c['hi'][:n].cpu().numpy(), 2, R, T,
                                                   det.q, c['inst'][:n].cpu().numpy().astype(np.int64) & 0xFFFFFFFF,
                                                   det.key, c['thr'][:n].cpu().numpy().view(np.uint32), gate=gate)
            assert (c['spins'][:n].cpu().numpy() == os_).all() and (c['de'][:n].cpu().numpy() == ode).all()
            assert (c['used'][:n].cpu().numpy() == oused).all()
    with pytest.raises(ValueError):
        det.run_c(mixed)


def test_whole_chain_c_entry_equals_python_step():
    """arp_detect_batched (one C call: front ends, persistent anneal, decodes) gives the Python step's decisions and
    bit errors; arp_tune_batched returns one of its candidate shapes; a HIP graph of the C call replays equal."""
    from arpeggio_amd.batch import BatchedDetector, draw_mixed, prepare
    fr = prepare(draw_mixed(21, 3000, snr_lo=10.0, snr_hi=22.0))
    det = BatchedDetector([dict(nt=f['nt'], M=f['M'], B=f['l'].shape[0], idx=f['idx']) for f in fr], R=4, passes=32,
                          key=0x5151)
    det.run(fr)
    det.run_c(fr)
    torch.cuda.synchronize()
    for c in det.cls:
        assert (c['lhat_c'] == c['lhat']).all() and (c['biterr_c'] == c['biterr']).all()
    shape, ms = det.tune_c(fr)
    assert shape[1] in (4, 8, 16) and shape[0] > 0 and shape[3] in (1, 2) and ms > 0
    det.run(fr)
    torch.cuda.synchronize()
    want = [c['lhat'].clone() for c in det.cls]
    g = torch.cuda.CUDAGraph()
    det.run_c(fr)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        det.run_c(fr)
    for _ in range(3):
        for c in det.cls:
            c['lhat_c'].fill_(0)
        g.replay()
        torch.cuda.synchronize()
        for c, w in zip(det.cls, want):
            assert (c['lhat_c'] == w).all()
