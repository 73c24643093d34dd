#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ by running the reference's own functions IN THE BUILD CONTAINER
(where the reference repository is mounted read-only at /root/reference). The fixtures are data only: inputs and
the reference's outputs. Nothing here runs on the GPU box, and no test imports the reference.

Reference functions run (paths relative to the reference repository):
  arpeggio/sim/local_code_a59/arpeggio_ref59.py  anneal, philox, bal_pat  (loaded as arpeggio/sim/local_code_a513/
                                                 gate_a513.py:26-31 loads it: the source up to its command block)
  arpeggio/mimo/chordfx.py                        philox, KAT, to_fixed, chord_fx
  arpeggio/mimo/mimo.py                           mmse, _mu, mmse_hard, quantize_levels, ising_delta_at, ising,
                                                  spins_to_levels_delta, select, bit_errors, geom
  arpeggio/sim/x3model/chord_ref.py               compress
  arpeggio/sim/arpx.py                            draw_seed, draw  (the two function definitions only: the module
                                                  imports numba through tree.py, which is absent here)
  arpeggio/sim/a56an.py                           boot_ratio, mcnemar
Outputs: golden.npz (arrays) and golden.json (case index, scalars, and the B200 hash check).
Usage: python tests/golden/make_golden.py [--quick]"""
import ast
import hashlib
import json
import math
import os
import random
import sys
import time

import numpy as np
import torch

REF = '/root/reference'
sys.path.insert(0, os.path.join(REF, 'arpeggio', 'mimo'))
sys.path.insert(0, os.path.join(REF, 'arpeggio', 'sim', 'x3model'))
import chordfx  # noqa: E402
import mimo  # noqa: E402
import chord_ref  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
torch.set_num_threads(8)


def load_arpeggio_ref():
    p = os.path.join(REF, 'arpeggio', 'sim', 'local_code_a59', 'arpeggio_ref59.py')
    src = open(p).read()
    cut = src.index("\nif sys.argv[1] == 'gen':")
    ns = {'__name__': 'arpeggio_ref59'}
    exec(compile(src[:cut], p, 'exec'), ns)
    return ns


def load_functions(path, names, extra):
    src = open(path).read()
    tree = ast.parse(src)
    keep = [n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name in names]
    mod = ast.Module(body=keep, type_ignores=[])
    ns = dict(extra)
    exec(compile(mod, path, 'exec'), ns)
    return ns


ARX = load_functions(os.path.join(REF, 'arpeggio', 'sim', 'arpx.py'), {'draw_seed', 'draw'},
                     dict(torch=torch, math=math, mimo=mimo, np=np))
A56 = load_functions(os.path.join(REF, 'arpeggio', 'sim', 'a56an.py'), {'boot_ratio', 'mcnemar'},
                     dict(np=np, math=math))
A59 = load_arpeggio_ref()
arrays, index = {}, {}


def put(key, a):
    arrays[key] = np.ascontiguousarray(a)


def sha(t):
    return hashlib.sha256(t.detach().cpu().contiguous().numpy().tobytes()).hexdigest()


def front(seed, snr, bi, B, nt, M, K):
    b = ARX['draw'](seed, snr, bi, B, nt, M, dev='cpu')
    z, Ai, _ = mimo.mmse(b)
    lm = mimo.quantize_levels(z / mimo._mu(b, z), M)
    I = mimo.ising_delta_at(b, lm, K)
    return b, z, Ai, lm, I


# ---------------------------------------------------------------- 1. Philox
def g_philox():
    words = [list(c) + list(k) for c, k, _ in chordfx.KAT]
    want = [list(w) for _, _, w in chordfx.KAT]
    rng = random.Random(2024)
    extra = [[rng.getrandbits(32) for _ in range(6)] for _ in range(64)]
    got = []
    for w in extra:
        got.append(list(A59['philox'](*w)))
    put('philox/in', np.array(words + extra, dtype=np.uint32))
    put('philox/out', np.array(want + got, dtype=np.uint32))
    index['philox'] = dict(n_kat=len(words), n=len(words) + len(extra))


# ---------------------------------------------------------------- 2. Arpeggio fixed point (arpeggio_ref59.anneal)
def bits_to_spins(bs, P, D):
    s = np.zeros(P * D, dtype=np.int8)
    for p in range(P):
        for jj, ch in enumerate(bs[p]):
            j = D - 1 - jj
            s[p * D + j] = 1 if ch == '1' else -1
    return s


def arp_case(name, G, h, P, T, q, inst, key, start, mode, R, prob):
    B, D = h.shape
    A59['SEED'] = key
    rec = dict(P=P, q=int(q), key='%x' % key, start=start, mode=mode, R=R, B=B, D=D, L=len(T), prob=prob)
    best, flips, tests, n1, init, nacc = [], [], [], [], [], []
    spins, accs, tks = [], [], []
    for b in range(B):
        Gl = [[int(v) for v in row] for row in G[b]]
        hl = [int(v) for v in h[b]]
        for r in range(R):
            out = A59['anneal'](Gl, hl, P, [int(t) for t in T], int(q), r, int(inst[b]), start, mode)
            best.append(out['best']); flips.append(len(out['acc'])); tests.append(sum(out['tests']))
            n1.append(out['n1']); init.append(out['init'])
            spins.append(bits_to_spins(out['bs'], P, D))
            accs.append(np.array(out['acc'], dtype=np.int32).reshape(-1, 2))
            nacc.append(len(out['acc']))
            tks.append(np.array(out['tests'], dtype=np.int32))
    if 'arpprob/%s/G' % prob not in arrays:
        put('arpprob/%s/G' % prob, np.asarray(G, dtype=np.int32))
        put('arpprob/%s/h' % prob, np.asarray(h, dtype=np.int32))
    put('arp/%s/T' % name, np.asarray(T, dtype=np.int32))
    put('arp/%s/inst' % name, np.asarray(inst, dtype=np.int64))
    put('arp/%s/best' % name, np.array(best, dtype=np.int64).reshape(B, R))
    put('arp/%s/flips' % name, np.array(flips, dtype=np.int64).reshape(B, R))
    put('arp/%s/tests' % name, np.array(tests, dtype=np.int64).reshape(B, R))
    put('arp/%s/n1' % name, np.array(n1, dtype=np.int64).reshape(B, R))
    put('arp/%s/init' % name, np.array(init, dtype=np.int64).reshape(B, R))
    put('arp/%s/spins' % name, np.stack(spins).reshape(B, R, P * D))
    put('arp/%s/acc' % name, np.concatenate(accs, 0))
    put('arp/%s/nacc' % name, np.array(nacc, dtype=np.int64))
    put('arp/%s/tests_k' % name, np.stack(tks).reshape(B, R, len(T)))
    index.setdefault('arp', {})[name] = rec
    print('arp', name, 'flips', sum(flips), flush=True)


def compress_problem(I, K):
    F = chordfx.to_fixed(I, 128 * 2 * K, 0.3, 0.3, 0.4, 0.0)
    G, h = chord_ref.compress(F, K)
    return G.numpy(), h.numpy()


def sched(S, T0, T1, q, K):
    Q = chordfx.QSCALE
    Tsw = [int(round(v * Q)) for v in mimo.geom(T0, T1, S, 'cpu').double().tolist()]
    return [t for t in Tsw for _ in range(2 * K)], int(round(q * Q))


def g_arpeggio(quick):
    rnd = random.Random(513)
    keys = (0x1234567890ABCDEF, rnd.getrandbits(64))
    # tiny problem: exhaustive-friendly, arbitrary symmetric integers with a diagonal
    rs = np.random.default_rng(7)
    for D, P in ((4, 2), (5, 4), (3, 6)):
        A = rs.integers(-3000, 3000, (2, D, D))
        G = (A + A.transpose(0, 2, 1)) // 2
        h = rs.integers(-4000, 4000, (2, D))
        T = [int(v) for v in rs.integers(200, 3000, 24)]
        for start in ('rand', 'bal'):
            for mode in ('arpeggio', 'par'):
                arp_case('tiny_D%d_P%d_%s_%s' % (D, P, start, mode), G, h, P, T, 300, [5, 2 ** 32 - 1], keys[1], start,
                         mode, 3, 'tiny_D%d_P%d' % (D, P))
    shapes = [('x16', 16, 20.0, 6), ('x32', 32, 22.0, 3), ('x64', 64, 22.0, 2)]
    scheds = {'pick': (16, 0.1, 0.1, 0.0), 'hot': (8, 1.2, 0.1, 0.2)}
    for tag, nt, snr, B in shapes:
        Ks = (1, 2, 3) if tag == 'x16' else (1,) if quick else ((1, 2) if tag == 'x32' else (1,))
        for K in Ks:
            _, _, _, _, I = front(9600, snr, 0, B, nt, 16, K)
            G, h = compress_problem(I, K)
            nb = 2 if tag != 'x16' else 3
            G, h = G[:nb], h[:nb]
            P = 2 * K
            for sname, (S, T0, T1, q) in scheds.items():
                T, qi = sched(S if tag != 'x64' else S // 2, T0, T1, q, K)
                for ki, key in enumerate(keys):
                    for start in ('rand', 'bal'):
                        if quick and (ki or start == 'bal') and tag != 'x16':
                            continue
                        inst = [1000 * K + bb for bb in range(G.shape[0])]
                        arp_case('%s_K%d_%s_k%d_%s' % (tag, K, sname, ki, start), G, h, P, T, qi, inst, key, start,
                                 'arpeggio', 2, '%s_K%d' % (tag, K))
            T, qi = sched(8, 0.5, 0.1, 0.1, K)
            arp_case('%s_K%d_par' % (tag, K), G, h, P, T, qi, [7 + bb for bb in range(G.shape[0])], keys[0], 'rand',
                     'par', 2, '%s_K%d' % (tag, K))


# ---------------------------------------------------------------- 3. Chord fixed point (chordfx.chord_fx)
def chord_case(name, F, R, seed, P, inst0):
    t0 = time.time()
    cs, ce = chordfx.chord_fx(F, R, seed, P=P, inst0=inst0)
    bs, be = chordfx.chord_fx(F, R, seed, P=P, inst0=inst0, mode='best')
    for k in ('J', 'h', 'h0', 'T', 'c'):
        put('chord/%s/%s' % (name, k), F[k].numpy())
    put('chord/%s/chains_s' % name, cs.numpy())
    put('chord/%s/chains_e' % name, ce.numpy())
    put('chord/%s/best_s' % name, bs.numpy())
    put('chord/%s/best_e' % name, be.numpy())
    index.setdefault('chord', {})[name] = dict(R=R, seed=seed, P=P, inst0=inst0, q=int(F['q']), S=int(F['S']))
    print('chord', name, '%.1fs' % (time.time() - t0), flush=True)


def g_chord(quick):
    def ri(nt, M, snr, B):
        b = ARX['draw'](9700, snr, 0, B, nt, M, dev='cpu')
        z, _, _ = mimo.mmse(b)
        return mimo.ising(b, z)

    def ddi(nt, snr, B, K):
        return front(9700, snr, 1, B, nt, 16, K)[4]

    cases = [('q16_RI', lambda: ri(16, 4, 10.0, 6), 4, 64, 0.7, 0.5, 0.6, 0.0, 1, 7, 0),
             ('q16_RI_kappa', lambda: ri(16, 4, 10.0, 6), 4, 64, 2.0, 0.05, 0.5, 0.3, 1, 8, 1000),
             ('x16_K1', lambda: ddi(16, 20.0, 6, 1), 4, 128, 0.3, 0.3, 0.4, 0.0, 2, 9, 0),
             ('x32_K1_kappa', lambda: ddi(32, 22.0, 3, 1), 3, 96, 0.3, 0.2, 0.4, 0.3, 2, 10, 77),
             ('x16_K3', lambda: ddi(16, 18.0, 3, 3), 2, 48, 0.3, 0.3, 0.4, 0.0, 6, 11, 5),
             ('x64_K1', lambda: ddi(64, 24.0, 2, 1), 2, 24, 0.3, 0.3, 0.4, 0.0, 2, 12, 0),
             ('q16_RI_R5_seed64', lambda: ri(16, 4, 12.0, 3), 5, 32, 0.7, 0.5, 0.6, 0.0, 1, 2 ** 40 + 3, 123456)]
    for name, mk, R, S, t0, t1, q, kappa, P, seed, inst0 in cases:
        I = mk()
        F = chordfx.to_fixed(I, S, t0, t1, q, kappa)
        chord_case(name, F, R, seed, P, inst0)


# ---------------------------------------------------------------- 4. B200 kernel hashes (kernel_b200_v2.json)
def g_b200(quick):
    ref = json.load(open(os.path.join(REF, 'data', 'ber', 'equivalence', 'kernel_b200_v2.json')))
    SH = {'x16': dict(nt=16, M=16), 'q32': dict(nt=32, M=4), 'q16': dict(nt=16, M=4)}
    res = {}
    todo = (('x16', 20.0, 1), ('x16', 20.0, 2), ('x16', 20.0, 3), ('q32', 12.0, 1), ('q16', 12.0, 1))
    if quick:
        todo = todo[:1]
    for shape, snr, K in todo:
        s = SH[shape]
        b = ARX['draw'](9400, snr, K, 64, s['nt'], s['M'], dev='cpu')
        z, _, _ = mimo.mmse(b)
        lm = mimo.quantize_levels(z / mimo._mu(b, z), s['M'])
        I = mimo.ising_delta_at(b, lm, K)
        key = '%s_K%d' % (shape, K)
        inp = sha(I['J']) + sha(I['h0'])
        match = inp == ref['inputs'][key]
        F = chordfx.to_fixed(I, 128 * 2 * K, 0.3, 0.3, 0.4, 0.0)
        G, h = chord_ref.compress(F, K)
        put('b200/%s/G' % key, G.numpy().astype(np.int16))
        put('b200/%s/h' % key, h.numpy().astype(np.int32))
        put('b200/%s/T' % key, F['T'].numpy())
        put('b200/%s/c' % key, F['c'].numpy())
        r = dict(inputs_match=bool(match), K=K, q=int(F['q']), S=int(F['S']), cases={})
        for P in (2 * K, 1):
            cname = 'chord_%s_P%d' % (key, P)
            t0 = time.time()
            ss, ee = chordfx.chord_fx(F, 12, 123456789 + K, P=P)
            got = [sha(ss), sha(ee)]
            want = ref['cases'][cname]
            r['cases'][cname] = dict(R=12, seed=123456789 + K, P=P, b200=want, chord_fx=got, chord_fx_equals_b200=got == want)
            print('b200', cname, 'inputs', match, 'chord_fx==B200', got == want, '%.1fs' % (time.time() - t0), flush=True)
        res[key] = r
    index['b200'] = res


# ---------------------------------------------------------------- 5. Front end (float) and decisions
def g_front():
    out = {}
    for tag, nt, M, snr, K in (('x16', 16, 16, 18.0, 1), ('x32', 32, 16, 20.0, 1), ('q16', 16, 4, 8.0, 1), ('x16K2', 16, 16, 16.0, 2)):
        B = 8
        b = ARX['draw'](9800, snr, 3, B, nt, M, dev='cpu')
        z, Ai, el2 = mimo.mmse(b)
        mu = mimo._mu(b, z)
        lm_solve = mimo.quantize_levels(z / mu, M)
        lm_hard = mimo.mmse_hard(b, z, Ai)
        I = mimo.ising_delta_at(b, lm_hard, K)
        F = chordfx.to_fixed(I, 64, 0.3, 0.2, 0.4, 0.3)
        G, h = chord_ref.compress(F, K)
        g = torch.Generator(device='cpu'); g.manual_seed(5)
        S = torch.where(torch.rand(B, 5, 2 * K * 2 * nt, generator=g) < 0.5, -1.0, 1.0)
        Lc = mimo.spins_to_levels_delta(S, I, M)
        lh, allc, d = mimo.select(b, Lc, lm_hard)
        e, nb = mimo.bit_errors(b, lh)
        E0 = mimo.energy0(S, I['h0'], I['J'])
        p = 'front/%s/' % tag
        for k, v in dict(H=b['H'], y=b['y'], l=b['l'], z=z, Ainvdiag=torch.diagonal(Ai, dim1=-2, dim2=-1), mu=mu,
                         lm_solve=lm_solve, lm_hard=lm_hard, h0=I['h0'], Jrow0=I['J'][:, 0, :], scale=I['scale'],
                         hint=F['h'], T=F['T'], c=F['c'],
                         G=G, hG=h, S=S, Lc=Lc, lh=lh, d=d, E0=E0).items():
            put(p + k, v.numpy())
        out[tag] = dict(nt=nt, M=M, snr=snr, K=K, N0=b['N0'], el2=el2, q=int(F['q']), bit_errors=e, bits=nb,
                        seed=9800, bi=3, draw_seed=ARX['draw_seed'](9800, snr, 3))
    # to_fixed rounding ties: values exactly halfway between integers after scaling
    Jt = torch.tensor([[[0.0, 0.5 / 4096, 1.5 / 4096], [0.5 / 4096, 0.0, -2.5 / 4096], [1.5 / 4096, -2.5 / 4096, 0.0]]])
    ht = torch.tensor([[0.5 / 4096, -0.5 / 4096, 3.5 / 4096]])
    Ft = chordfx.to_fixed(dict(J=Jt, h=ht, h0=ht), 10, 0.7, 0.05, 0.45, 0.3)
    for k in ('J', 'h', 'T', 'c'):
        put('tofixed/' + k, Ft[k].numpy())
    put('tofixed/Jin', Jt.numpy()); put('tofixed/hin', ht.numpy())
    out['tofixed'] = dict(q=int(Ft['q']), S=10, t0=0.7, t1=0.05, qf=0.45, kappa=0.3)
    put('geom/a', mimo.geom(0.3, 0.2, 256, 'cpu').numpy())
    put('geom/b', mimo.geom(1.2, 0.1, 7, 'cpu').numpy())
    index['front'] = out


def g_stats():
    rs = np.random.default_rng(3)
    res = []
    for i in range(3):
        a = (rs.random(5000) < 0.01 * (i + 1)).astype(np.int64) * rs.integers(1, 3, 5000)
        b = (rs.random(5000) < 0.015).astype(np.int64) * rs.integers(1, 3, 5000)
        put('stats/%d/a' % i, a); put('stats/%d/b' % i, b)
        res.append(dict(boot=list(A56['boot_ratio'](a, b, B=2000)), mcnemar=list(A56['mcnemar'](a, b))))
    index['stats'] = res


def main():
    quick = '--quick' in sys.argv
    t0 = time.time()
    g_philox()
    g_stats()
    g_front()
    g_arpeggio(quick)
    g_chord(quick)
    g_b200(quick)
    np.savez_compressed(os.path.join(OUT, 'golden.npz'), **arrays)
    json.dump(index, open(os.path.join(OUT, 'golden.json'), 'w'), indent=1, sort_keys=True)
    print('wrote', len(arrays), 'arrays in %.0fs' % (time.time() - t0))


if __name__ == '__main__':
    main()
