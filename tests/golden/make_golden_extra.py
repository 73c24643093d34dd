#!/usr/bin/env python3
"""Writes tests/golden/golden_extra.npz / golden_extra.json by running more of the reference's own functions IN THE
BUILD CONTAINER (the reference repository is mounted read-only at /root/reference there). Data only: inputs and the
reference's outputs. Nothing here runs on the GPU box, and no test imports the reference.

Reference functions run (paths relative to the reference repository):
  arpeggio/sim/x3model/chord_ref.py          compress, lane_seeds, xs32 (production xorshift streams)
  arpeggio/mimo/mimo.py                       ising (RI), spins_to_levels, mmse_sic, mdi_radius, mdi, mmse_soft_llr,
                                              list_llr, select, dist
  arpeggio/sim/local_code_a513/run_a513.py   arpeggio_schedule (the function definition only, loaded from its AST:
                                              the module imports the CUDA loader)
  arpeggio/sim/arpx.py                        draw_seed, draw (function definitions only; the module imports numba)
Usage: python tests/golden/make_golden_extra.py"""
import ast
import json
import math
import os
import sys

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
arrays, index = {}, {}


def load_functions(path, names, extra):
    src = open(path).read()
    tree = ast.parse(src)
    keep = [n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name in names]
    ns = dict(extra)
    exec(compile(ast.Module(body=keep, type_ignores=[]), path, 'exec'), ns)
    return ns


ARX = load_functions(os.path.join(REF, 'arpeggio', 'sim', 'arpx.py'), {'draw_seed', 'draw'},
                     dict(torch=torch, math=math, mimo=mimo, np=np))
A513 = load_functions(os.path.join(REF, 'arpeggio', 'sim', 'local_code_a513', 'run_a513.py'), {'arpeggio_schedule'},
                      dict(torch=torch, math=math, mimo=mimo, chordfx=chordfx, np=np))


def put(key, a):
    if torch.is_tensor(a):
        a = a.detach().cpu().numpy()
    arrays[key] = np.ascontiguousarray(a)


def g_compress():
    """chord_ref.compress on to_fixed outputs of dDI problems (K = 1..3) around the MMSE levels."""
    out = {}
    for tag, nt, snr, K in (('x16K1', 16, 18.0, 1), ('x16K2', 16, 16.0, 2), ('x16K3', 16, 20.0, 3),
                            ('x32K1', 32, 22.0, 1)):
        b = ARX['draw'](9810, snr, K, 4, nt, 16, dev='cpu')
        z, _, _ = mimo.mmse(b)
        lm = mimo.quantize_levels(z / mimo._mu(b, z), 16)
        I = mimo.ising_delta_at(b, lm, K)
        F = chordfx.to_fixed(I, 128 * 2 * K, 0.3, 0.3, 0.4, 0.0)
        G, h = chord_ref.compress(F, K)
        put('compress/%s/J' % tag, F['J'])
        put('compress/%s/h' % tag, F['h'])
        put('compress/%s/h0' % tag, F['h0'])
        put('compress/%s/G' % tag, G)
        put('compress/%s/hG' % tag, h)
        out[tag] = dict(K=K)
    index['compress'] = out


def g_ising_ri():
    """mimo.ising (RI-MIMO mapping with the MMSE penalty) and mimo.spins_to_levels."""
    out = {}
    for tag, nt, M, snr in (('q16', 16, 4, 10.0), ('x16', 16, 16, 18.0), ('q32', 32, 4, 12.0)):
        b = ARX['draw'](9820, snr, 0, 6, nt, M, dev='cpu')
        z, _, _ = mimo.mmse(b)
        I = mimo.ising(b, z)
        g = torch.Generator(device='cpu')
        g.manual_seed(11)
        N = I['h'].shape[1]
        S = torch.where(torch.rand(6, 3, N, generator=g) < 0.5, -1.0, 1.0)
        Lc = mimo.spins_to_levels(S, M, 2 * nt)
        for k, v in dict(H=b['H'], y=b['y'], z=z, h=I['h'], h0=I['h0'], J=I['J'], sP=I['sP'], scale=I['scale'], S=S,
                         Lc=Lc).items():
            put('ri/%s/%s' % (tag, k), v)
        out[tag] = dict(nt=nt, M=M, snr=snr, N0=b['N0'], r=I['r'])
    index['ri'] = out


def g_schedule():
    """run_a513.arpeggio_schedule for the reference's picks (x16, x32, x64) and a hot schedule, K = 1..3."""
    out = []
    for S, T0, T1, q in ((64, 0.1, 0.1, 0.0), (32, 0.1, 0.1, 0.0), (8, 1.2, 0.1, 0.2), (16, 0.5, 0.05, 0.35)):
        for K in (1, 2, 3):
            T, qi = A513['arpeggio_schedule'](dict(S=S, T0=T0, T1=T1, q=q), K)
            out.append(dict(S=S, T0=T0, T1=T1, q=q, K=K, T=[int(t) for t in T], qi=int(qi)))
    index['schedule'] = out


def g_sic_radius():
    """mimo.mmse_sic (per real stream and per user) and mimo.mdi_radius (mean / max / min eigenvalue rules)."""
    out = {}
    for tag, nt, M, snr in (('x16', 16, 16, 18.0), ('x32', 32, 16, 22.0), ('q16', 16, 4, 8.0)):
        B = 16
        b = ARX['draw'](9830, snr, 1, B, nt, M, dev='cpu')
        z, _, _ = mimo.mmse(b)
        lm = mimo.quantize_levels(z / mimo._mu(b, z), M)
        xs_r = mimo.mmse_sic(b, per='real')
        xs_c = mimo.mmse_sic(b, per='complex')
        for k, v in dict(H=b['H'], y=b['y'], l=b['l'], lm=lm, sic_real=xs_r, sic_complex=xs_c).items():
            put('sic/%s/%s' % (tag, k), v)
        rec = dict(nt=nt, M=M, snr=snr, N0=b['N0'], B=B)
        for gname, xg in (('mmse', lm), ('sic', xs_r)):
            for heur in ('mean', 'max', 'min'):
                R, Gam = mimo.mdi_radius(b, xg, 3, heur)
                put('radius/%s/%s/%s/R' % (tag, gname, heur), R)
                put('radius/%s/%s/%s/Gam' % (tag, gname, heur), Gam)
        out[tag] = rec
    index['sic'] = out


def det_solver(I, na, K):
    """A deterministic stand-in solver for pinning the MDI wrapper: anneal a of a sub-problem returns the spins
    sign(h0_n + 0.3 sin(1.3 a + 0.7 n + 0.11 K)) (ties to +1)."""
    h0 = I['h0']
    n = torch.arange(h0.shape[1], dtype=torch.float32)
    a = torch.arange(na, dtype=torch.float32)
    pert = 0.3 * torch.sin(1.3 * a[:, None] + 0.7 * n[None, :] + 0.11 * K)
    return torch.where(h0[:, None, :] + pert[None] >= 0, 1.0, -1.0)


def g_mdi():
    """mimo.mdi (MMSE + SIC guesses, radius rule, K = 1..Rmax sub-problems, per-stage selection) with the
    deterministic solver above; also the per-instance anneal count it reports through stats."""
    out = {}
    for tag, nt, M, snr, Ns, Rmax, S, guesses in (('x16', 16, 16, 18.0, 16, 3, 8, ('mmse', 'sic')),
                                                  ('x16s2', 16, 16, 14.0, 7, 2, 3, ('mmse', 'sic')),
                                                  ('q16', 16, 4, 6.0, 8, 1, 1, ('mmse',))):
        B = 24
        b = ARX['draw'](9840, snr, 2, B, nt, M, dev='cpu')
        calls = []

        def solver(I, na, K):
            calls.append([int(I['h0'].shape[0]), int(na), int(K)])
            return det_solver(I, na, K)
        st = {}
        lh = mimo.mdi(b, solver, Ns, Rmax, S, guesses=guesses, stats=st)
        z, _, _ = mimo.mmse(b)
        start = dict(start_mmse=mimo.quantize_levels(z / mimo._mu(b, z), M))    # the guesses mimo.mdi starts from
        if 'sic' in guesses:
            start['start_sic'] = mimo.mmse_sic(b, per='real')
        for k, v in dict(H=b['H'], y=b['y'], l=b['l'], lhat=lh, **start).items():
            put('mdi/%s/%s' % (tag, k), v)
        out[tag] = dict(nt=nt, M=M, snr=snr, N0=b['N0'], Ns=Ns, Rmax=Rmax, S=S, guesses=list(guesses),
                        anneals=st['anneals'], calls=calls)
    index['mdi'] = out


def g_soft():
    """mimo.mmse_soft_llr and mimo.list_llr (max-log LLRs from a candidate list and the best candidate's single-bit
    neighbours)."""
    out = {}
    for tag, nt, M, snr in (('x16', 16, 16, 16.0), ('q16', 16, 4, 8.0)):
        B = 8
        b = ARX['draw'](9850, snr, 0, B, nt, M, dev='cpu')
        z, Ai, el2 = mimo.mmse(b)
        llr_m = mimo.mmse_soft_llr(b, z, Ai, el2)
        lm = mimo.mmse_hard(b, z, Ai)
        g = torch.Generator(device='cpu')
        g.manual_seed(21)
        lv = torch.tensor(mimo.PAM[M]['levels'], dtype=torch.float32)
        cand = lv[torch.randint(0, len(lv), (B, 4, 2 * nt), generator=g)]
        cand[:, 0] = b['l']       # the truth among the candidates: most bits get both hypotheses from the list
        lh, allc, d = mimo.select(b, cand, lm)
        llr_l = mimo.list_llr(b, allc, d, lh, clip=50.0)
        for k, v in dict(H=b['H'], y=b['y'], l=b['l'], z=z, llr_mmse=llr_m, cand=cand, lm=lm, lh=lh, d=d,
                         llr_list=llr_l).items():
            put('soft/%s/%s' % (tag, k), v)
        out[tag] = dict(nt=nt, M=M, snr=snr, N0=b['N0'], el2=el2)
    index['soft'] = out


def g_xs():
    """chord_ref.lane_seeds / xs32: the production engine's per-lane xorshift32 streams."""
    out = []
    for seed, inst, r, D in ((0x1234567890ABCDEF, 7, 0, 32), (99, 0, 3, 64), (2 ** 63 + 5, 2 ** 32 - 1, 11, 16)):
        st = chord_ref.lane_seeds(seed, torch.tensor([[[inst]]], dtype=torch.int64),
                                  torch.tensor([[[r]]], dtype=torch.int64), D)[0, 0]
        seq = []
        x = st.clone()
        for _ in range(5):
            x = chord_ref.xs32(x)
            seq.append(x.clone())
        out.append(dict(seed='%x' % seed, inst=inst, r=r, D=D))
        put('xs/%d/seeds' % (len(out) - 1), st)
        put('xs/%d/seq' % (len(out) - 1), torch.stack(seq))
    index['xs'] = out


def main():
    g_compress()
    g_ising_ri()
    g_schedule()
    g_sic_radius()
    g_mdi()
    g_soft()
    g_xs()
    np.savez_compressed(os.path.join(OUT, 'golden_extra.npz'), **arrays)
    json.dump(index, open(os.path.join(OUT, 'golden_extra.json'), 'w'), indent=1, sort_keys=True)
    print('wrote', len(arrays), 'arrays')


if __name__ == '__main__':
    main()
