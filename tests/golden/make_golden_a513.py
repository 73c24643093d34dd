#!/usr/bin/env python3
"""Copies the reference's per-use fixtures of its fixed-point Arpeggio MDI runs (data only) into
tests/golden/a513_fixed.npz / a513_fixed.json. Run IN THE BUILD CONTAINER, where the reference is mounted read-only.

Source: data/results/a513_fixed/<tag>/<method>/snr_<snr>/bNNNNN.npz of the reference repository, written by
arpeggio/sim/local_code_a513/run_a513.py:160-166 (per use: bit errors, anneals; per batch: counters per K =
[flips, tests, anneals], draw seed, batch size). Each batch bi is the draw arpx.draw(seed, snr, bi, m, nt, 16)."""
import glob
import json
import os

import numpy as np

REF = '/root/reference/data/results/a513_fixed'
OUT = os.path.dirname(os.path.abspath(__file__))
RUNS = (('x16', 'arpeggiofx_64_0.1_0.1_0_rand_32', 16, 64),
        ('x32', 'arpeggiofx_32_0.1_0.1_0_rand_8', 32, 300),
        ('x64', 'arpeggiofx_64_0.1_0.1_0_bal_8', 64, 120))


def main():
    arrays, index = {}, {}
    for tag, method, nt, nb in RUNS:
        d = glob.glob(os.path.join(REF, tag, method, 'snr_*'))[0]
        files = sorted(glob.glob(os.path.join(d, 'b*.npz')))[:nb]
        be, ann, cnt = [], [], []
        rec = None
        for f in files:
            z = np.load(f, allow_pickle=False)
            info = json.loads(str(z['info']))
            assert info['bi'] == len(be), f
            be.append(z['biterr'].astype(np.uint8))
            ann.append(z['anneals'].astype(np.float32))
            cnt.append([info['cnt'].get(str(K), [0, 0, 0]) for K in (1, 2, 3)])
            rec = dict(method=method, nt=nt, seed=info['seed'], snr=info['snr'], m=info['m'], pick=info['pick'])
        arrays['%s/biterr' % tag] = np.stack(be)
        arrays['%s/anneals' % tag] = np.stack(ann)
        arrays['%s/cnt' % tag] = np.asarray(cnt, dtype=np.int64)          # (batches, K = 1..3, [flips, tests, anneals])
        rec['batches'] = len(files)
        index[tag] = rec
        print(tag, len(files), 'batches', int(arrays['%s/biterr' % tag].sum()), 'bit errors')
    np.savez_compressed(os.path.join(OUT, 'a513_fixed.npz'), **arrays)
    json.dump(index, open(os.path.join(OUT, 'a513_fixed.json'), 'w'), indent=1, sort_keys=True)


if __name__ == '__main__':
    main()
