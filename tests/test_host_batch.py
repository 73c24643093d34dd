"""Host-side logic of the mixed-size batch and the bench's command line (no GPU work)."""
import os
import subprocess
import sys

import torch

from conftest import ROOT


def test_draw_mixed_classes_and_arrival_order():
    from arpeggio_amd.batch import draw_mixed, gram_stride
    fr = draw_mixed(3, 1000, dev='cpu')
    sizes = {(f['nt'], f['M']): f['l'].shape[0] for f in fr}
    assert sum(sizes.values()) == 1000
    assert sizes[(16, 4)] == 200 and sizes[(32, 16)] == 200 and sizes[(64, 4)] == 100      # 40 / 40 / 20 %, half each
    assert [f['nt'] for f in fr] == sorted([f['nt'] for f in fr], reverse=True)     # largest frames first
    idx = torch.cat([f['idx'] for f in fr]).sort().values
    assert torch.equal(idx, torch.arange(1000))                                      # arrival indices, each once
    for f in fr:
        D = 2 * f['nt']
        assert f['H'].shape == (f['l'].shape[0], D, D) and f['y'].shape == (f['l'].shape[0], D)
        snr = 10 * torch.log10(f['nt'] / f['N0'])
        assert float(snr.min()) >= 8.0 - 1e-4 and float(snr.max()) <= 24.0 + 1e-4
    assert gram_stride(64) == 64 * 64 + 64 and gram_stride(128) == 128 * 128 + 128


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'], capture_output=True, text=True)
    assert r.returncode == 2 and 'one GPU' in r.stderr
