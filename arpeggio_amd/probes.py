"""Chip probes behind the stated bounds (no reference counterpart); prints one JSON object.

    python3 -m arpeggio_amd.probes

lds: the LDS read bandwidth of the whole chip (arp_probe_lds: conflict-free 16-byte reads from every CU), in GB/s and
bytes per clock per CU at the 2.4 GHz peak clock. issue: cycles per iteration and SIMD of 8 independent FP32 MFMAs
(16x16x4), of 64 independent vector FMAs, and of both in each wave (arp_probe_issue); `overlap` is the fraction of
the shorter of the two hidden when they share a wave (0: their times add, 1: fully overlapped)."""
import json

import torch

from . import _lib

CLOCK_HZ = 2.4e9


def lds(workgroups=2048, iters=20000):
    L = _lib.load()
    out = torch.empty(workgroups * 256, device='cuda')
    _lib.check(L.arp_probe_lds(_lib.ptr(out), workgroups, 100, _lib.stream_ptr()), 'probe_lds')
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    _lib.check(L.arp_probe_lds(_lib.ptr(out), workgroups, iters, _lib.stream_ptr()), 'probe_lds')
    e1.record()
    torch.cuda.synchronize()
    s = e0.elapsed_time(e1) * 1e-3
    nbytes = workgroups * 4 * iters * 8 * 1024   # 4 waves, 8 reads of 1 KiB per wave and iteration
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    return dict(GBs=nbytes / s / 1e9, bytes_per_clock_per_cu=nbytes / s / cus / CLOCK_HZ, seconds=s, bytes=nbytes,
                workgroups=workgroups, compute_units=cus)


def issue(workgroups=4096, iters=2000):
    L = _lib.load()
    out = torch.empty(workgroups * 256, device='cuda')
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    res = {}
    for mode, name in ((0, 'mfma'), (1, 'fma'), (2, 'both')):
        _lib.check(L.arp_probe_issue(mode, _lib.ptr(out), workgroups, 100, _lib.stream_ptr()), 'probe_issue')
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        _lib.check(L.arp_probe_issue(mode, _lib.ptr(out), workgroups, iters, _lib.stream_ptr()), 'probe_issue')
        e1.record()
        torch.cuda.synchronize()
        s = e0.elapsed_time(e1) * 1e-3
        waves_per_simd = workgroups * 4 / (cus * 4)
        res[name] = s * CLOCK_HZ / waves_per_simd / iters
    lo, hi = sorted((res['mfma'], res['fma']))
    res['overlap'] = (res['mfma'] + res['fma'] - res['both']) / lo
    return res


def main():
    print(json.dumps(dict(lds=lds(), issue=issue(), clock_hz=CLOCK_HZ)))


if __name__ == '__main__':
    main()
