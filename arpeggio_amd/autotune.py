"""Launch-shape autotuner for the persistent mixed-size anneal (arp_anneal_arpeggio_batched).

The shape is (workgroups, waves per workgroup, LDS budget, anneals per wave): the budget fixes how many frames of each
class share a tile (budget // staged frame size) and, with the CU's 160 KiB of LDS, how many workgroups stay resident
per CU; the grid is exactly the resident capacity (a persistent grid). Anneals per wave (copies per wave) is 1 or 2
for the classes of D <= 32: two anneals share a wave64, one per 32-lane half. Each candidate is timed on the batch itself with HIP
events on the launch stream; the fastest is kept and cached per (class sizes, R, L)."""
import torch

from .batch import gram_stride

LDS_PER_CU = 160 * 1024
_CACHE = {}


def candidates(det, waves=(4, 8, 16), apws=None):
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    need = max(4 * gram_stride(c['D']) for c in det.cls)
    budgets = sorted({need, min(LDS_PER_CU, 2 * need), LDS_PER_CU // 2, LDS_PER_CU})
    if apws is None:
        apws = (1, 2) if any(c['D'] <= 32 for c in det.cls) else (1,)
    out = []
    for b in budgets:
        if b < need:
            continue
        per_cu = max(1, LDS_PER_CU // b)
        for a in apws:
            for w in waves:
                out.append((cus * per_cu, w, b, a))
    return out


def time_shape(det, shape, reps=3, stream=None):
    det.set_shape(shape)
    det.anneal(stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s = stream or torch.cuda.current_stream()
    e0.record(s)
    for _ in range(reps):
        det.anneal(stream)
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_batched(det, reps=3, cands=None, log=None):
    """Sets det's fastest launch shape; returns (best_ms, best_shape, [(ms, shape), ...])."""
    key = (tuple((c['D'], c['B']) for c in det.cls), det.R, det.L)
    if key in _CACHE:
        det.set_shape(_CACHE[key][1])
        return _CACHE[key]
    res = []
    for shape in cands or candidates(det):
        ms = time_shape(det, shape, reps)
        res.append((ms, shape))
        if log:
            log('autotune: workgroups %d, waves %d, LDS budget %d B, anneals per wave %d: %.3f ms' % (
                tuple(shape) + (ms,)))
    best = min(res)
    det.set_shape(best[1])
    _CACHE[key] = (best[0], best[1], res)
    return _CACHE[key]
