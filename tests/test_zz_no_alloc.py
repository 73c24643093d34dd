"""The allocation-free whole-chain entry point (arp_detect_adaptive_noalloc) against the workspace entry point
(arp_detect_adaptive), through the C ABI: bit-equal on two calls in a row, the second with fewer frames (so the second
call runs on the library's array as the first left it), for the learned two-stage chain and the random-start one-stage
chain; and a shape beyond the array's 32 MiB is refused with a clear error before anything runs."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print('[case]', *a, flush=True)


def test_noalloc_equals_workspace_entry_point():
    from arpeggio_amd import frames
    from arpeggio_amd.detector import detect_chain
    from arpeggio_amd.learned import DEFAULT_WEIGHTS, WarmStartMLP
    model = WarmStartMLP.load(DEFAULT_WEIGHTS)
    for (B, snr, start, R2) in ((1024, 21.0, 'learned', 16), (300, 19.0, 'learned', 16), (700, 22.0, 'rand', 0),
                                (64, 18.0, 'rand', 8)):
        say('noalloc', B, snr, start, R2)
        b = frames.draw(811, snr, B, B, 32, 16, dev='cuda')
        lt = b['l'].to(torch.int8).contiguous()
        m = model if start == 'learned' else None
        ref = detect_chain(b['H'], b['y'], b['N0'], 16, R1=4, L1=24, R2=R2, L2=64, c=3.0, model=m, key=77, ltrue=lt)
        got = detect_chain(b['H'], b['y'], b['N0'], 16, R1=4, L1=24, R2=R2, L2=64, c=3.0, model=m, key=77, ltrue=lt,
                           noalloc=True)
        torch.cuda.synchronize()
        assert torch.equal(ref[0], got[0]) and torch.equal(ref[1], got[1]), (B, start, R2)


def test_noalloc_refuses_oversized_shape():
    from arpeggio_amd import frames
    from arpeggio_amd.detector import detect_chain
    say('noalloc oversized')
    b = frames.draw(812, 20.0, 0, 4096, 32, 16, dev='cuda')
    with pytest.raises(ValueError, match='scratch'):
        detect_chain(b['H'], b['y'], b['N0'], 16, R1=4, L1=24, R2=16, L2=64, noalloc=True)
