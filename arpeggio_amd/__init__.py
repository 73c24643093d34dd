"""arpeggio_amd: MI355X-native (gfx950, CDNA4) kernel path of the delta-encoded Ising annealing MIMO detector.

Drop-in modules (same call shapes as the reference's Python API):
  chordfx    Kernel(...).chord / .philox_kat, to_fixed, philox, KAT, QSCALE
  arpeggiofx solve(G, h, P, R, T, q, inst, key, start)
The kernels live in libarpeggio_hip.so (hand-written HIP, built in-tree by arpeggio_amd.build)."""
from . import _lib  # noqa: F401

__all__ = ['chordfx', 'arpeggiofx']
