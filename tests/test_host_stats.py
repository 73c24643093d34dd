"""Host statistics used by the bench's fixed-BER comparison: the vectorised paired bootstrap against the oracle's
replicate-by-replicate form (which the reference's published outputs pin), plus closed-form cases."""
import numpy as np

from arpeggio_amd.stats import boot_ratio


def test_boot_ratio_closed_forms():
    a = np.array([0, 3, 0, 1, 0, 0, 2])
    assert boot_ratio(a, a) [0] == 1.0 and boot_ratio(a, a)[1:] == (1.0, 1.0)
    assert boot_ratio(np.zeros(5), np.zeros(5)) == (1.0, 1.0, 1.0)
    r, lo, hi = boot_ratio(2 * a, a)
    assert r == 2.0 and lo == 2.0 and hi == 2.0              # proportional vectors: every replicate is exactly 2
    assert boot_ratio([], []) == (None, None, None)


def test_boot_ratio_matches_oracle_distribution(orc):
    rs = np.random.default_rng(4)
    U = 20000
    b = rs.poisson(0.02, U) * rs.integers(1, 4, U)
    a = np.where(rs.random(U) < 0.7, b, rs.poisson(0.02, U))
    r0, lo0, hi0 = orc.boot_ratio(a, b, B=2000)
    r1, lo1, hi1 = boot_ratio(a, b, reps=2000, seed=9)
    assert r0 == r1
    width = hi0 - lo0
    assert abs(lo1 - lo0) < 0.1 * width and abs(hi1 - hi0) < 0.1 * width
