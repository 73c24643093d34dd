"""The C-ABI library loads and exports every entry point include/arpeggio_hip.h declares (no GPU work)."""
import ctypes
import os
import re

from conftest import ROOT


def declared():
    txt = open(os.path.join(ROOT, 'include', 'arpeggio_hip.h')).read()
    return sorted(set(re.findall(r'^\s*(?:const\s+)?\w+\s*\*?\s*(arp_\w+)\s*\(', txt, re.M)))


def test_header_symbols_exported():
    from arpeggio_amd import _lib
    names = declared()
    assert len(names) >= 5
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from arpeggio_amd import _lib
    names = set(declared()) - {'arp_last_error', 'arp_abi_version'} - _lib.SPECIAL
    assert names <= set(_lib.SIGNATURES), names - set(_lib.SIGNATURES)
    L = _lib.load()
    assert L.arp_abi_version() >= 1
    assert L.arp_last_error() == b''


def test_argument_errors_without_gpu():
    """Invalid sizes are refused before any HIP call, with a message."""
    from arpeggio_amd import _lib
    L = _lib.load()
    rc = L.arp_anneal_arpeggio(None, None, None, 8, 0, 3, 2, 16, 1, None, 0, 0, 0, 0, None, None, None, None)
    assert rc == -1 and b'P = 3' in L.arp_last_error()
    rc = L.arp_anneal_chord(None, None, None, None, None, 0, 8, 2, 2, 48, 1, 0, 0, 0, 0, None, None, None, None)
    assert rc == -3
    assert L.arp_anneal_arpeggio(None, None, None, 8, 0, 2, 2, 16, 0, None, 0, 0, 0, 0, None, None, None, None) == 0


def test_batched_argument_errors_without_gpu():
    import ctypes as ct
    from arpeggio_amd import _lib
    L = _lib.load()
    arr = (_lib.BatchClass * 2)()
    arr[0] = _lib.BatchClass(None, None, None, None, None, None, None, 200, 0, 8, 0, 0, 1)
    rc = L.arp_anneal_arpeggio_batched(ct.cast(arr, ct.c_void_p), 1, 4, 0, 0, ct.c_void_p(16), 4, 4, None)
    assert rc == -3 and b'D = 200' in L.arp_last_error()
    assert L.arp_anneal_arpeggio_batched(ct.cast(arr, ct.c_void_p), 0, 4, 0, 0, ct.c_void_p(16), 4, 4, None) == -1
    arr[0] = _lib.BatchClass(None, None, None, None, None, None, None, 64, 0, 8, 0, 0, 1)
    assert L.arp_anneal_arpeggio_batched(ct.cast(arr, ct.c_void_p), 1, 4, 0, 0, ct.c_void_p(16), 4, 4, None) == 0
    arr[0].apw = 3
    rc = L.arp_anneal_arpeggio_batched(ct.cast(arr, ct.c_void_p), 1, 4, 0, 0, ct.c_void_p(16), 4, 4, None)
    assert rc == -1 and b'apw = 3' in L.arp_last_error()


def test_anneals_per_wave_setting_without_gpu():
    from arpeggio_amd import _lib
    L = _lib.load()
    assert L.arp_set_anneals_per_wave(3) == -1 and b'outside 0..2' in L.arp_last_error()
    assert L.arp_set_anneals_per_wave(-1) == -1
    for a in (2, 1, 0):
        assert L.arp_set_anneals_per_wave(a) == 0


def test_oversized_workgroup_refused_without_gpu():
    """A launch shape above a kernel's compiled workgroup bound is an argument error on the host, before any HIP call:
    16 anneals per workgroup (1,024 threads) against the first-accept kernels' 512 and the FP32 Chord's 256. The
    pointers are host placeholders that no kernel ever sees."""
    import ctypes as ct
    from arpeggio_amd import _lib
    L = _lib.load()
    buf = ct.create_string_buffer(64)
    p = ct.cast(buf, ct.c_void_p)
    try:
        assert L.arp_set_anneal_waves(17) == -1
        assert L.arp_set_anneal_waves(16) == 0
        rc = L.arp_anneal_arpeggio(p, p, p, 8, 0, 2, 16, 128, 1, p, 0, 0, 0, 0, p, p, p, None)
        assert rc == -1 and b'exceed the kernel' in L.arp_last_error(), L.arp_last_error()
        rc = L.arp_anneal_arpeggio_ws(p, p, p, 8, 0, 2, 16, 64, 1, p, 0, 0, p, 1, p, p, p, None, None)
        assert rc == -1 and b'exceed the kernel' in L.arp_last_error()
        rc = L.arp_anneal_arpeggio_f32(p, p, p, 8, 0.0, 2, 16, 64, 1, p, 0, 0, 0, p, p, p, None, None)
        assert rc == -1 and b'exceed the kernel' in L.arp_last_error()
        rc = L.arp_anneal_chord_f32(p, p, p, p, p, 0.0, 8, 2, 16, 64, 1, 0, 0, 0, p, p, None, None, None)
        assert rc == -1 and b'exceed the kernel' in L.arp_last_error()
        arr = (_lib.BatchClass * 1)()
        arr[0] = _lib.BatchClass(p.value, p.value, p.value, p.value, p.value, p.value, p.value, 64, 4, 8, 0, 0, 1)
        rc = L.arp_anneal_arpeggio_batched(ct.cast(arr, ct.c_void_p), 1, 4, 0, 0, p, 4, 17, None)
        assert rc == -1 and b'waves = 17' in L.arp_last_error()
    finally:
        assert L.arp_set_anneal_waves(0) == 0


def test_detect_workspace_and_argument_errors_without_gpu():
    """The whole-chain entry point sizes its workspace on the host and refuses a short one before any HIP call."""
    import ctypes as ct
    from arpeggio_amd import _lib
    L = _lib.load()
    n = L.arp_detect_workspace_bytes(1000, 64, 4, 16, 1)
    assert n >= 1000 * (64 * 64 * 2 + 20 * 128) and n % 256 == 0   # int16 Gram, spins of 20 anneals
    assert L.arp_detect_workspace_bytes(1000, 64, 4, 0, 0) < n
    assert L.arp_detect_workspace_bytes(10, 200, 4, 16, 0) == 0
    buf = ct.create_string_buffer(64)
    p = ct.cast(buf, ct.c_void_p)
    rc = L.arp_detect_adaptive(p, p, None, 0.1, 1000, 64, 64, 16, p, 8, 4, p, 8, 16, 0, 40.0, p, 0, 0, None, 1, None, p,
                               None, p, 64, None)
    assert rc == -3 and b'workspace' in L.arp_last_error()
    rc = L.arp_detect_adaptive(p, p, None, 0.1, 10, 64, 64, 8, p, 8, 4, p, 8, 16, 0, 40.0, p, 0, 0, None, 1, None, p,
                               None, p, 1 << 30, None)
    assert rc == -1 and b'M must be' in L.arp_last_error()


def test_chord_lane_argument_errors_without_gpu():
    """arp_anneal_chord_lane refuses a bad copy count, dimension or stream choice before any HIP call."""
    from arpeggio_amd import _lib
    L = _lib.load()
    args = lambda P, D, rng: (None, None, None, None, 8, 0, P, 2, D, 1, None, 0, 0, rng, None, None, None, None, None,
                             None)
    assert L.arp_anneal_chord_lane(*args(3, 16, 1)) == -1 and b'P = 3' in L.arp_last_error()
    assert L.arp_anneal_chord_lane(*args(2, 129, 1)) == -3
    assert L.arp_anneal_chord_lane(*args(2, 16, 2)) == -1 and b'rng' in L.arp_last_error()
    assert L.arp_anneal_chord_lane(*args(2, 16, 1)) == -1 and b'null' in L.arp_last_error()


def test_list_anneal_gram_forms_without_gpu():
    """arp_anneal_arpeggio_list takes the Gram as int32 (Gt) or int16 (Gh), exactly one of them."""
    import ctypes
    from arpeggio_amd import _lib
    L = _lib.load()
    one = ctypes.c_int(0)
    args = lambda gt, gh: (gt, gh, None, None, 8, 0, 2, 2, 16, 1, None, 0, 0, 0, None, 0, 0, None, None, None, None,
                           None, None, None)
    p = ctypes.addressof(one)
    for gt, gh in ((None, None), (p, p)):
        assert L.arp_anneal_arpeggio_list(*args(gt, gh)) == -1 and b'exactly one' in L.arp_last_error()
    assert L.arp_abi_version() == 5


def test_fused_mlp_feature_width_without_gpu():
    """arp_mlp_start_fused takes feature rows up to 1016 wide (the tile's activations fit the 160 KiB of LDS)."""
    from arpeggio_amd import _lib
    L = _lib.load()
    args = lambda F, N: (None, 4, F, None, None, None, None, None, None, 256, N, 1.0, None, None)
    assert L.arp_mlp_start_fused(*args(1048, 256)) == -1 and b'32..1016' in L.arp_last_error()
    assert L.arp_mlp_start_fused(*args(544, 320)) == -1 and b'N = 320' in L.arp_last_error()
    assert L.arp_mlp_start_fused(*args(544, 256)) == -1 and b'null' in L.arp_last_error()


def test_soft_argument_errors_without_gpu():
    from arpeggio_amd import _lib
    L = _lib.load()
    assert L.arp_soft_mmse_llr(None, None, 4, 32, 8, None, None) == -1 and b'M must be' in L.arp_last_error()
    assert L.arp_soft_mmse_llr(None, None, 0, 32, 16, None, None) == 0
    assert L.arp_soft_list_llr(None, None, None, None, 4, 8, 2, 200, 64, 16, 1, None, 1.0, 50.0, None, None, None,
                               None) == -3
    assert L.arp_soft_list_llr(None, None, None, None, 4, 8, 2, 64, 64, 16, 1, None, 1.0, 0.0, None, None, None,
                               None) == -1 and b'clip' in L.arp_last_error()
    assert L.arp_soft_list_llr(None, None, None, None, 4, 64, 2, 64, 64, 16, 1, None, 1.0, 50.0, None, None, None,
                               None) == -3
    assert L.arp_soft_list_llr(None, None, None, None, 0, 8, 2, 64, 64, 16, 1, None, 1.0, 50.0, None, None, None,
                               None) == 0


def test_ising_ri_argument_errors_without_gpu():
    from arpeggio_amd import _lib
    L = _lib.load()
    assert L.arp_ising_ri(None, None, None, 4, 64, 64, 8, 0.15, 4096.0, None, None, None, None, None) == -1
    assert L.arp_ising_ri(None, None, None, 4, 64, 200, 16, 0.15, 4096.0, None, None, None, None, None) == -3
    assert L.arp_ising_ri(None, None, None, 4, 64, 64, 16, 0.15, 4096.0, None, None, None, None, None) == -1
    assert b'null' in L.arp_last_error()
    assert L.arp_ising_ri(None, None, None, 0, 64, 64, 16, 0.15, 4096.0, None, None, None, None, None) == 0


def test_batched_chain_argument_errors_without_gpu():
    from arpeggio_amd import _lib
    L = _lib.load()
    arr = (_lib.BatchFrames * 1)()
    arr[0] = _lib.BatchFrames(None, None, None, 1.0, None, None, None, None, 80, 16, 4)
    p = ctypes.cast(arr, ctypes.c_void_p)
    assert L.arp_detect_batched_workspace_bytes(p, 0, 8) == 0
    assert L.arp_detect_batched(p, 1, None, 64, 0, 8, 4096.0, 0, 0, 512, 8, 66048, None, 0, None) == -1
    T = (ctypes.c_int * 64)()
    assert L.arp_detect_batched(p, 1, T, 64, 0, 8, 4096.0, 0, 0, 512, 8, 66048, None, 0, None) == -3
    arr[0] = _lib.BatchFrames(None, None, None, 1.0, None, None, None, None, 32, 16, 4)
    assert L.arp_detect_batched(p, 1, T, 64, 0, 8, 4096.0, 0, 0, 512, 8, 66048, None, 0, None) == -1
    assert b'null' in L.arp_last_error()
    arr[0] = _lib.BatchFrames(None, None, None, 1.0, None, None, None, None, 32, 16, 0)
    assert L.arp_detect_batched(p, 1, T, 64, 0, 8, 4096.0, 0, 0, 512, 8, 66048, None, 0, None) == -1
    assert b'workspace' in L.arp_last_error()
