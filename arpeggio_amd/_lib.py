"""ctypes binding of libarpeggio_hip.so (the C ABI declared in include/arpeggio_hip.h).

The library is built in-tree by arpeggio_amd.build. There is no fallback: if the library or a GPU is missing, the
entry points raise."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libarpeggio_hip.so')

_P = ctypes.c_void_p
_i32, _u32, _u64 = ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64
_l, _f = ctypes.c_long, ctypes.c_float

# name -> argtypes; every entry point returns int (0 = success)
class MlpWeights(ctypes.Structure):
    """arp_mlp_weights: the warm-start MLP's device weights for the whole-chain entry point."""
    _fields_ = [('W1t', ctypes.c_void_p), ('b1', ctypes.c_void_p), ('W2t', ctypes.c_void_p), ('b2', ctypes.c_void_p),
                ('W3t', ctypes.c_void_p), ('b3', ctypes.c_void_p), ('F', ctypes.c_int), ('H', ctypes.c_int),
                ('inv_temp', ctypes.c_float)]


class BatchClass(ctypes.Structure):
    """arp_batch_class: one size class of a mixed batch (device pointers and sizes)."""
    _fields_ = [('Gt', ctypes.c_void_p), ('h', ctypes.c_void_p), ('inst', ctypes.c_void_p), ('T', ctypes.c_void_p),
                ('out_s', ctypes.c_void_p), ('out_de', ctypes.c_void_p), ('out_st', ctypes.c_void_p),
                ('D', ctypes.c_int), ('B', ctypes.c_int), ('L', ctypes.c_int), ('q', ctypes.c_int),
                ('start', ctypes.c_int), ('tile', ctypes.c_int), ('thr', ctypes.c_void_p), ('out_used', ctypes.c_void_p),
                ('gate', ctypes.c_int), ('apw', ctypes.c_int)]


class BatchFrames(ctypes.Structure):
    """arp_batch_frames: one class of a mixed batch for the whole-chain entry point."""
    _fields_ = [('H', ctypes.c_void_p), ('y', ctypes.c_void_p), ('N0', ctypes.c_void_p), ('N0_scalar', ctypes.c_float),
                ('inst', ctypes.c_void_p), ('ltrue', ctypes.c_void_p), ('lhat', ctypes.c_void_p),
                ('biterr', ctypes.c_void_p), ('nt', ctypes.c_int), ('M', ctypes.c_int), ('B', ctypes.c_int)]


SIGNATURES = {
    'arp_philox_kat': [_P, _P, _i32, _P],
    'arp_set_anneal_waves': [_i32],
    'arp_set_anneals_per_wave': [_i32],
    'arp_probe_lds': [_P, _i32, _i32, _P],
    'arp_probe_issue': [_i32, _P, _i32, _i32, _P],
    'arp_set_front_end_form': [_i32],
    'arp_anneal_arpeggio': [_P, _P, _P, _i32, _i32, _i32, _i32, _i32, _i32, _P, _u32, _u32, _i32, _i32, _P, _P, _P,
                            _P],
    'arp_anneal_chord': [_P, _P, _P, _P, _P, _i32, _i32, _i32, _i32, _i32, _i32, _u32, _u32, _i32, _i32, _P, _P, _P,
                         _P],
    'arp_front_end': [_P, _P, _P, ctypes.c_float, _i32, _i32, _i32, _i32, _i32, ctypes.c_float, _P, _P, _P, _P, _P, _P,
                      _P, _P, _P, _P, _P, _P, _P, _i32, _P, _P],
    'arp_anneal_arpeggio_ws': [_P, _P, _P, _i32, _i32, _i32, _i32, _i32, _i32, _P, _u32, _u32, _P, _i32, _P, _P, _P, _P,
                               _P],
    'arp_anneal_arpeggio_batched': [_P, _i32, _i32, _u32, _u32, _P, _i32, _i32, _P],
    'arp_mlp_start_fused': [_P, _i32, _i32, _P, _P, _P, _P, _P, _P, _i32, _i32, _f, _P, _P],
    'arp_mlp_gemm': [_P, _l, _l, _P, _l, _l, _i32, _i32, _i32, _P, _i32, _P, _P, _P, _l, _P],
    'arp_mlp_gelu_bwd': [_P, _P, _l, _P, _P, _P],
    'arp_mlp_bce': [_P, _l, _P, _i32, _i32, _P, _P, _P, _P],
    'arp_mlp_colsum': [_P, _l, _i32, _i32, _P, _P],
    'arp_mlp_adam': [_P, _P, _P, _P, _l, _f, _f, _f, _f, _i32, _P, _P],
    'arp_cast_bf16': [_P, _l, _P, _P],
    'arp_start_thresholds': [_P, _l, _i32, _i32, _f, _P, _P],
    'arp_anneal_arpeggio_f32': [_P, _P, _P, _i32, ctypes.c_float, _i32, _i32, _i32, _i32, _P, _u32, _u32, _i32, _P, _P,
                                _P, _P, _P],
    'arp_anneal_chord_f32': [_P, _P, _P, _P, _P, ctypes.c_float, _i32, _i32, _i32, _i32, _i32, _u32, _u32, _i32, _P, _P,
                             _P, _P, _P],
    'arp_decode_select': [_P, _P, _P, _P, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _P, _P, _P, _P, _P],
    'arp_anneal_arpeggio_list': [_P, _P, _P, _P, _i32, _i32, _i32, _i32, _i32, _i32, _P, _u32, _u32, _i32, _P, _i32,
                                 _i32, _P, _P, _P, _P, _P, _P, _P],
    'arp_decode_select_list': [_P, _P, _P, _P, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _P, _P, _P, _P, _P, _P, _P,
                               _P, _P],
    'arp_certify': [_P, _P, _f, _f, _i32, _P, _P, _P],
    'arp_anneal_chord_lane': [_P, _P, _P, _P, _i32, _i32, _i32, _i32, _i32, _i32, _P, _u32, _u32, _i32, _P, _P, _P,
                              _P, _P, _P],
    'arp_mmse_sic': [_P, _P, _P, _f, _i32, _i32, _i32, _i32, _i32, _P, _P],
    'arp_mdi_radius': [_P, _P, _P, _i32, _i32, _i32, _i32, _i32, _P, _P, _P, _P],
    'arp_mdi_build': [_P, _P, _P, _P, _i32, _i32, _i32, _f, _P, _P, _P],
    'arp_detect_adaptive': [_P, _P, _P, _f, _i32, _i32, _i32, _i32, _P, _i32, _i32, _P, _i32, _i32, _i32, _f, _P, _u32,
                            _u32, _P, _i32, _P, _P, _P, _P, ctypes.c_size_t, _P],
    'arp_detect_adaptive_noalloc': [_P, _P, _P, _f, _i32, _i32, _i32, _i32, _P, _i32, _i32, _P, _i32, _i32, _i32, _f,
                                    _P, _u32, _u32, _P, _i32, _P, _P, _P, _P],
    'arp_ising_ri': [_P, _P, _P, _i32, _i32, _i32, _i32, _f, _f, _P, _P, _P, _P, _P],
    'arp_detect_batched': [_P, _i32, _P, _i32, _i32, _i32, _f, _u32, _u32, _i32, _i32, _i32, _P, ctypes.c_size_t, _P],
    'arp_tune_batched': [_P, _i32, _P, _i32, _i32, _i32, _f, _u32, _u32, _i32, _P, ctypes.c_size_t, _P, _P, _P],
    'arp_soft_mmse_llr': [_P, _P, _i32, _i32, _i32, _P, _P],
    'arp_soft_list_llr': [_P, _P, _P, _P, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _P, _f, _f, _P, _P, _P, _P],
}

# entry points whose return type is not int (bound in load())
SPECIAL = {'arp_detect_workspace_bytes', 'arp_detect_batched_workspace_bytes'}

_lib = None


class ArpeggioError(RuntimeError):
    pass


def load(path=None):
    """Loads the library (once). Raises ArpeggioError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ArpeggioError('%s is not built: run arpeggio_amd.build.build() (hipcc, gfx950)' % p)
    L = ctypes.CDLL(p)
    L.arp_last_error.restype = ctypes.c_char_p
    L.arp_last_error.argtypes = []
    L.arp_abi_version.restype = ctypes.c_int
    L.arp_detect_workspace_bytes.restype = ctypes.c_size_t
    L.arp_detect_workspace_bytes.argtypes = [_i32, _i32, _i32, _i32, _i32]
    L.arp_detect_batched_workspace_bytes.restype = ctypes.c_size_t
    L.arp_detect_batched_workspace_bytes.argtypes = [_P, _i32, _i32]
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _lib = L
    return L


def check(rc, what):
    if rc != 0:
        msg = load().arp_last_error().decode(errors='replace')
        if rc == -1 or rc == -3:
            raise ValueError('%s: %s' % (what, msg))
        raise ArpeggioError('%s failed (%d): %s' % (what, rc, msg))


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
