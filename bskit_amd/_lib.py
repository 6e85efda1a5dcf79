"""ctypes binding of libbskhip.so, the gfx950 C ABI declared in include/bskhip.h.

The library is built in-tree by ``bskit_amd/csrc/Makefile`` (``build()``).
There is no CPU fallback: every compute entry point runs on the GPU, and a
missing or unloadable library raises immediately.
"""
import ctypes
import os
import subprocess
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libbskhip.so')
CSRC = os.path.join(_HERE, 'csrc')

# constants of include/bskhip.h
OK, EINVAL, ENOMEM, EUNSUPPORTED, EHIP = 0, -1, -2, -3, -4
DT_F32, DT_F64 = 0, 1
FMT_F32, FMT_BF16X2, FMT_F64 = 0, 1, 2
KIND_DELTA, KIND_UNIT, KIND_KMAG, KIND_KPOW = 0, 1, 2, 3
MODE_AUTO, MODE_AAB, MODE_ABC = 0, 1, 2
STAGE_X, STAGE_YZ, STAGE_Y, STAGE_LEVELS = 1, 2, 4, 8
LEVELS = 4
ABI_VERSION = 7

_c_int, _c_ll, _c_double, _vp = ctypes.c_int, ctypes.c_longlong, ctypes.c_double, ctypes.c_void_p
_c_size_p = ctypes.POINTER(ctypes.c_size_t)
_int_p = ctypes.POINTER(ctypes.c_int)
_dbl_p = ctypes.POINTER(ctypes.c_double)

# name -> argtypes (all return int status unless listed in _RESTYPE)
_SIGS = {
    'bsk_version': [],
    'bsk_last_error': [],
    'bsk_shell_thresholds': [_c_int, _c_double, _c_int, _dbl_p, _dbl_p, _int_p, _int_p],
    'bsk_plan_create': [ctypes.POINTER(_vp), _c_int, _c_double, _c_int, _dbl_p, _dbl_p],
    'bsk_plan_destroy': [_vp],
    'bsk_plan_geometry': [_vp, _int_p, _int_p, _int_p, _int_p, _int_p],
    'bsk_plan_thresholds': [_vp, _int_p, _int_p],
    'bsk_plan_shell_mesh': [_vp, _int_p],
    'bsk_plan_set_shell_mesh': [_vp, ctypes.c_int],
    'bsk_plan_levels': [_vp, ctypes.POINTER(_c_ll), ctypes.POINTER(_c_ll),
                        ctypes.POINTER(_c_ll)],
    'bsk_plan_triangle_levels': [_vp, _c_int, _int_p, _int_p],
    'bsk_plan_set_variant': [_vp, _c_int, _c_int],
    'bsk_plan_set_compensation': [_vp, _c_int],
    'bsk_plan_set_kpower': [_vp, _c_double],
    'bsk_forward_workspace': [_vp, _c_int, _c_size_p],
    'bsk_forward': [_vp, _c_int, _vp, _vp, _vp, _vp],
    'bsk_shells_workspace': [_vp, _c_int, _c_int, _c_int, _c_size_p],
    'bsk_shells': [_vp, _c_int, _c_int, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp,
                   _vp, _c_ll, _c_ll, _vp, _vp],
    'bsk_shells_levels': [_vp, _c_int, _c_int, _vp, _c_int, _c_int, _c_int, _vp, _vp, _c_ll, _vp,
                          _vp, _vp],
    'bsk_pairs_shell_parts': [_vp, ctypes.POINTER(ctypes.c_ubyte)],
    'bsk_power_spectrum_workspace': [_vp, _c_size_p],
    'bsk_power_spectrum': [_vp, _c_int, _vp, _vp, _vp, _vp],
    'bsk_triples_plain_workspace': [_c_int, _c_int, _c_size_p],
    'bsk_triples_plain': [_c_int, _vp, _vp, _vp, _c_ll, _c_ll, _c_int, _vp, _c_int, _c_double, _vp,
                          _vp, _vp],
    'bsk_triples_tiled_workspace': [_c_int, _c_int, _c_int, _c_size_p],
    'bsk_triples_tiled': [_c_int, _vp, _c_ll, _c_int, _c_ll, _c_int, _vp, _c_int, _c_int, _c_double,
                          _vp, _vp, _vp],
    'bsk_pairs_create': [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _int_p],
    'bsk_pairs_create_levels': [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _int_p, _int_p],
    'bsk_pairs_segments': [_vp, _int_p, _int_p, _int_p, _int_p, _int_p, _int_p, _int_p],
    'bsk_pairs_upload': [_vp],
    'bsk_pairs_destroy': [_vp],
    'bsk_pairs_info': [_vp, _int_p, _int_p, _int_p, _int_p, _int_p],
    'bsk_pairs_slots': [_vp, _int_p],
    'bsk_pairs_tables': [_vp, _int_p, _int_p, _int_p],
    'bsk_triples_mfma_workspace': [_vp, _c_int, _c_size_p],
    'bsk_triples_mfma': [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_ll, _c_ll, _c_int, _c_double, _vp,
                         _vp, _vp],
    'bsk_triples_mfma_reduce': [_vp, _c_int, _c_double, _vp, _vp, _vp],
    'bsk_triples_mfma_levels': [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_ll,
                                ctypes.POINTER(_c_ll), ctypes.POINTER(_c_ll), _c_int, _c_double,
                                _vp, _vp, _vp],
    'bsk_triples_mfma_levels_noalloc': [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _c_ll,
                                        ctypes.POINTER(_c_ll), ctypes.POINTER(_c_ll), _c_int,
                                        _c_double, _vp, _vp],
    'bsk_triples_zfused_supported': [_vp, _vp, ctypes.POINTER(_c_int)],
    'bsk_triples_zfused': [_vp, _vp, _vp, _c_int, _c_int, _c_double, _vp, _vp, _vp],
    'bsk_cic_workspace': [_c_int, _c_ll, _vp],
    'bsk_cic_paint': [_c_int, _c_double, _c_ll, _vp, _vp, _c_int, _vp, _vp],
    'bsk_direct_triangles': [_c_int, _c_int, _vp, _c_int, _vp, _c_int, _c_int, _c_int, _vp, _vp,
                             _vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp],
}
_RESTYPE = {'bsk_last_error': ctypes.c_char_p}

_lock = threading.Lock()
_lib = None


class BskError(RuntimeError):
    """Error status returned by libbskhip."""

    def __init__(self, name, code, message):
        super().__init__('%s failed (status %d): %s' % (name, code, message))
        self.code = code


def build(verbose=False):
    """Compile libbskhip.so for gfx950 with the in-tree Makefile."""
    cmd = ['make', '-C', CSRC, '-j8']
    if verbose:
        subprocess.check_call(cmd)
    else:
        subprocess.check_call(cmd, stdout=subprocess.DEVNULL)


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError('libbskhip.so is not built (%s); run '
                                  '__graft_entry__.build() or make -C %s' % (LIB_PATH, CSRC))
            h = ctypes.CDLL(LIB_PATH)
            for name, args in _SIGS.items():
                fn = getattr(h, name)
                fn.argtypes = args
                fn.restype = _RESTYPE.get(name, _c_int)
            if h.bsk_version() != ABI_VERSION:
                raise ImportError('%s has ABI version %d, this package needs %d; rebuild it'
                                  % (LIB_PATH, h.bsk_version(), ABI_VERSION))
            _lib = h
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def last_error():
    return lib().bsk_last_error().decode(errors='replace')


def call(name, *args):
    """Call an ABI function and raise BskError on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    if rc != OK:
        raise BskError(name, rc, last_error())
    return rc


def ptr(t):
    """Device (or host) address of a tensor/array as c_void_p (None -> NULL)."""
    if t is None:
        return None
    if hasattr(t, 'data_ptr'):
        return _vp(t.data_ptr())
    return t.ctypes.data_as(_vp)


def int_array(values):
    arr = (ctypes.c_int * len(values))(*[int(v) for v in values])
    return arr


def dbl_array(values):
    return (ctypes.c_double * len(values))(*[float(v) for v in values])


def size_out():
    return ctypes.c_size_t(0)
