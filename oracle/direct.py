"""ctypes wrapper of the oracle's exact periodic triangle enumerator (direct.c).

Test infrastructure only.  ``direct.c`` is built by ``oracle/Makefile`` into
``oracle/_build/liboracle_direct.so``.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, '_build', 'liboracle_direct.so')
_lib = None


def build():
    subprocess.check_call(['make', '-s', '-C', _HERE])


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oracle_direct.restype = ctypes.c_int
        _lib.oracle_shell_sizes.restype = ctypes.c_int
        _lib.oracle_count_sym.restype = ctypes.c_int
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def shell_sizes(N, n2lo, n2hi):
    lib = _load()
    lo = np.ascontiguousarray(n2lo, dtype=np.int64)
    hi = np.ascontiguousarray(n2hi, dtype=np.int64)
    out = np.zeros(len(lo), dtype=np.int64)
    rc = lib.oracle_shell_sizes(ctypes.c_int(N), ctypes.c_int(len(lo)), _p(lo), _p(hi), _p(out))
    if rc:
        raise RuntimeError('oracle_shell_sizes failed: %d' % rc)
    return out


def enumerate_triangles(N, n2lo, n2hi, tri, dk_a=None, dk_b=None, dk_c=None):
    """Exact closed-triangle enumeration.

    ``tri`` is an (T, 3) array of shell indices.  ``dk_*`` are optional full
    complex N^3 grids (FFT index order) of delta(k) for the three sides.
    Returns (count int64[T], ksum float64[T,3] in k_f units, bsum complex128[T]
    or None).
    """
    lib = _load()
    lo = np.ascontiguousarray(n2lo, dtype=np.int64)
    hi = np.ascontiguousarray(n2hi, dtype=np.int64)
    t = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
    T = len(t)
    count = np.zeros(T, dtype=np.int64)
    ksum = np.zeros((T, 3), dtype=np.float64)
    bsum = None
    parts = [None] * 6
    if dk_a is not None:
        dk_b = dk_a if dk_b is None else dk_b
        dk_c = dk_b if dk_c is None else dk_c
        keep = []
        for g in (dk_a, dk_b, dk_c):
            g = np.asarray(g)
            assert g.shape == (N, N, N)
            re = np.ascontiguousarray(g.real, dtype=np.float64)
            im = np.ascontiguousarray(g.imag, dtype=np.float64)
            keep += [re, im]
        parts = keep
        bsum = np.zeros((T, 2), dtype=np.float64)
    rc = lib.oracle_direct(ctypes.c_int(N), ctypes.c_int(len(lo)), _p(lo), _p(hi),
                           ctypes.c_int(T), _p(t), *[_p(x) for x in parts],
                           _p(count), _p(ksum), _p(bsum))
    if rc:
        raise RuntimeError('oracle_direct failed: %d' % rc)
    if bsum is not None:
        bsum = bsum[:, 0] + 1j * bsum[:, 1]
    return count, ksum, bsum


def count_triangles(N, n2lo, n2hi, tri):
    """Exact periodic closed-triangle counts (int64[T]) by direct enumeration
    of mode pairs, one pass per distinct side pair with the 48-fold cubic
    symmetry (shells below the Nyquist plane); equal to
    ``enumerate_triangles(...)[0]``, fast enough for 256^3-512^3 shell sets."""
    lib = _load()
    lo = np.ascontiguousarray(n2lo, dtype=np.int64)
    hi = np.ascontiguousarray(n2hi, dtype=np.int64)
    t = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
    count = np.zeros(len(t), dtype=np.int64)
    rc = lib.oracle_count_sym(ctypes.c_int(N), ctypes.c_int(len(lo)), _p(lo), _p(hi),
                              ctypes.c_int(len(t)), _p(t), _p(count))
    if rc:
        raise RuntimeError('oracle_count_sym failed: %d' % rc)
    return count
