"""Device engine of the FFT bispectrum path on one MI355X (gfx950).

``GridEngine`` holds one libbskhip plan for an (Nmesh, BoxSize, bin set) key
and drives the kernels on torch-allocated HBM buffers:

    field --forward--> delta(k) compact cube --shells--> nb shell fields
          --triples--> per-triangle FP64 sums

torch is used only for device memory and the current stream; every
arithmetic step is a libbskhip kernel.  There is no CPU fallback: without a
GPU or without the built library these calls raise.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib as L

# Triple-product paths
PATH_PLAIN = 'plain'      # one sweep per triangle in the shell precision
PATH_MFMA = 'mfma'        # split-BF16 MFMA, FP32 accumulate, FP64 reduce
PATH_TILED = 'tiled'      # LDS tiled VALU in the shell precision
PATHS = (PATH_PLAIN, PATH_MFMA, PATH_TILED)

_MODE = {1: L.MODE_AUTO, 2: L.MODE_AAB, 3: L.MODE_ABC}


def require_device(device=None):
    if not torch.cuda.is_available():
        raise RuntimeError('bskit_amd needs a ROCm GPU (torch.cuda.is_available() is False); '
                           'there is no CPU fallback')
    L.lib()
    return torch.device(device if device is not None else 'cuda')


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _dt(dtype):
    if dtype == torch.float32:
        return L.DT_F32
    if dtype == torch.float64:
        return L.DT_F64
    raise TypeError('fields must be float32 or float64, got %s' % dtype)


def shell_limits(nmesh, box_size, k_low, k_high):
    """Integer |n|^2 limits of the closed shells (host call into the ABI)."""
    k_low = np.atleast_1d(np.asarray(k_low, dtype=np.float64))
    k_high = np.atleast_1d(np.asarray(k_high, dtype=np.float64))
    n = len(k_low)
    lo = (ctypes.c_int * n)()
    hi = (ctypes.c_int * n)()
    L.call('bsk_shell_thresholds', nmesh, float(box_size), n, L.dbl_array(k_low),
           L.dbl_array(k_high), lo, hi)
    return np.array(lo[:], dtype=np.int64), np.array(hi[:], dtype=np.int64)


class Workspace:
    """Grow-only scratch buffers (device bytes) keyed by purpose."""

    def __init__(self, device):
        self.device = device
        self._bufs = {}

    def get(self, key, nbytes):
        nbytes = max(int(nbytes), 256)
        b = self._bufs.get(key)
        if b is None or b.numel() < nbytes:
            b = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            self._bufs[key] = b
        return b

    def release(self):
        self._bufs.clear()


class LevelLayout:
    """Sublattice level layout of a shell row on an S^3 mesh (S % 4 == 0):
    level L's cells are data[off[L]:off[L] + ncell[L]] (include/bskhip.h,
    bsk_plan_levels); the parts C0 minus E, E, O, C1 minus O (C0 / C1: even /
    odd x+y+z, E / O: x, y, z all even / all odd), each zero-padded to a
    multiple of 128 cells."""

    def __init__(self, S, off, ncell, stride):
        self.S = S
        self.off = tuple(int(v) for v in off)
        self.ncell = tuple(int(v) for v in ncell)
        self.stride = int(stride)
        self._perm = None

    def __eq__(self, other):
        return isinstance(other, LevelLayout) and (self.S, self.off, self.ncell, self.stride) == (
            other.S, other.off, other.ncell, other.stride)

    def part_cells(self):
        """Padded cell counts of the parts C0 minus E, E, O, C1 minus O."""
        o, n = self.off, self.ncell
        return (o[3] - o[1], n[3], o[2] + n[2] - o[3] - n[3], o[0] + n[0] - o[2] - n[2])

    def c_order_index(self, device):
        """Row position of every cell in C order (x, y, z), for inspection."""
        if self._perm is None or self._perm.device != torch.device(device):
            S, nh = self.S, self.S // 2
            nc, ne, e = self.off[2], self.ncell[3], nh ** 3  # P0 / P3 span, E / O span
            i = torch.arange(S, device=device)
            x, y, z = i.view(S, 1, 1), i.view(1, S, 1), i.view(1, 1, S)
            xy = 2 * (x % 2) + y % 2  # row class: 0 both even .. 3 both odd
            pos = ((x // 2) * nh + y // 2) * nh + z // 2
            c0 = (x + y + z) % 2 == 0
            part_off = torch.where(
                c0, torch.where((xy == 0) & (z % 2 == 0), nc, (xy - 1) * e),
                torch.where((xy == 3) & (z % 2 == 1), nc + ne, nc + 2 * ne + xy * e))
            self._perm = (part_off + pos).reshape(-1)
        return self._perm


class ShellParts:
    """Layout parts (bit p) to write per shell: device bytes and a host copy."""

    def __init__(self, host, device):
        self.host = np.ascontiguousarray(host, dtype=np.uint8)
        self.dev = torch.from_numpy(self.host.copy()).to(device)


class ShellSet:
    """nb shell fields of one input field in one storage format: rows of
    ``stride`` elements holding the S^3 cells in C order, or (``layout``) in
    the sublattice level layout."""

    def __init__(self, fmt, data, lo=None, ncell=None, layout=None):
        self.fmt = fmt
        self.data = data      # F32/F64: [nb, stride] float; BF16X2: hi plane int16
        self.lo = lo          # BF16X2 lo plane
        self.nb = data.shape[0]
        self.stride = data.shape[1]
        self.ncell = self.stride if ncell is None else int(ncell)
        self.layout = layout
        self.parts = None  # level layout: parts written per shell (None: all)

    @property
    def sum_cells(self):
        """Leading elements of a row that hold every cell (and zeros)."""
        return self.ncell if self.layout is None else self.layout.ncell[0]

    @staticmethod
    def _f(v):
        return (v.to(torch.int32) << 16).view(torch.float32)

    def _values(self, a, b):
        if self.fmt != L.FMT_BF16X2:
            return self.data[:, a:b]
        return self._f(self.data[:, a:b]) + self._f(self.lo[:, a:b])

    def as_float(self):
        """FP32/FP64 values (hi + lo for BF16X2) of the S^3 cells in C order,
        for inspection."""
        if self.layout is None:
            return self._values(0, self.ncell)
        return self._values(0, self.layout.ncell[0])[:, self.layout.c_order_index(self.data.device)]

    def level_values(self, level):
        """Row range of a sublattice level as FP32/FP64 values."""
        o, n = self.layout.off[level], self.layout.ncell[level]
        return self._values(o, o + n)


class PairPlan:
    """Owner of a bsk_pairs plan: the MFMA kernel's tiles for one triangle list
    (with ``levels``: the sublattice level each triangle is summed over)."""

    def __init__(self, tri, nshell, mode, upload=True, levels=None):
        tri = np.ascontiguousarray(np.asarray(tri, dtype=np.int32).reshape(-1, 3))
        h = ctypes.c_void_p()
        ip = ctypes.POINTER(ctypes.c_int)
        if levels is None:
            L.call('bsk_pairs_create', ctypes.byref(h), int(mode), int(nshell), len(tri),
                   tri.ctypes.data_as(ip))
        else:
            levels = np.ascontiguousarray(np.asarray(levels, dtype=np.int32).reshape(-1))
            L.call('bsk_pairs_create_levels', ctypes.byref(h), int(mode), int(nshell), len(tri),
                   tri.ctypes.data_as(ip), levels.ctypes.data_as(ip))
        self.handle = h
        ns = ctypes.c_int()
        arr = [(ctypes.c_int * 4)() for _ in range(6)]
        L.call('bsk_pairs_segments', h, ctypes.byref(ns), *arr)
        # launch segments: (layout parts bitmask, first shell, shells staged,
        # launches, items per wave, replicas of the wave lists)
        self.segments = [tuple(a[g] for a in arr) for g in range(ns.value)]
        self.list_entries = sum(8 * g[3] * g[4] for g in self.segments)
        if upload:
            L.call('bsk_pairs_upload', h)
        self.ntri = len(tri)
        v = [ctypes.c_int() for _ in range(5)]
        L.call('bsk_pairs_info', h, *[ctypes.byref(x) for x in v])
        self.npair, self.ntile, self.nitem, self.items_per_wave, self.npass = [x.value for x in v]

    def tables(self):
        """(items, partial positions, pair rows) host tables of the launches."""
        items = np.zeros(self.list_entries, dtype=np.int32)
        outpos = np.zeros(self.list_entries, dtype=np.int32)
        pairs = np.zeros((self.ntile + 1) * 16, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        L.call('bsk_pairs_tables', self.handle, items.ctypes.data_as(ip),
               outpos.ctypes.data_as(ip), pairs.ctypes.data_as(ip))
        return items, outpos, pairs

    def shell_parts(self, nshell):
        """Layout parts (bit p) in which the plan's launches stage each shell."""
        out = np.zeros(nshell, dtype=np.uint8)
        L.call('bsk_pairs_shell_parts', self.handle,
               out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
        return out

    def slots(self):
        """(T, 4) partial slots of each triangle, one per launch segment
        (-1: none); the triangle's sum is the sum of its slots x 2^level."""
        out = np.zeros((self.ntri, 4), dtype=np.int32)
        L.call('bsk_pairs_slots', self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out

    def __del__(self):
        try:
            if self.handle:
                L.lib().bsk_pairs_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class GridEngine:
    """Plan + kernels for one (Nmesh, BoxSize, bins) key on one device.

    The forward transform runs on the N^3 input mesh; shells and triple sums
    run on the shell mesh S^3 (``self.S``, ``self.ncell = S^3``).  S is the
    smallest supported size above 3m (m = largest retained |n| component),
    or N: the cell sums on S^3 equal (S/N)^3 times those on N^3 exactly
    (no product of three shells holds a frequency that aliases to zero on
    either mesh), so every sum is scaled by 1/S^3 where the reference
    divides by N^3.  ``shell_mesh=N`` evaluates on the reference's own mesh.
    """

    def __init__(self, nmesh, box_size, k_low, k_high, device=None, cic_window=0,
                 shell_mesh=None):
        self.device = require_device(device)
        self.N = int(nmesh)
        self.box = float(box_size)
        self.k_low = np.atleast_1d(np.asarray(k_low, dtype=np.float64))
        self.k_high = np.atleast_1d(np.asarray(k_high, dtype=np.float64))
        self.nb = len(self.k_low)
        if self.nb == 0:
            raise ValueError('at least one k bin is needed')
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            L.call('bsk_plan_create', ctypes.byref(h), self.N, self.box, self.nb,
                   L.dbl_array(self.k_low), L.dbl_array(self.k_high))
        self._plan = h
        g = [ctypes.c_int() for _ in range(5)]
        L.call('bsk_plan_geometry', self._plan, *[ctypes.byref(x) for x in g])
        self.m, self.MX, self.MY, self.MZ, self.nlo = [x.value for x in g]
        lo = (ctypes.c_int * self.nb)()
        hi = (ctypes.c_int * self.nb)()
        L.call('bsk_plan_thresholds', self._plan, lo, hi)
        self.n2lo = np.array(lo[:], dtype=np.int64)
        self.n2hi = np.array(hi[:], dtype=np.int64)
        self._set_shell_mesh(0 if shell_mesh is None else int(shell_mesh))
        self.work = Workspace(self.device)
        self._tri_cache = {}
        self.tuned_nwg = None
        self.tuned_zfused_nwg = None
        self.prefer_zfused = False  # set by the autotuner when the fused kernel is faster
        # BF16X2 shell sets use the sublattice level layout and the MFMA sums
        # run each triangle over the sparsest sublattice exact for it
        self.levels = True
        self.cic_window = 0
        if cic_window:
            self.set_compensation(cic_window)

    # ------------------------------------------------------------------ plan
    def close(self):
        if self._plan:
            L.lib().bsk_plan_destroy(self._plan)
            self._plan = None
        self.work.release()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _set_shell_mesh(self, shell_mesh):
        L.call('bsk_plan_set_shell_mesh', self._plan, int(shell_mesh))
        s = ctypes.c_int()
        L.call('bsk_plan_shell_mesh', self._plan, ctypes.byref(s))
        self.S = s.value
        self.ncell = self.S ** 3
        self.layout = None  # sublattice level layout of a shell row
        if self.S % 4 == 0:
            o, n = (ctypes.c_longlong * L.LEVELS)(), (ctypes.c_longlong * L.LEVELS)()
            st = ctypes.c_longlong()
            L.call('bsk_plan_levels', self._plan, o, n, ctypes.byref(st))
            self.layout = LevelLayout(self.S, o, n, st.value)

    def set_shell_mesh(self, shell_mesh=0):
        """Shell mesh size (0 = automatic, Nmesh = the reference's own mesh)."""
        self._set_shell_mesh(shell_mesh)
        self._tri_cache.clear()

    def set_compensation(self, nmesh_window):
        L.call('bsk_plan_set_compensation', self._plan, int(nmesh_window))
        self.cic_window = int(nmesh_window)

    def autotune(self, tri=None, field=None, force=False, log=None):
        """Tune radix variants and the MFMA workgroup count (cached per shape)."""
        from . import tune
        return tune.tune(self, tri=tri, field=field, force=force, log=log)

    def set_kpower(self, power):
        """Exponent p of KIND_KPOW shells (|k|^p on the shell)."""
        L.call('bsk_plan_set_kpower', self._plan, float(power))

    def set_variant(self, fft_pass, variant):
        L.call('bsk_plan_set_variant', self._plan, int(fft_pass), int(variant))

    @property
    def kf(self):
        return 2.0 * math.pi / self.box

    # --------------------------------------------------------------- forward
    def forward(self, field, stream=None, work=None, out=None, window=None):
        """delta(k) on the compact cube [MY, MX, MZ] (complex, 1/N^3 scaled).

        window: mesh size of the CIC window deconvolved in this transform
        (0: none; None: the engine's current setting).  It is read when the
        transform is enqueued, so fields with different windows share the
        engine."""
        if window is not None and int(window) != self.cic_window:
            self.set_compensation(window)
        if field.device.type != 'cuda':
            raise ValueError('field must be a device tensor')
        if tuple(field.shape) != (self.N,) * 3:
            raise ValueError('field shape %s does not match Nmesh=%d' % (tuple(field.shape), self.N))
        field = field.contiguous()
        dt = _dt(field.dtype)
        ctype = torch.complex64 if dt == L.DT_F32 else torch.complex128
        dk = out
        if dk is None:
            dk = torch.empty((self.MY, self.MX, self.MZ), dtype=ctype, device=self.device)
        elif dk.dtype != ctype or tuple(dk.shape) != (self.MY, self.MX, self.MZ):
            raise ValueError('out must be a %s [%d, %d, %d] tensor'
                             % (ctype, self.MY, self.MX, self.MZ))
        nbytes = L.size_out()
        L.call('bsk_forward_workspace', self._plan, dt, ctypes.byref(nbytes))
        w = (work or self.work).get('fwd', nbytes.value)
        L.call('bsk_forward', self._plan, dt, L.ptr(field), L.ptr(dk), L.ptr(w),
               stream_handle(stream))
        return dk

    # ---------------------------------------------------------------- shells
    def shells(self, dk=None, kind=L.KIND_DELTA, fmt=L.FMT_F32, dtype=torch.float32,
               shell0=0, nshell=None, out=None, stream=None, work=None, parts=None):
        """All (or a range of) shell fields of one spectrum on the full S^3 mesh.

        parts (level layout only): the ShellParts of shell_parts(tri), the
        layout parts each shell must hold for those triangles; MFMA sums over
        these shells then accept only triangle lists whose plans read no other
        part."""
        nshell = self.nb - shell0 if nshell is None else nshell
        dt = _dt(dtype)
        if kind == L.KIND_DELTA:
            if dk is None:
                raise ValueError('delta shells need the spectrum')
            dt = L.DT_F32 if dk.dtype == torch.complex64 else L.DT_F64
        if out is None:
            out = self.alloc_shells(fmt, self.nb)
        nbytes = L.size_out()
        L.call('bsk_shells_workspace', self._plan, dt, nshell, self.S, ctypes.byref(nbytes))
        w = (work or self.work).get('shell', nbytes.value)
        if parts is not None:
            if out.layout is None:
                raise ValueError('part masks need shells in the level layout')
            pm = parts.dev[shell0:shell0 + nshell]
            L.call('bsk_shells_levels', self._plan, dt, kind, L.ptr(dk), shell0, nshell, fmt,
                   L.ptr(out.data), L.ptr(out.lo), out.stride, L.ptr(pm), L.ptr(w),
                   stream_handle(stream))
            if out.parts is None:
                out.parts = np.full(out.nb, 0xf, np.uint8)
            out.parts[shell0:shell0 + nshell] = parts.host[shell0:shell0 + nshell]
            return out
        stages = L.STAGE_X | L.STAGE_YZ | (L.STAGE_LEVELS if out.layout else 0)
        L.call('bsk_shells', self._plan, dt, kind, L.ptr(dk), shell0, nshell, 0, self.S,
               stages, fmt, L.ptr(out.data), L.ptr(out.lo), out.stride, 0,
               L.ptr(w), stream_handle(stream))
        if out.parts is not None:
            out.parts[shell0:shell0 + nshell] = 0xf
        return out

    def shell_parts(self, tri, nfield=1):
        """ShellParts: the layout parts each shell must hold for the MFMA sums
        of tri with nfield fields (cached); pass it to shells(..., parts=)."""
        tri = np.ascontiguousarray(np.asarray(tri, dtype=np.int32).reshape(-1, 3))
        key = (tri.tobytes(), 'parts', int(nfield))
        t = self._tri_cache.get(key)
        if t is None:
            plan = self.pair_plan(tri, self.nb, _MODE[nfield], levels=True)
            t = ShellParts(plan.shell_parts(self.nb), self.device)
            self._tri_cache[key] = t
        return t

    def alloc_shells(self, fmt, nb, levels=None):
        """Shell storage for nb shells: C order, or the sublattice level
        layout (default for BF16X2, the MFMA path; not for S % 4 != 0)."""
        if levels is None:
            levels = self.uses_levels(fmt)
        lay = self.layout if levels else None
        stride = self.ncell if lay is None else lay.stride
        if fmt == L.FMT_BF16X2:
            hi = torch.empty((nb, stride), dtype=torch.int16, device=self.device)
            lo = torch.empty((nb, stride), dtype=torch.int16, device=self.device)
            return ShellSet(fmt, hi, lo, self.ncell, lay)
        dtype = torch.float32 if fmt == L.FMT_F32 else torch.float64
        return ShellSet(fmt, torch.empty((nb, stride), dtype=dtype, device=self.device),
                        None, self.ncell, lay)

    def triangle_levels(self, tri):
        """Sublattice level (0..3) each triangle (bin indices) is summed over:
        the sparsest one on which its sum is exact."""
        tri = np.ascontiguousarray(np.asarray(tri, dtype=np.int32).reshape(-1, 3))
        out = np.zeros(len(tri), dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        L.call('bsk_plan_triangle_levels', self._plan, len(tri), tri.ctypes.data_as(ip),
               out.ctypes.data_as(ip))
        return out

    # --------------------------------------------------------------- triples
    def triangles_device(self, tri):
        tri = np.ascontiguousarray(np.asarray(tri, dtype=np.int32).reshape(-1, 3))
        key = (tri.tobytes(), 'tri')
        t = self._tri_cache.get(key)
        if t is None:
            t = torch.from_numpy(tri.copy()).to(self.device)
            self._tri_cache[key] = t
        return t

    def pair_plan(self, tri, nshell, mode, levels=False):
        """Cached MFMA pair plan of a triangle list (host tables uploaded once);
        levels=True sums each triangle over its sublattice level."""
        tri = np.ascontiguousarray(np.asarray(tri, dtype=np.int32).reshape(-1, 3))
        key = (tri.tobytes(), 'pairs', nshell, mode, bool(levels))
        r = self._tri_cache.get(key)
        if r is None:
            lv = self.triangle_levels(tri) if levels else None
            with torch.cuda.device(self.device):
                r = PairPlan(tri, nshell, mode, levels=lv)
            self._tri_cache[key] = r
        return r

    def mfma_nwg(self, ncell=None):
        """Workgroups of the MFMA kernel: the tuned count, else up to 1024 with
        >= 8 chunks each."""
        if self.tuned_nwg and (ncell is None or ncell == self.ncell):
            return self.tuned_nwg
        nchunk = (self.ncell if ncell is None else ncell) // 128
        return int(max(1, min(1024, nchunk // 8)))

    @staticmethod
    def mfma_supported(nshell, nfield, tri):
        """Whether the MFMA kernel covers this shell count / field roles."""
        return nshell <= (96 if nfield == 3 else 128)

    def zfused_supported(self, tri):
        """Whether the fused z c2r + split-BF16 triple-sum kernel covers the
        auto spectrum of all nb shells with this triangle list (shell mesh
        size, shell count, pair-plan shape)."""
        plan = self.pair_plan(tri, self.nb, L.MODE_AUTO)
        ok = ctypes.c_int()
        L.call('bsk_triples_zfused_supported', self._plan, plan.handle, ctypes.byref(ok))
        return bool(ok.value)

    def bispectrum_zfused(self, dk, tri, scale=1.0, out=None, stream=None, nwg=None, work=None,
                          kernel_events=None, after_shells=None):
        """scale * sum_x d_i d_j d_k over the shell mesh for every (i, j, k) of
        tri, from the FP32 spectrum dk, with the z transform of the shells
        inside the triple-sum kernel: the quantity triples([shells(dk,
        fmt=BF16X2)], tri) computes, without the shell fields in HBM."""
        tri = np.asarray(tri, dtype=np.int32).reshape(-1, 3)
        T = len(tri)
        if dk is None or dk.dtype != torch.complex64:
            raise ValueError('the fused path needs an FP32 (complex64) spectrum')
        if T and (tri.min() < 0 or tri.max() >= self.nb):
            raise IndexError('triangle shell index out of range')
        if out is None:
            out = torch.empty(T, dtype=torch.float64, device=self.device)
        if T == 0:
            return out
        plan = self.pair_plan(tri, self.nb, L.MODE_AUTO)
        s = stream_handle(stream)
        nbytes = L.size_out()
        L.call('bsk_shells_workspace', self._plan, L.DT_F32, self.nb, self.S, ctypes.byref(nbytes))
        w = (work or self.work).get('shell', nbytes.value)
        L.call('bsk_shells', self._plan, L.DT_F32, L.KIND_DELTA, L.ptr(dk), 0, self.nb, 0, self.S,
               L.STAGE_X | L.STAGE_Y, L.FMT_BF16X2, None, None, self.ncell, 0, L.ptr(w), s)
        if after_shells is not None:  # e.g. a timing mark between the two stages
            after_shells()
        nwg = self.zfused_nwg() if nwg is None else int(nwg)
        L.call('bsk_triples_mfma_workspace', plan.handle, nwg, ctypes.byref(nbytes))
        pw = (work or self.work).get('mfma', nbytes.value)
        if kernel_events is None:
            L.call('bsk_triples_zfused', self._plan, plan.handle, L.ptr(w), 0, nwg, float(scale),
                   L.ptr(out), L.ptr(pw), s)
            return out
        kernel_events[0].record(stream)
        L.call('bsk_triples_zfused', self._plan, plan.handle, L.ptr(w), 0, nwg, float(scale), None,
               L.ptr(pw), s)
        kernel_events[1].record(stream)
        L.call('bsk_triples_mfma_reduce', plan.handle, nwg, float(scale), L.ptr(out), L.ptr(pw), s)
        return out

    def zfused_nwg(self):
        """Workgroups of the fused kernel: the tuned count, else 512 (two per
        CU) with at least 8 rows each."""
        if self.tuned_zfused_nwg:
            return self.tuned_zfused_nwg
        return int(max(1, min(512, self.S * self.S // 8)))

    def plain_nchunk(self, ncell=None):
        ncell = self.ncell if ncell is None else ncell
        return int(max(1, min(4096, ncell // 65536)))

    def triples(self, sets, tri, path=PATH_MFMA, scale=1.0, out=None, stream=None, nwg=None,
                work=None, kernel_events=None):
        """scale * sum_x A_i B_j C_k for every (i, j, k) row of tri.

        sets: [A] (auto), [A, B] (<AAB>: sides A, A, B) or [A, B, C].
        Returns a float64 device tensor of length T.
        """
        tri = np.asarray(tri, dtype=np.int32).reshape(-1, 3)
        T = len(tri)
        if T and (tri.min() < 0 or tri.max() >= min(x.nb for x in sets)):
            raise IndexError('triangle shell index out of range')
        if out is None:
            out = torch.empty(T, dtype=torch.float64, device=self.device)
        if T == 0:
            return out
        nf = len(sets)
        if nf == 1:
            sa = sb = sc = sets[0]
        elif nf == 2:
            sa, sb, sc = sets[0], sets[0], sets[1]
        elif nf == 3:
            sa, sb, sc = sets
        else:
            raise ValueError('1, 2 or 3 shell sets')
        if len({x.stride for x in sets}) != 1 or len({x.ncell for x in sets}) != 1:
            raise ValueError('shell sets of one call need the same layout')
        s = stream_handle(stream)
        ncell = sa.ncell
        if path == PATH_MFMA:
            for x in sets:
                if x.fmt != L.FMT_BF16X2:
                    raise ValueError('the MFMA path needs BF16X2 shells')
            mode = _MODE[nf]
            nwg = self.mfma_nwg(ncell) if nwg is None else int(nwg)
            lay = sa.layout
            if any(x.layout != lay for x in sets):
                raise ValueError('shell sets of one call need the same layout')
            plan = self.pair_plan(tri, sa.nb, mode, levels=lay is not None)
            if lay is not None and any(x.parts is not None for x in sets):
                need = plan.shell_parts(sa.nb)
                for x in sets:
                    if x.parts is not None and np.any(need & ~x.parts):
                        raise ValueError('the shells were written for another triangle list: '
                                         'parts this plan reads are missing')
            if lay is None:
                off = (ctypes.c_longlong * L.LEVELS)(0, 0, 0, 0)
                cells = (ctypes.c_longlong * L.LEVELS)(ncell, 0, 0, 0)
            else:
                off = (ctypes.c_longlong * L.LEVELS)(*lay.off)
                cells = (ctypes.c_longlong * L.LEVELS)(*lay.ncell)
            nbytes = L.size_out()
            L.call('bsk_triples_mfma_workspace', plan.handle, nwg, ctypes.byref(nbytes))
            w = (work or self.work).get('mfma', nbytes.value)
            args = (plan.handle, L.ptr(sa.data), L.ptr(sa.lo), L.ptr(sb.data), L.ptr(sb.lo),
                    L.ptr(sc.data), L.ptr(sc.lo), sa.stride, off, cells, nwg, float(scale))
            if kernel_events is None:
                L.call('bsk_triples_mfma_levels', *args, L.ptr(out), L.ptr(w), s)
                return out
            # events around the MFMA kernels alone (the reduction follows)
            kernel_events[0].record(stream)
            L.call('bsk_triples_mfma_levels', *args, None, L.ptr(w), s)
            kernel_events[1].record(stream)
            L.call('bsk_triples_mfma_reduce', plan.handle, nwg, float(scale), L.ptr(out),
                   L.ptr(w), s)
            return out
        if path == PATH_PLAIN:
            dt = {L.FMT_F32: L.DT_F32, L.FMT_F64: L.DT_F64}.get(sa.fmt)
            if dt is None:
                raise ValueError('the plain path needs F32 or F64 shells')
            nchunk = self.plain_nchunk(ncell)
            tri_d = self.triangles_device(tri)
            nbytes = L.size_out()
            L.call('bsk_triples_plain_workspace', min(T, 65535), nchunk, ctypes.byref(nbytes))
            w = (work or self.work).get('plain', nbytes.value)
            for t0 in range(0, T, 65535):
                t1 = min(T, t0 + 65535)
                L.call('bsk_triples_plain', dt, L.ptr(sa.data), L.ptr(sb.data), L.ptr(sc.data),
                       sa.stride, sa.sum_cells, t1 - t0, L.ptr(tri_d[t0:t1]), nchunk, float(scale),
                       L.ptr(out[t0:t1]), L.ptr(w), s)
            return out
        if path == PATH_TILED:
            if nf != 1:
                raise ValueError('the tiled path takes one shell set (auto spectra)')
            dt = {L.FMT_F32: L.DT_F32, L.FMT_F64: L.DT_F64}[sa.fmt]
            nbatch = (T + 511) // 512
            nchunk = int(max(1, min(ncell // 32, 2048 // nbatch)))
            tri_d = self.triangles_device(tri)
            nbytes = L.size_out()
            L.call('bsk_triples_tiled_workspace', T, 1, nchunk, ctypes.byref(nbytes))
            w = (work or self.work).get('tiled', nbytes.value)
            L.call('bsk_triples_tiled', dt, L.ptr(sa.data), sa.stride, sa.nb, sa.sum_cells, T,
                   L.ptr(tri_d), 1,
                   nchunk, float(scale), L.ptr(out), L.ptr(w), s)
            return out
        raise ValueError('unknown triple path %r' % (path,))

    # ------------------------------------------------------- x-slab streaming
    def slab_planes(self, nfield, fmt, max_bytes=None):
        """x planes per slab so that nfield sets of nb shell slabs fit in
        max_bytes (default 40% of free HBM); a divisor of S whose slab cell
        count is a multiple of the pair kernel's 128-cell chunk."""
        S, plane = self.S, self.S * self.S
        if max_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            max_bytes = int(free * 0.4)
        per = {L.FMT_F64: 8, L.FMT_F32: 4, L.FMT_BF16X2: 4}[fmt]
        xs = int(max(1, min(S, max_bytes // max(1, nfield * self.nb * plane * per))))
        while xs > 1 and (S % xs or (xs * plane) % 128):
            xs -= 1
        if S % xs or (xs * plane) % 128:
            raise ValueError('no x slab of the %d^3 shell mesh fits %d bytes' % (S, max_bytes))
        return xs

    def bispectrum_slabs(self, dks, tri, path=PATH_MFMA, scale=1.0, max_bytes=None, out=None,
                         stream=None, nwg=None, xslab=None):
        """scale * sum_x A_i B_j C_k over the shell mesh for every (i, j, k) of
        tri with the shells made and summed one x slab at a time, so the shell
        fields never exist whole: for shell sets larger than HBM (the
        reference's production 1024^3 / 80-bin runs).  dks: spectra from
        forward(), one (auto) or two (<AAB>: sides A, A, B).  path 'mfma'
        (split-BF16, FP32 spectra) or 'plain' (the spectra's precision).
        Slab sums are added in FP64 in a fixed order (deterministic)."""
        tri = np.ascontiguousarray(np.asarray(tri, dtype=np.int32).reshape(-1, 3))
        T = len(tri)
        nf = len(dks)
        if nf not in (1, 2):
            raise ValueError('one (auto) or two (<AAB>) spectra')
        if T and (tri.min() < 0 or tri.max() >= self.nb):
            raise IndexError('triangle shell index out of range')
        f64 = dks[0].dtype == torch.complex128
        if path == PATH_MFMA:
            if f64:
                raise ValueError('the MFMA path needs FP32 (complex64) spectra')
            fmt = L.FMT_BF16X2
        elif path == PATH_PLAIN:
            fmt = L.FMT_F64 if f64 else L.FMT_F32
        else:
            raise ValueError('path must be mfma or plain')
        dt = L.DT_F64 if f64 else L.DT_F32
        if out is None:
            out = torch.zeros(T, dtype=torch.float64, device=self.device)
        else:
            out.zero_()
        if T == 0:
            return out
        S, nb = self.S, self.nb
        xs = int(xslab) if xslab else self.slab_planes(nf, fmt, max_bytes)
        if S % xs or (xs * S * S) % 128:
            raise ValueError('x slab of %d planes does not tile the %d^3 shell mesh' % (xs, S))
        cells = xs * S * S
        st = stream_handle(stream)
        nbytes = L.size_out()
        L.call('bsk_shells_workspace', self._plan, dt, nb, xs, ctypes.byref(nbytes))
        works = [self.work.get('slab_w%d' % f, nbytes.value) for f in range(nf)]
        sets = []  # one slab of every shell per field, C order
        for f in range(nf):
            if fmt == L.FMT_BF16X2:
                hi = torch.empty((nb, cells), dtype=torch.int16, device=self.device)
                lo = torch.empty((nb, cells), dtype=torch.int16, device=self.device)
                sets.append(ShellSet(fmt, hi, lo, cells))
            else:
                sets.append(ShellSet(fmt, torch.empty((nb, cells), dtype=torch.float64 if f64
                                                      else torch.float32, device=self.device),
                                     None, cells))
        # inverse x of every shell once per field (kept in its workspace)
        for f in range(nf):
            L.call('bsk_shells', self._plan, dt, L.KIND_DELTA, L.ptr(dks[f]), 0, nb, 0, xs,
                   L.STAGE_X, fmt, None, None, cells, 0, L.ptr(works[f]), st)
        part = torch.empty(T, dtype=torch.float64, device=self.device)
        if path == PATH_MFMA:
            plan = self.pair_plan(tri, nb, _MODE[nf], levels=False)
            nwg = int(nwg) if nwg else int(max(1, min(1024, cells // 128 // 8)))
            L.call('bsk_triples_mfma_workspace', plan.handle, nwg, ctypes.byref(nbytes))
            wm = self.work.get('slab_mfma', nbytes.value)
        for x0 in range(0, S, xs):
            for f in range(nf):
                sh = sets[f]
                L.call('bsk_shells', self._plan, dt, L.KIND_DELTA, L.ptr(dks[f]), 0, nb, x0, xs,
                       L.STAGE_YZ, fmt, L.ptr(sh.data), L.ptr(sh.lo), cells, 0,
                       L.ptr(works[f]), st)
            if path == PATH_MFMA:
                a, c = sets[0], sets[-1]
                L.call('bsk_triples_mfma', plan.handle, L.ptr(a.data), L.ptr(a.lo),
                       L.ptr(a.data), L.ptr(a.lo), L.ptr(c.data), L.ptr(c.lo), cells, cells,
                       nwg, float(scale), L.ptr(part), L.ptr(wm), st)
            else:
                self.triples(sets, tri, path=PATH_PLAIN, scale=scale, out=part, stream=stream)
            with torch.cuda.stream(stream if stream is not None else
                                   torch.cuda.current_stream(self.device)):
                out += part
        return out

    # ------------------------------------------------------------- grid info
    def grid_info(self, tri, max_bytes=None, stream=None):
        """(N_tri, kbar) for each triangle (i, j, k) of bin indices, FP64.

        The unit and |k| shells are made in FP64 x-slab by x-slab and reduced
        with the tiled FP64 kernel; N_tri is the rounded sum / S^3 (exact
        integer), kbar = sum(|k| n n) / S^3 / N_tri (NaN where N_tri == 0).
        """
        tri = np.asarray(tri, dtype=np.int32).reshape(-1, 3)
        T = len(tri)
        if T == 0:
            return np.zeros(0), np.zeros((0, 3))
        if tri.min() < 0 or tri.max() >= self.nb:
            raise IndexError('triangle bin index out of range')
        N, nb = self.S, self.nb
        if max_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            max_bytes = int(free * 0.5)
        plane = N * N * 8 * 2 * nb
        xs = int(max(1, min(N, max_bytes // max(plane, 1))))
        while N % xs:
            xs -= 1
        s = stream_handle(stream)
        nbytes = L.size_out()
        L.call('bsk_shells_workspace', self._plan, L.DT_F64, nb, xs, ctypes.byref(nbytes))
        wu = self.work.get('gi_u', nbytes.value)
        wk = self.work.get('gi_k', nbytes.value)
        scells = xs * N * N
        buf = torch.empty((2 * nb, scells), dtype=torch.float64, device=self.device)
        tri_d = self.triangles_device(tri)
        nbatch = (T + 511) // 512
        nchunk = int(max(1, min(scells // 32, 2048 // nbatch)))
        L.call('bsk_triples_tiled_workspace', T, 4, nchunk, ctypes.byref(nbytes))
        wt = self.work.get('gi_t', nbytes.value)
        total = torch.zeros(4 * T, dtype=torch.float64, device=self.device)
        part = torch.empty(4 * T, dtype=torch.float64, device=self.device)
        L.call('bsk_shells', self._plan, L.DT_F64, L.KIND_UNIT, None, 0, nb, 0, xs, L.STAGE_X,
               L.FMT_F64, None, None, 0, 0, L.ptr(wu), s)
        L.call('bsk_shells', self._plan, L.DT_F64, L.KIND_KMAG, None, 0, nb, 0, xs, L.STAGE_X,
               L.FMT_F64, None, None, 0, 0, L.ptr(wk), s)
        for x0 in range(0, N, xs):
            L.call('bsk_shells', self._plan, L.DT_F64, L.KIND_UNIT, None, 0, nb, x0, xs,
                   L.STAGE_YZ, L.FMT_F64, L.ptr(buf), None, scells, 0, L.ptr(wu), s)
            L.call('bsk_shells', self._plan, L.DT_F64, L.KIND_KMAG, None, 0, nb, x0, xs,
                   L.STAGE_YZ, L.FMT_F64, L.ptr(buf[nb:]), None, scells, 0, L.ptr(wk), s)
            L.call('bsk_triples_tiled', L.DT_F64, L.ptr(buf), scells, 2 * nb, scells, T,
                   L.ptr(tri_d), 4, nchunk, 1.0, L.ptr(part), L.ptr(wt), s)
            total += part
        sums = (total / float(self.ncell)).view(T, 4).cpu().numpy()
        ntri = np.rint(sums[:, 0])
        kbar = np.full((T, 3), np.nan)
        nz = ntri != 0
        kbar[nz] = sums[nz, 1:4] / ntri[nz, None]
        return ntri, kbar

    # -------------------------------------------------------- power spectrum
    def power_spectrum(self, dk, stream=None, work=None):
        """(P, N_modes, k_mean) per bin from the compact delta(k) cube (FP64)."""
        dt = L.DT_F32 if dk.dtype == torch.complex64 else L.DT_F64
        nbytes = L.size_out()
        L.call('bsk_power_spectrum_workspace', self._plan, ctypes.byref(nbytes))
        w = (work or self.work).get('pk', nbytes.value)
        out = torch.empty(3 * self.nb, dtype=torch.float64, device=self.device)
        L.call('bsk_power_spectrum', self._plan, dt, L.ptr(dk), L.ptr(out), L.ptr(w),
               stream_handle(stream))
        s = out.view(self.nb, 3).cpu().numpy()
        nmodes = s[:, 1]
        with np.errstate(divide='ignore', invalid='ignore'):
            p = np.where(nmodes > 0, self.box ** 3 * s[:, 0] / nmodes, np.nan)
            kmean = np.where(nmodes > 0, s[:, 2] / nmodes, np.nan)
        return p, nmodes, kmean

    # ---------------------------------------------------------------- sizing
    def uses_levels(self, fmt=L.FMT_BF16X2):
        """Whether default shell sets of this format use the sublattice level
        layout (and MFMA sums over them the level plan)."""
        return fmt == L.FMT_BF16X2 and self.levels and self.layout is not None

    def shell_bytes(self, fmt, nb=None):
        """Device bytes of a default shell set (alloc_shells) of nb shells."""
        nb = self.nb if nb is None else nb
        per = 8 if fmt == L.FMT_F64 else 4
        stride = self.layout.stride if self.uses_levels(fmt) else self.ncell
        return nb * stride * per


def cic_paint(positions, nmesh, box_size, to_contrast=True, out=None, stream=None, work=None):
    """CIC-paint (npart, 3) float32 device positions onto an nmesh^3 mesh
    (deterministic: fixed-point integer sums per node).  work: an optional
    Workspace whose 'cic' buffer is reused between calls."""
    require_device()
    pos = positions.contiguous()
    if pos.dtype != torch.float32 or pos.dim() != 2 or pos.shape[1] != 3:
        raise ValueError('positions must be a (npart, 3) float32 tensor')
    if out is None:
        out = torch.empty((nmesh,) * 3, dtype=torch.float32, device=pos.device)
    nbytes = L.size_out()
    L.call('bsk_cic_workspace', int(nmesh), pos.shape[0], ctypes.byref(nbytes))
    if work is None:
        w = torch.empty(max(int(nbytes.value), 256), dtype=torch.uint8, device=pos.device)
    else:
        w = work.get('cic', nbytes.value)
    L.call('bsk_cic_paint', int(nmesh), float(box_size), pos.shape[0], L.ptr(pos), L.ptr(out),
           1 if to_contrast else 0, L.ptr(w), stream_handle(stream))
    return out
