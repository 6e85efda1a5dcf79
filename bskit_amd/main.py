"""FFTBispectrum: drop-in for bskit's FFT bispectrum estimator on MI355X.

Same class, constructor arguments, measurement methods, result dict and
output files as the reference (bskit/main.py:1392-2161), and the same
module functions (main.py:227-831); the arithmetic runs in libbskhip's
gfx950 kernels:

* the input mesh is painted to HBM once (array upload or GPU CIC);
* delta(k) comes from a pruned forward FFT that keeps only the modes up to
  the largest bin edge (normalised 1/N^3 like pmesh's paint(mode='complex'));
* all nb shell fields of a bin set come from one masked inverse-FFT launch
  sequence (replacing nb view/mask/paint(mode='real') calls);
* the triangle sums run in the input's precision by default, as the
  reference's numpy sums do (main.py:1875): FP32 sweeps ('fp32') for FP32
  input, FP64 sweeps ('fp64') for FP64 input; precision='bf16x3' selects
  the split-BF16 MFMA GEMMs (FP32 accumulation, relative error ~1e-5, the
  fast path); every path ends in a fixed-order FP64 reduction;
* N_tri and k-means come from FP64 unit and |k| shells (exact integer N_tri).

Deliberate differences from the reference, each documented where it occurs:
shell membership is decided on integer |n|^2 (ties on a bin edge belong to
both bins); N_tri is rounded to the exact integer and k-means are NaN where
N_tri == 0; cross spectra (<AAB>, <ABC>) also work in the fast methods;
imin/imax default to the full list; a k_edges input also works with the fast
methods (bins are derived from the edge rows); both fast methods may be used
on one object (k_mean is always (T, 3)); single-row output files are written.
"""
import logging
import math
import warnings

import numpy as np
import torch

from . import _lib as L
from . import io as bio
from .bins import (generate_bin_edge_list, generate_equilateral_triangle_bin_list,
                   generate_isosceles_triangle_bin_list, generate_squeezed_triangle_bin_list,
                   generate_triangle_bin_list, geometric_k_mean)
from .engine import GridEngine, require_device
from .mesh import ArrayMesh, MeshSource, ParticleMesh, cast_source

__all__ = ['FFTBispectrum', 'k_field', 'number_field', 'compute_Nbin', 'compute_k_means_on_grid',
           'compute_bk_FFT_value', 'bk_FFT_full', 'bk_FFT_grid_info', 'bk_FFT_unnormalized_value',
           'bk_binned', 'ArrayMesh', 'ParticleMesh', 'MeshSource',
           'PRECISIONS']

PRECISIONS = ('bf16x3', 'fp32', 'fp64')

# --------------------------------------------------------------------- engines
_ENGINES = {}
_ENGINE_LIMIT = 8


def _bins_key(kLU):
    return tuple((float(a), float(b)) for a, b in np.asarray(kLU, dtype=np.float64))


def get_engine(nmesh, box, kLU, device=None, shell_mesh=None):
    """Shared GridEngine per (Nmesh, BoxSize, bins, shell mesh) key (LRU).

    shell_mesh None picks the smallest exact shell mesh; Nmesh keeps the
    shells on the input mesh (needed where the shell fields themselves are
    returned).  The CIC window compensation is chosen per forward transform
    (each field's own), not per engine.
    """
    dev = require_device(device)
    key = (int(nmesh), float(box), _bins_key(kLU), str(dev),
           None if shell_mesh is None else int(shell_mesh))
    eng = _ENGINES.pop(key, None)
    if eng is None:
        kLU = np.asarray(kLU, dtype=np.float64)
        eng = GridEngine(nmesh, box, kLU[:, 0], kLU[:, 1], device=dev, shell_mesh=shell_mesh)
        eng.gridinfo_cache = {}
    _ENGINES[key] = eng
    while len(_ENGINES) > _ENGINE_LIMIT:
        old = next(iter(_ENGINES))
        _ENGINES.pop(old).close()
    return eng


def clear_engines():
    for e in _ENGINES.values():
        e.close()
    _ENGINES.clear()


def _unique_bins(edges6):
    """Unique (kL, kU) bins of (T, 6) triangle edges and per-side indices."""
    e = np.asarray(edges6, dtype=np.float64).reshape(-1, 6)
    sides = e.reshape(-1, 3, 2)
    flat = sides.reshape(-1, 2)
    uniq, inv = np.unique(flat, axis=0, return_inverse=True)
    return uniq, inv.reshape(-1, 3).astype(np.int64)


def _grid_info(eng, tri):
    """Cached exact N_tri and k-means (k_f units handled by the engine)."""
    tri = np.asarray(tri, dtype=np.int64).reshape(-1, 3)
    key = tri.tobytes()
    cache = getattr(eng, 'gridinfo_cache', None)
    if cache is not None and key in cache:
        return cache[key]
    res = eng.grid_info(tri)
    if cache is not None:
        cache[key] = res
    return res


class FFTBispectrum:
    """Bispectrum measurements from one grid (or 2-3 grids for cross spectra).

    Parameters follow bskit/main.py:1485-1506; see the reference docstring
    for their meaning.  Extra keyword-only parameters:

    precision : {'bf16x3', 'fp32', 'fp64'}, optional
        triple-product arithmetic: split-BF16 MFMA with FP32 accumulation
        (the fast path), FP32 sweeps (default for float32 input, the
        reference's own arithmetic) or FP64 (default for float64 input).
        Final reductions are FP64 in every case.
    device : torch device, optional
    """
    logger = logging.getLogger('FFTBispectrum')
    # shell sets above this fraction of free HBM are made and summed x slab
    # by x slab instead of resident
    slab_threshold = 0.6

    def __init__(self, source, Nmesh=None, BoxSize=None, dk=None, kmin=None, kmax=None,
                 num_lowk_bins=0, dk_high=-1., dmu=None, mu_min=None, mu_max=None,
                 pos_units_mpcoverh=1., k_edges=None, second=None, third=None,
                 triangle_type='all', isos_mult=0, isos_tol=0.1, squeezed_bin_index=0,
                 for_grid_info_only=False, *, precision=None, device=None):
        self.first = cast_source(source, Nmesh, BoxSize)
        self.mesh = self.first
        self.attrs = {'Nmesh': self.first.attrs['Nmesh'].copy(),
                      'BoxSize': self.first.attrs['BoxSize'].copy()}
        self.second = None
        self.third = None
        self.num_fields = 1
        if second is not None:
            self.second = cast_source(second, Nmesh, BoxSize)
            self.num_fields += 1
        if third is not None:
            if second is None:
                raise ValueError('Need second source defined if third source is defined!')
            self.third = cast_source(third, Nmesh, BoxSize)
            self.num_fields += 1
        for other, name in ((self.second, 'second'), (self.third, 'third')):
            if other is None:
                continue
            if not np.array_equal(other.attrs['BoxSize'], self.attrs['BoxSize']):
                raise ValueError('BoxSize mismatch between first and %s mesh' % name)
            if not np.array_equal(other.attrs['Nmesh'], self.attrs['Nmesh']):
                raise ValueError('Nmesh mismatch between first and %s mesh' % name)

        if dk is None:
            dk = 2 * np.pi / self.attrs['BoxSize'].min()
        if (kmin is None or kmax is None) and k_edges is None:
            raise ValueError('Must specify either {kmin,kmax} values or k_edges array!')
        self.attrs.update(dk=dk, kmin=kmin, kmax=kmax, dmu=dmu, mu_min=mu_min, mu_max=mu_max,
                          pos_units_mpcoverh=pos_units_mpcoverh, triangle_type=triangle_type,
                          num_lowk_bins=num_lowk_bins, dk_high=dk_high, isos_mult=isos_mult,
                          isos_tol=isos_tol, squeezed_bin_index=squeezed_bin_index)

        self.k_indices = None
        if k_edges is not None:
            self.k_edges = np.asarray(k_edges, dtype=np.float64)
        else:
            args = dict(num_lowk_bins=num_lowk_bins, dk_high=dk_high)
            if triangle_type == 'all':
                kw = dict(kmin=kmin, kmax=kmax, dk=dk, dmu=dmu, mu_min=mu_min, mu_max=mu_max,
                          num_fields=self.num_fields, **args)
                self.k_edges = generate_triangle_bin_list(**kw)
                self.k_indices = generate_triangle_bin_list(return_indices=True, **kw)
            elif triangle_type == 'equilateral':
                self.k_edges = generate_equilateral_triangle_bin_list(kmin, kmax, dk, **args)
                self.k_indices = generate_equilateral_triangle_bin_list(kmin, kmax, dk, **args,
                                                                        return_indices=True)
            elif triangle_type == 'squeezed':
                self.k_edges = generate_squeezed_triangle_bin_list(
                    kmin, kmax, dk, squeezed_bin_index=squeezed_bin_index, **args)
                self.k_indices = generate_squeezed_triangle_bin_list(
                    kmin, kmax, dk, squeezed_bin_index=squeezed_bin_index, **args,
                    return_indices=True)
            elif triangle_type == 'isosceles':
                self.k_edges = generate_isosceles_triangle_bin_list(
                    kmin, kmax, dk, isos_mult=isos_mult, isos_tol=isos_tol, **args)
                self.k_indices = generate_isosceles_triangle_bin_list(
                    kmin, kmax, dk, isos_mult=isos_mult, isos_tol=isos_tol, **args,
                    return_indices=True)
            else:
                raise ValueError("triangle_type must be 'all', 'equilateral', 'isosceles' or "
                                 "'squeezed'")

        self.b = None
        self._device = device
        self._fields = None
        self._dk = {}
        self._shells = {}
        if precision is None:
            any64 = any(m.dtype == torch.float64 for m in self._meshes())
            precision = 'fp64' if any64 else 'fp32'
        if precision not in PRECISIONS:
            raise ValueError('precision must be one of %s' % (PRECISIONS,))
        self.precision = precision
        if not for_grid_info_only:
            self._paint_meshes()
        else:
            self.attrs['painted'] = False

    # ----------------------------------------------------------- state
    def __getstate__(self):
        return dict(b=self.b, k_edges=self.k_edges, attrs=self.attrs, precision=self.precision)

    def __setstate__(self, state):
        self.attrs = state['attrs']
        self.k_edges = state['k_edges']
        self.b = state['b']
        self.k_indices = None
        self._fields = None
        self._dk = {}
        self._shells = {}
        self._device = None
        self.precision = state.get('precision', 'fp32')

    def _meshes(self):
        return [m for m in (self.first, self.second, self.third) if m is not None]

    def _paint_meshes(self):
        """Put the input field(s) in HBM (upload or GPU CIC); bskit/main.py:1608-1621."""
        dev = require_device(self._device)
        self._fields = [m.device_field(dev) for m in self._meshes()]
        if self.precision != 'fp64':
            self._fields = [f.to(torch.float32) if f.dtype != torch.float32 else f
                            for f in self._fields]
        else:
            self._fields = [f.to(torch.float64) for f in self._fields]
        self._dk = {}
        self._shells = {}
        self.attrs['painted'] = True

    def release(self):
        """Free the device caches (painted fields, spectra, shells)."""
        self._fields = None
        self._dk = {}
        self._shells = {}
        self.attrs['painted'] = False

    def set_k_edges(self, k_edges):
        """Replace the triangle-bin edge array (bskit/main.py:1624-1634).

        The bin indices are re-derived from the new edges (the reference keeps
        the old k_indices, which the fast methods would then misuse).
        """
        self.k_edges = np.asarray(k_edges, dtype=np.float64)
        self.k_indices = None

    @property
    def bin_edges(self):
        """(nbin, 2) k bins used by the fast methods."""
        return self._fast_bins()[0]

    # ----------------------------------------------------------- helpers
    @property
    def _N(self):
        return int(self.attrs['Nmesh'][0])

    @property
    def _L(self):
        return float(self.attrs['BoxSize'][0])

    def _slice(self, imin, imax):
        n = len(self.k_edges)
        imin = 0 if imin is None else int(imin)
        imax = n if imax is None else int(imax)
        return imin, min(imax, n)

    def _fast_bins(self):
        """k bins and triangle indices of the fast methods (main.py:1828-1839)."""
        if self.k_indices is not None and self.attrs.get('kmin') is not None:
            kLU = generate_bin_edge_list(self.attrs['kmin'], self.attrs['kmax'], self.attrs['dk'],
                                         self.attrs['num_lowk_bins'], self.attrs['dk_high'])
            return kLU, np.asarray(self.k_indices, dtype=np.int64).reshape(-1, 3)
        return _unique_bins(self.k_edges)

    def _engine(self, kLU):
        return get_engine(self._N, self._L, kLU, self._device)

    def _shell_format(self, kLU, tri):
        fmt = {'bf16x3': L.FMT_BF16X2, 'fp32': L.FMT_F32, 'fp64': L.FMT_F64}[self.precision]
        if fmt == L.FMT_BF16X2 and not GridEngine.mfma_supported(len(kLU), len(self._fields), tri):
            warnings.warn('%d bins with %d fields exceed the MFMA kernel (128 shells, 96 for '
                          '<ABC>): using the FP32 path' % (len(kLU), len(self._fields)))
            fmt = L.FMT_F32
        return fmt

    def _shellsets(self, eng, kLU, tri):
        if not self.attrs.get('painted') or self._fields is None:
            raise RuntimeError('meshes are not painted (for_grid_info_only=True)')
        fmt = self._shell_format(kLU, tri)
        key = _bins_key(kLU)
        sets = []
        meshes = self._meshes()
        for fi, f in enumerate(self._fields):
            # one set of shells per field; keep the last bin set only.  Each
            # mesh is deconvolved by its own CIC window (measure_bs_fast.py
            # compensates each mesh it reads, :209-223), none for a plain
            # array.
            win = int(meshes[fi].compensation_window)
            skey = (key, fi, fmt, win)
            s = self._shells.get(skey)
            if s is None:
                for k in [k for k in self._shells if k[1] == fi]:
                    del self._shells[k]
                dk = eng.forward(f, window=win)
                s = eng.shells(dk, kind=L.KIND_DELTA, fmt=fmt,
                               dtype=torch.float32 if fmt != L.FMT_F64 else torch.float64)
                self._shells[skey] = s
            sets.append(s)
        path = 'mfma' if fmt == L.FMT_BF16X2 else 'plain'
        return sets, path

    def _unnorm_sums(self, kLU, tri, factor):
        """factor * sum_x d_i d_j d_k / ncell: the reference's L^6/N^3 sum
        (main.py:1875-1882) evaluated on the engine's shell mesh."""
        eng = self._engine(kLU)
        if not self.attrs.get('painted') or self._fields is None:
            raise RuntimeError('meshes are not painted (for_grid_info_only=True)')
        nf = len(self._fields)
        fmt = self._shell_format(kLU, tri)
        free, _ = torch.cuda.mem_get_info(eng.device)
        if nf <= 2 and nf * eng.shell_bytes(fmt) > self.slab_threshold * free:
            # the shell sets do not fit in HBM: make and sum them x slab by x
            # slab (each field's spectrum kept, shells never whole)
            meshes = self._meshes()
            dks = [eng.forward(f, window=int(m.compensation_window))
                   for f, m in zip(self._fields, meshes)]
            path = 'mfma' if fmt == L.FMT_BF16X2 else 'plain'
            out = eng.bispectrum_slabs(dks, tri, path=path, scale=factor / float(eng.ncell))
            return out.cpu().numpy()
        sets, path = self._shellsets(eng, kLU, tri)
        out = eng.triples(sets, tri, path=path, scale=factor / float(eng.ncell))
        return out.cpu().numpy()

    def _append_b(self, index, edges, kmean, B, ntri):
        new = {'index': np.asarray(index), 'k_edge': np.asarray(edges).reshape(-1, 6),
               'k_mean': np.asarray(kmean, dtype=np.float64).reshape(-1, 3),
               'B': np.asarray(B, dtype=np.float64), 'N_tri': np.asarray(ntri, dtype=np.float64)}
        if self.b is None:
            self.b = {k: v.copy() for k, v in new.items()}
        else:
            for k, v in new.items():
                self.b[k] = np.append(self.b[k], v, axis=0)
        n = len(index)
        return {k: v[len(v) - n:] for k, v in self.b.items()}

    # ------------------------------------------------------- slow path
    def measure_bispectrum(self, imin=None, imax=None, kmeas_min=None, kmeas_max=None,
                           out_file=None, verbose=0, meas_type='full'):
        """Per-triangle B, N_tri and k-means; bskit/main.py:1637-1780.

        Bin edges are scaled by pos_units_mpcoverh before measuring and the
        results converted back (k-means / u, B * u^6), as the reference does.
        """
        if (kmeas_min is not None) and (kmeas_max is not None):
            raise NotImplementedError('kmeas_min/kmeas_max not implemented yet!')
        if meas_type not in ('full', 'grid_info', 'unnorm_b_value'):
            raise ValueError("meas_type must be 'full', 'grid_info' or 'unnorm_b_value'")
        i0, i1 = self._slice(imin, imax)
        tri_edges = np.asarray(self.k_edges[i0:i1], dtype=np.float64).reshape(-1, 6)
        index = np.arange(i0, i1)
        T = len(index)
        u = float(self.attrs['pos_units_mpcoverh'])
        B = np.zeros(T)
        ntri = np.zeros(T)
        kmean = np.zeros((T, 3))
        if T:
            kLU, tri = _unique_bins(tri_edges * u)
            eng = self._engine(kLU)
            if meas_type in ('full', 'grid_info'):
                ntri, kmean = _grid_info(eng, tri)
                kmean = kmean / u
            if meas_type in ('full', 'unnorm_b_value'):
                s = self._unnorm_sums(kLU, tri, self._L ** 6 * u ** 6)
                if meas_type == 'full':
                    with np.errstate(divide='ignore', invalid='ignore'):
                        B = np.where(ntri != 0, s / np.where(ntri != 0, ntri, 1.0), np.nan)
                else:
                    B = s
        if verbose > 0:
            for t in range(T):
                print('FFT Bk: %d %g %g %g %g %g' % (t, kmean[t, 0], kmean[t, 1], kmean[t, 2],
                                                    B[t], ntri[t]))
        if out_file is not None and T:
            kind = {'full': bio.SLOW_FULL, 'grid_info': bio.SLOW_GRIDINFO,
                    'unnorm_b_value': bio.SLOW_UNNORM}[meas_type]
            bio.append_rows(out_file, kind, index, tri_edges, kmean, B, ntri, chunk=1)
        return self._append_b(index, tri_edges, kmean, B, ntri)

    # ------------------------------------------------------- fast path
    def measure_bispectrum_faster(self, imin=None, imax=None, kmeas_min=None, kmeas_max=None,
                                  out_file=None, verbose=0):
        """Unnormalised B for triangles [imin, imax); bskit/main.py:1784-1938.

        All shells of the bin list are made once; B_unnorm =
        L^6/N^3 sum_x d_i d_j d_k * u^6 (bins are not scaled by u, as in the
        reference).  k_mean and N_tri are stored as zeros.
        """
        i0, i1 = self._slice(imin, imax)
        kLU, tri_all = self._fast_bins()
        tri = tri_all[i0:i1]
        edges = np.asarray(self.k_edges[i0:i1], dtype=np.float64).reshape(-1, 6)
        index = np.arange(i0, i1)
        T = len(index)
        u = float(self.attrs['pos_units_mpcoverh'])
        B = np.zeros(T)
        if T:
            B = self._unnorm_sums(kLU, tri, self._L ** 6 * u ** 6)
        if verbose > 0:
            for t in range(T):
                print('FFT Bk: %d %d %g' % (t, index[t], B[t]))
        if out_file is not None and T:
            bio.append_rows(out_file, bio.FAST_B, index, edges, B=B)
        return self._append_b(index, edges, np.zeros((T, 3)), B, np.zeros(T))

    def measure_gridinfo_faster(self, imin=None, imax=None, kmeas_min=None, kmeas_max=None,
                                out_file=None, verbose=0):
        """N_tri and k-means for triangles [imin, imax); bskit/main.py:1941-2132.

        N_tri = sum_x n_i n_j n_k / N^3 (rounded: it is an integer count),
        kbar_1 = sum_x kk_i n_j n_k / N^3 / N_tri etc. from FP64 unit and
        |k| shells, divided by u; NaN where N_tri == 0.
        """
        i0, i1 = self._slice(imin, imax)
        kLU, tri_all = self._fast_bins()
        tri = tri_all[i0:i1]
        edges = np.asarray(self.k_edges[i0:i1], dtype=np.float64).reshape(-1, 6)
        index = np.arange(i0, i1)
        T = len(index)
        ntri = np.zeros(T)
        kmean = np.zeros((T, 3))
        if T:
            eng = self._engine(kLU)
            ntri, kmean = _grid_info(eng, tri)
            kmean = kmean / float(self.attrs['pos_units_mpcoverh'])
        if verbose > 0:
            for t in range(T):
                if (t + 1) % 100 == 0:
                    print('FFT Bk: %d %d %g %g %g %g' % (t, index[t], kmean[t, 0], kmean[t, 1],
                                                         kmean[t, 2], ntri[t]))
        if out_file is not None and T:
            bio.append_rows(out_file, bio.FAST_GRIDINFO, index, edges, kmean, ntri=ntri)
        return self._append_b(index, edges, kmean, np.zeros(T), ntri)

    def save_bispectrum(self, out_file):
        """Write the b dict to an ASCII file; bskit/main.py:2135-2161."""
        bio.save_bispectrum(out_file, self.b)


# ---------------------------------------------------------------------------
# Module functions (bskit/main.py:227-831).  'bins' is a (3, 2) array of
# [kmin, kmax] per side; meshes are anything cast_source accepts.
# ---------------------------------------------------------------------------
def _box_n(box_size, n_mesh):
    b = np.asarray(box_size, dtype=np.float64).reshape(-1)
    n = np.asarray(n_mesh).reshape(-1)
    if len(b) != 3:
        raise ValueError('box_size must be 3-element list!')
    if len(n) != 3:
        raise ValueError('n_mesh must be 3-element list!')
    if not (b[0] == b[1] == b[2] and n[0] == n[1] == n[2]):
        raise NotImplementedError('only cubic meshes are supported')
    return float(b[0]), int(n[0])


def _single_shell(box_size, n_mesh, kmin, kmax, kind, power=None):
    box, N = _box_n(box_size, n_mesh)
    eng = get_engine(N, box, [[kmin, kmax]], shell_mesh=N)
    if power is not None:
        eng.set_kpower(power)
    s = eng.shells(None, kind=kind, fmt=L.FMT_F64, dtype=torch.float64)
    return s.data[0].view(N, N, N)


def k_field(box_size, n_mesh, kmin, kmax, p):
    """c2r of |k|^p on the closed shell [kmin, kmax]; bskit/main.py:227-277.

    Any real power p (the reference weights the shell by |k|**p at
    :264-268); p = 0 is the number field.  Returns an FP64 device tensor of
    shape n_mesh.
    """
    p = float(p)
    if p == 0:
        return _single_shell(box_size, n_mesh, kmin, kmax, L.KIND_UNIT)
    if p == 1:
        return _single_shell(box_size, n_mesh, kmin, kmax, L.KIND_KMAG)
    return _single_shell(box_size, n_mesh, kmin, kmax, L.KIND_KPOW, power=p)


def number_field(box_size, n_mesh, kmin, kmax):
    """c2r of the unit mask of [kmin, kmax]; bskit/main.py:280-329."""
    return _single_shell(box_size, n_mesh, kmin, kmax, L.KIND_UNIT)


def _tri_engine(box_size, n_mesh, bins):
    box, N = _box_n(box_size, n_mesh)
    b = np.asarray([bins[i] for i in range(3)], dtype=np.float64).reshape(3, 2)
    kLU, tri = _unique_bins(b.reshape(1, 6))
    return get_engine(N, box, kLU), tri


def compute_Nbin(box_size, n_mesh, bins, verbose, return_number_fields=False):
    """Number of triangles in a (k1,k2,k3) bin; bskit/main.py:431-499."""
    eng, tri = _tri_engine(box_size, n_mesh, bins)
    ntri, _ = _grid_info(eng, tri)
    if verbose > 0:
        print('Starting Nbin computations')
    if return_number_fields:
        box, N = _box_n(box_size, n_mesh)
        fields = {i: number_field(box_size, n_mesh, bins[i][0], bins[i][1]) for i in range(3)}
        return float(ntri[0]), fields
    return float(ntri[0])


def compute_k_means_on_grid(box_size, n_mesh, bins, verbose, Nbin, number_fields):
    """Mean |k_i| over a triangle bin; bskit/main.py:502-567."""
    eng, tri = _tri_engine(box_size, n_mesh, bins)
    _, kbar = _grid_info(eng, tri)
    return {0: float(kbar[0, 0]), 1: float(kbar[0, 1]), 2: float(kbar[0, 2])}


def _mesh_fb(mesh, bins, second_mesh=None, third_mesh=None, precision=None):
    b = np.asarray([bins[i] for i in range(3)], dtype=np.float64).reshape(1, 6)
    return FFTBispectrum(mesh, k_edges=b, second=second_mesh, third=third_mesh,
                         precision=precision)


def compute_bk_FFT_value(mesh, bins, Nbin=1, verbose=0, second_mesh=None, third_mesh=None):
    """B in one triangle bin for a given N_tri; bskit/main.py:570-672.

    Roles: <AAA> with one mesh, <AAB> with two, <ABC> with three.
    """
    fb = _mesh_fb(mesh, bins, second_mesh, third_mesh)
    r = fb.measure_bispectrum(0, 1, meas_type='unnorm_b_value')
    return float(r['B'][0]) / Nbin


def bk_FFT_full(mesh, bin0, bin1, bin2, verbose=0, second_mesh=None, third_mesh=None,
                approximate_k_means=False):
    """(B, N_tri, k1, k2, k3) in one triangle bin; bskit/main.py:675-750."""
    if third_mesh is not None and second_mesh is None:
        raise ValueError('Must specify second_mesh if specifying third_mesh!')
    fb = _mesh_fb(mesh, {0: bin0, 1: bin1, 2: bin2}, second_mesh, third_mesh)
    r = fb.measure_bispectrum(0, 1, meas_type='full')
    k = r['k_mean'][0]
    if approximate_k_means:
        k = [geometric_k_mean(b[0], b[1]) for b in (bin0, bin1, bin2)]
    return float(r['B'][0]), float(r['N_tri'][0]), float(k[0]), float(k[1]), float(k[2])


def bk_FFT_grid_info(mesh, bin0, bin1, bin2, verbose=0):
    """(N_tri, k1, k2, k3) in one triangle bin; bskit/main.py:753-791."""
    m = cast_source(mesh)
    box = m.attrs['BoxSize']
    n = m.attrs['Nmesh']
    bins = {0: bin0, 1: bin1, 2: bin2}
    eng, tri = _tri_engine(box, n, bins)
    ntri, kbar = _grid_info(eng, tri)
    return float(ntri[0]), float(kbar[0, 0]), float(kbar[0, 1]), float(kbar[0, 2])


def bk_FFT_unnormalized_value(mesh, bin0, bin1, bin2, verbose=0, second_mesh=None,
                              third_mesh=None):
    """Unnormalised B in one triangle bin; bskit/main.py:794-831."""
    return compute_bk_FFT_value(mesh, {0: bin0, 1: bin1, 2: bin2}, Nbin=1, verbose=verbose,
                                second_mesh=second_mesh, third_mesh=third_mesh)


def bk_binned(mesh, bin0, bin1, bin2, verbose=0, periodic=False):
    """Direct triangle enumeration estimator; bskit/main.py:834-1004.

    Enumerates mode triples (k1, k2, k3) in the three bins with
    k1 + k2 + k3 = 0 (exactly, as the reference's closure test; with
    periodic=True modulo N, the FFT estimator's closure) and returns
    (B, N_tri, k1_mean, k2_mean, k3_mean) with B = L^6 sum delta delta delta
    / N_tri (complex; the reference returns the complex sum).
    """
    from .direct import direct_bispectrum
    return direct_bispectrum(mesh, bin0, bin1, bin2, periodic=periodic, verbose=verbose)


# ---------------------------------------------------------------------------
# Power spectrum (bskit/main.py:333-405).
# ---------------------------------------------------------------------------
def power_spectrum(mesh, k_edges, precision=None):
    """Binned P(k), number of modes and mean |k| for each [kmin, kmax] row of
    k_edges (closed bins), from one forward transform of the mesh.

    The reference squares three shell fields per bin and sums over the cells
    (pk_FFT, main.py:333-405); by Parseval the same sums are the mode sums of
    |delta_k|^2, 1 and |k| over the bin, which one reduction over the compact
    spectrum computes for every bin at once (bsk_power_spectrum).  Returns
    float64 arrays (P, Nbin, k_mean).
    """
    m = cast_source(mesh)
    N = int(m.attrs['Nmesh'][0])
    box = float(m.attrs['BoxSize'][0])
    kLU = np.asarray(k_edges, dtype=np.float64).reshape(-1, 2)
    dev = require_device()
    eng = get_engine(N, box, kLU, dev)
    f = m.device_field(dev)
    dtype = torch.float64 if precision == 'fp64' else (torch.float32 if precision else f.dtype)
    return eng.power_spectrum(eng.forward(f.to(dtype), window=int(m.compensation_window)))


def pk_FFT(mesh, kmin, kmax):
    """Mean P(k), number of modes and mean k within one bin; main.py:333-405."""
    p, n, k = power_spectrum(mesh, [[kmin, kmax]])
    return float(p[0]), float(n[0]), float(k[0])
