"""Subbox cuts and per-subbox measurements (bskit/main.py:28-150 and
scripts/measure/measure_subbox_bs_fast.py:245-293).

A subbox of an N^3 mesh in a box L is the n^3 = (N/s)^3 block of cells
starting at cell box_multiindex * N/s, in a box L/s (s = nsub_per_side);
its values are the parent's cells themselves (the reference's
nearest-grid-point readout at integer offsets, field_subbox_pm).  On the
GPU the cut is a strided copy of the device field.  ``measure_subboxes``
runs the fast measurement on each subbox with one shared plan (every
subbox has the same Nmesh, BoxSize and bins) and writes
``<prefix>_subbox<i>.dat`` like the reference script.
"""
import numpy as np
import torch

from .engine import require_device
from .main import FFTBispectrum
from .mesh import ArrayMesh, cast_source


def subbox_multiindex_to_index(multiindex, nsub_per_side):
    """(a, b, c) -> a*n^2 + b*n + c; main.py:28-49."""
    assert len(multiindex) == 3
    return int(multiindex[0] * nsub_per_side ** 2 + multiindex[1] * nsub_per_side +
               multiindex[2])


def subbox_index_to_multiindex(i, nsub_per_side):
    """i -> float array (a, b, c) with i = a*n^2 + b*n + c; main.py:52-80."""
    m = np.zeros(3)
    a, r = divmod(i, nsub_per_side ** 2)
    b, c = divmod(r, nsub_per_side)
    m[0], m[1], m[2] = a, b, c
    return m


def _field_of(source):
    if isinstance(source, (np.ndarray, torch.Tensor)):
        return source, None
    m = cast_source(source)
    return m.device_field(require_device()), m


def field_subbox_pm(box_multiindex, nsub_per_side, source, BoxSize=None):
    """Mesh of the subbox with multi-index (a, b, c); main.py:83-150.

    source: a mesh (its BoxSize is used) or a 3-D array (pass BoxSize).
    Returns an ArrayMesh of shape (N/s)^3 in a box BoxSize/s.
    """
    field, m = _field_of(source)
    N = int(field.shape[0])
    s = int(nsub_per_side)
    if N % s:
        raise ValueError('Nmesh %d is not divisible by nsub_per_side %d' % (N, s))
    box = float(m.attrs['BoxSize'][0]) if m is not None else float(BoxSize)
    n = N // s
    a, b, c = (int(v) * n for v in box_multiindex)
    sub = field[a:a + n, b:b + n, c:c + n]
    sub = sub.contiguous() if isinstance(sub, torch.Tensor) else np.ascontiguousarray(sub)
    return ArrayMesh(sub, BoxSize=box / s)


def field_subbox_slice(field, bounds, BoxSize):
    """Mesh of the cells whose node coordinates x_i = i L/N (in [0, L)) lie
    in [bounds[d][0], bounds[d][1]) on every axis; main.py:153-194."""
    f, m = _field_of(field)
    N = int(f.shape[0])
    box = float(BoxSize if m is None else m.attrs['BoxSize'][0])
    x = np.arange(N) * (box / N)
    idx = [np.nonzero((x >= lo) & (x < hi))[0] for lo, hi in bounds]
    if len(idx) != 3:
        raise NotImplementedError('field_subbox_slice supports 3-D fields')
    if isinstance(f, torch.Tensor):
        sel = f[torch.as_tensor(idx[0], device=f.device)][:, torch.as_tensor(idx[1], device=f.device)]
        sel = sel[:, :, torch.as_tensor(idx[2], device=f.device)].contiguous()
    else:
        sel = f[np.ix_(idx[0], idx[1], idx[2])]
    sizes = [hi - lo for lo, hi in bounds]
    if not (sizes[0] == sizes[1] == sizes[2]):
        raise NotImplementedError('only cubic subboxes are supported')
    return ArrayMesh(sel, BoxSize=sizes[0])


def measure_subboxes(mesh, nsub_per_side, out_file_prefix=None, start=0, end=None, imin=None,
                     imax=None, second=None, third=None, meas_type='unnorm_b_value', **fb_kwargs):
    """Fast measurement on subboxes [start, end] (inclusive, as the reference
    script); returns {index: the b dict slice}.  fb_kwargs go to
    FFTBispectrum (binning, triangle type, ...)."""
    base = [cast_source(x) for x in (mesh, second, third) if x is not None]
    fields = [b.device_field(require_device()) for b in base]
    box = float(base[0].attrs['BoxSize'][0])
    s = int(nsub_per_side)
    end = s ** 3 - 1 if end is None else int(end)
    out = {}
    for ind in range(int(start), end + 1):
        multi = subbox_index_to_multiindex(ind, s)
        subs = [field_subbox_pm(multi, s, f, BoxSize=box) for f in fields]
        fb = FFTBispectrum(subs[0], second=subs[1] if len(subs) > 1 else None,
                           third=subs[2] if len(subs) > 2 else None,
                           for_grid_info_only=meas_type == 'grid_info', **fb_kwargs)
        path = None if out_file_prefix is None else '%s_subbox%d.dat' % (out_file_prefix, ind)
        if meas_type == 'grid_info':
            out[ind] = fb.measure_gridinfo_faster(imin=imin, imax=imax, out_file=path)
        else:
            out[ind] = fb.measure_bispectrum_faster(imin=imin, imax=imax, out_file=path)
    return out
