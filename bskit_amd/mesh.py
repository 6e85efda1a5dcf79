"""Input sources for FFTBispectrum.

The reference takes nbodykit CatalogSource / MeshSource objects and casts
them with nbodykit's ``_cast_source`` (bskit/main.py:1509-1524).  nbodykit
cannot exist on an MI355X box, so this module provides the sources the
drop-in accepts:

* a 3-D real array (numpy or torch, float32/float64) -- the painted field;
* ``ArrayMesh(array, BoxSize)`` -- the same with its box attached;
* ``ParticleMesh(positions, Nmesh, BoxSize, compensated=True)`` -- particles
  cloud-in-cell painted on the GPU, with the CIC window compensation of
  ``to_mesh(window='cic', compensated=True)`` (scripts/measure/
  measure_bs_fast.py:209-217) applied in the forward transform;
* any duck-typed mesh with ``attrs['Nmesh']``, ``attrs['BoxSize']`` and
  ``paint(mode='real')`` or ``compute(mode='real')`` returning an array.

Every source resolves to a device tensor on the GPU; attrs follow nbodykit's
layout (3-vectors for Nmesh and BoxSize).
"""
import numpy as np
import torch

from . import engine as E


def _vec3(v, name):
    if v is None:
        return None
    a = np.asarray(v, dtype=np.float64).reshape(-1)
    if a.size == 1:
        a = np.repeat(a, 3)
    if a.size != 3:
        raise ValueError('%s must be a scalar or a 3-vector' % name)
    return a


class MeshSource:
    """Base: attrs with Nmesh/BoxSize and a device field."""

    compensation_window = 0   # nmesh of the CIC window to deconvolve (0 = none)

    def __init__(self, Nmesh, BoxSize):
        self.attrs = {'Nmesh': _vec3(Nmesh, 'Nmesh').astype(int), 'BoxSize': _vec3(BoxSize, 'BoxSize')}

    def device_field(self, device):
        raise NotImplementedError

    @property
    def dtype(self):
        return torch.float32


class ArrayMesh(MeshSource):
    """A painted real field (numpy or torch) in a box."""

    def __init__(self, array, BoxSize, Nmesh=None):
        if array.ndim != 3:
            raise ValueError('the field must be a 3-D array')
        shape = tuple(int(s) for s in array.shape)
        if Nmesh is not None and not np.array_equal(_vec3(Nmesh, 'Nmesh').astype(int), shape):
            raise ValueError('Nmesh %s does not match the field shape %s' % (Nmesh, shape))
        super().__init__(shape, BoxSize)
        self.array = array

    @property
    def dtype(self):
        dt = self.array.dtype
        if dt in (np.float64, torch.float64):
            return torch.float64
        return torch.float32

    def device_field(self, device):
        a = self.array
        if isinstance(a, torch.Tensor):
            t = a.to(device=device)
        else:
            t = torch.from_numpy(np.ascontiguousarray(a)).to(device)
        if t.dtype not in (torch.float32, torch.float64):
            t = t.to(torch.float32)
        return t.contiguous()


class ParticleMesh(MeshSource):
    """Particles CIC-painted on the GPU into delta = rho/rho_mean - 1.

    compensated=True deconvolves the CIC window at this mesh resolution;
    window_nmesh overrides the resolution of that window (as the reference
    script's --nmesh_cic, measure_bs_fast.py:45-57).
    """

    def __init__(self, positions, Nmesh, BoxSize, compensated=True, window_nmesh=None):
        super().__init__(Nmesh, BoxSize)
        n = self.attrs['Nmesh']
        if not (n[0] == n[1] == n[2]):
            raise NotImplementedError('only cubic meshes are supported')
        self.positions = positions
        self.compensation_window = int(window_nmesh or n[0]) if compensated else 0

    def device_field(self, device):
        p = self.positions
        if isinstance(p, torch.Tensor):
            p = p.to(device=device, dtype=torch.float32)
        else:
            p = torch.from_numpy(np.ascontiguousarray(p, dtype=np.float32)).to(device)
        return E.cic_paint(p, int(self.attrs['Nmesh'][0]), float(self.attrs['BoxSize'][0]))


class _DuckMesh(MeshSource):
    def __init__(self, obj):
        super().__init__(obj.attrs['Nmesh'], obj.attrs['BoxSize'])
        self.obj = obj
        self.compensation_window = int(getattr(obj, 'compensation_window', 0) or 0)

    def _array(self):
        for name in ('paint', 'compute'):
            fn = getattr(self.obj, name, None)
            if fn is not None:
                return fn(mode='real')
        raise TypeError('mesh source needs paint(mode="real") or compute(mode="real")')

    def device_field(self, device):
        a = self._array()
        if not isinstance(a, torch.Tensor):
            a = np.asarray(a)
        return ArrayMesh(a, self.attrs['BoxSize']).device_field(device)


def cast_source(source, Nmesh=None, BoxSize=None):
    """Turn an input into a MeshSource (the role of nbodykit's _cast_source)."""
    if isinstance(source, MeshSource):
        m = source
    elif isinstance(source, (np.ndarray, torch.Tensor)):
        if BoxSize is None:
            raise ValueError('BoxSize is required for an array source')
        m = ArrayMesh(source, BoxSize, Nmesh)
    elif hasattr(source, 'attrs') and 'Nmesh' in source.attrs and 'BoxSize' in source.attrs:
        m = _DuckMesh(source)
    else:
        raise TypeError('unsupported source type %s' % type(source).__name__)
    if Nmesh is not None and not np.array_equal(m.attrs['Nmesh'], _vec3(Nmesh, 'Nmesh').astype(int)):
        raise ValueError('Nmesh mismatch: source has %s, requested %s' % (m.attrs['Nmesh'], Nmesh))
    if BoxSize is not None and not np.allclose(m.attrs['BoxSize'], _vec3(BoxSize, 'BoxSize')):
        raise ValueError('BoxSize mismatch: source has %s, requested %s' % (m.attrs['BoxSize'], BoxSize))
    n, b = m.attrs['Nmesh'], m.attrs['BoxSize']
    if not (n[0] == n[1] == n[2]) or not (b[0] == b[1] == b[2]):
        raise NotImplementedError('only cubic meshes (equal Nmesh and BoxSize per axis) are supported')
    return m
