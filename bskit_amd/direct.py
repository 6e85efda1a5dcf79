"""Direct triangle enumeration on the GPU (bk_binned of the reference).

bskit/main.py:834-1004 collects the modes of each bin with np.digitize on
|k|^2, i.e. kmin^2 <= |k|^2 < kmax^2 (half-open), from the half grid of the
complex field plus the conjugate of every mode of Hermitian weight 2
(:936-941), and accepts a triple when |k1+k2+k3|^2 < 1e-4 k_f -- exact
closure on the integer lattice, with the third mode one of the collected
modes of its bin.  Here the mode lists are built on the host from integer
|n|^2 limits with the same half-open semantics and the same mode sets
(``reference_modes``), and the pair loop runs in libbskhip
(bsk_direct_triangles).  With periodic=True the closure is taken modulo N
over the FFT grid's modes, which is what the FFT estimator counts; that form
gives exact N_tri checks of the fast path.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib as L
from .engine import GridEngine, require_device, stream_handle
from .mesh import cast_source

TIE_RTOL = 1e-9


def half_open_limits(kmin, kmax, kf):
    """Integer |n|^2 range of kmin <= |k| < kmax."""
    a = (kmin / kf) ** 2 if kmin > 0 else 0.0
    b = (kmax / kf) ** 2
    lo = int(math.ceil(a * (1 - TIE_RTOL))) if a > 0 else 0
    hi = int(math.ceil(b * (1 - TIE_RTOL))) - 1
    return lo, hi


def reference_modes(N, lo, hi):
    """(M, 3) int32 lattice vectors of the modes bk_binned collects for a bin
    (bskit/main.py:887-941): every stored mode of the half grid (n_x, n_y in
    [-N/2, N/2), 0 <= n_z <= N/2) with lo <= |n|^2 <= hi, plus the conjugate
    -n of each mode of Hermitian weight 2 (0 < n_z < N/2); the planes n_z = 0
    and n_z = N/2 are taken once.  So for n_z < 0 the x and y components run
    over (-N/2, N/2]."""
    if hi < 0:
        return np.zeros((0, 3), dtype=np.int32)
    h = N // 2
    r = np.arange(-h, h + 1)
    x, y, z = np.meshgrid(r, r, r, indexing='ij')
    n2 = x * x + y * y + z * z
    up = (z >= 0) & (x < h) & (y < h)                # stored half grid
    down = (z < 0) & (z > -h) & (x > -h) & (y > -h)  # conjugates of weight-2 modes
    sel = (n2 >= lo) & (n2 <= hi) & (up | down)
    return np.ascontiguousarray(np.stack([x[sel], y[sel], z[sel]], axis=1).astype(np.int32))


def shell_modes(N, lo, hi):
    """(M, 3) int32 lattice vectors with components in [-N/2, N/2) and lo <= |n|^2 <= hi."""
    if hi < 0:
        return np.zeros((0, 3), dtype=np.int32)
    m = min(int(math.isqrt(max(hi, 0))) + 1, N // 2)
    r = np.arange(-m, min(m, N // 2 - 1) + 1)
    x, y, z = np.meshgrid(r, r, r, indexing='ij')
    n2 = x * x + y * y + z * z
    sel = (n2 >= lo) & (n2 <= hi)
    return np.ascontiguousarray(np.stack([x[sel], y[sel], z[sel]], axis=1).astype(np.int32))


def count_triangles(N, lims, periodic=True, stream=None, device=None):
    """Exact closed-triangle count and k sums for three bins given by |n|^2 limits."""
    dev = require_device(device)
    ma = torch.from_numpy(shell_modes(N, *lims[0])).to(dev)
    mb = torch.from_numpy(shell_modes(N, *lims[1])).to(dev)
    nblocks = 1024
    part = torch.zeros((nblocks, 6), dtype=torch.float64, device=dev)
    L.call('bsk_direct_triangles', N, 1 if periodic else 0, L.ptr(ma), ma.shape[0], L.ptr(mb),
           mb.shape[0], int(lims[2][0]), int(lims[2][1]), None, None, None, 0, 0, 0, nblocks,
           L.ptr(part), stream_handle(stream))
    s = part.sum(0).cpu().numpy()
    return int(round(s[0])), s[1:4]


def direct_bispectrum(mesh, bin0, bin1, bin2, periodic=False, verbose=0, second_mesh=None,
                      third_mesh=None):
    """(B, N_tri, k1_mean, k2_mean, k3_mean) by direct enumeration; B complex."""
    m = cast_source(mesh)
    N = int(m.attrs['Nmesh'][0])
    box = float(m.attrs['BoxSize'][0])
    kf = 2 * np.pi / box
    bins = [tuple(map(float, b)) for b in (bin0, bin1, bin2)]
    lims = [half_open_limits(b[0], b[1], kf) for b in bins]
    dev = require_device()
    # spectrum on a compact cube covering all three bins
    kmax = max(b[1] for b in bins)
    eng = GridEngine(N, box, [kf * 0.5], [kmax])
    meshes = [m] + [cast_source(x) for x in (second_mesh, third_mesh) if x is not None]
    cubes = []
    for mm in meshes:
        f = mm.device_field(dev).to(torch.float64)
        cubes.append(eng.forward(f, window=int(mm.compensation_window)))
    if len(cubes) == 1:
        ca = cb = cc = cubes[0]
    elif len(cubes) == 2:
        ca, cb, cc = cubes[0], cubes[0], cubes[1]
    else:
        ca, cb, cc = cubes
    # open closure: the reference's own mode sets (half grid plus conjugates);
    # periodic closure: the FFT grid's modes, k3 taken modulo N
    modes = shell_modes if periodic else reference_modes
    ma = torch.from_numpy(modes(N, *lims[0])).to(dev)
    mb = torch.from_numpy(modes(N, *lims[1])).to(dev)
    nblocks = 1024
    part = torch.zeros((nblocks, 6), dtype=torch.float64, device=dev)
    L.call('bsk_direct_triangles', N, 1 if periodic else 0, L.ptr(ma), ma.shape[0], L.ptr(mb),
           mb.shape[0], int(lims[2][0]), int(lims[2][1]), L.ptr(ca), L.ptr(cb), L.ptr(cc),
           eng.MX, eng.MZ, eng.nlo, nblocks, L.ptr(part), stream_handle())
    s = part.sum(0).cpu().numpy()
    eng.close()
    nbin = int(round(s[0]))
    B = complex(s[4], s[5]) * box ** 6
    km = s[1:4] * kf
    if nbin > 0:
        B /= nbin
        km = km / nbin
    if verbose > 0:
        print('B, Nbin = %g %g' % (B.real, nbin))
    return B, nbin, float(km[0]), float(km[1]), float(km[2])
