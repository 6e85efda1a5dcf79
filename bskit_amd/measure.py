"""Command-line fast measurement, the counterpart of
scripts/measure/measure_bs_fast.py (argument names and meaning kept).

    python -m bskit_amd.measure npy_grid grid.npy 512 512 1000. 0 3655 bk.dat \\
        --k_bin_info 0.00628 0.2 0.01257 --meas_type unnorm_b_value

Inputs (``snap_type``): ``npy_grid`` (a real Nmesh^3 array saved with
numpy.save), ``npy_particles`` (an (n, 3) array of positions in units of
``pos_units_mpcoverh`` Mpc/h, CIC-painted on the GPU with the window
compensated, as to_mesh(window='cic', compensated=True) of
measure_bs_fast.py:209-217).  The reference's gadget / illustris /
bigfile_grid readers belong to nbodykit and are not provided; convert such
snapshots to .npy first.

``-cic`` divides a grid's delta(k) by the CIC window of ``nmesh_cic``
(CompensateCICShotnoiseNgrid, measure_bs_fast.py:45-57, :218-223).
Triangles [start_bin_index, end_bin_index) are measured in blocks and each
block's rows are appended and flushed to ``out_file`` (the reference
appends every 100 rows, main.py:1913-1929); ``--resume`` continues after
the last complete row of an existing ``out_file``, so an interrupted run
loses at most one block.
"""
import argparse
import os
import sys
import time

import numpy as np

from . import io as bio
from .main import FFTBispectrum
from .mesh import ArrayMesh, ParticleMesh


def _status(t0, msg):
    print('[%8.2fs] %s' % (time.time() - t0, msg), flush=True)


def parser():
    p = argparse.ArgumentParser(description='FFT bispectrum measurement on one GPU')
    p.add_argument('snap_type', choices=['npy_grid', 'npy_particles', 'bigfile_grid', 'gadget',
                                         'illustris'], help='type of input')
    p.add_argument('snap_prefix1', help='path of the input for mesh 1')
    p.add_argument('nmesh', type=int, help='Nmesh for FFTs')
    p.add_argument('nmesh_cic', type=int, help='Nmesh to be used for CIC correction')
    p.add_argument('boxsize', type=float, help='BoxSize in Mpc/h')
    p.add_argument('start_bin_index', type=int, help='first bin index to use')
    p.add_argument('end_bin_index', type=int, help='last bin index to use (exclusive)')
    p.add_argument('out_file', help='file path to output file')
    p.add_argument('--k_bin_file', help='file with k bin definitions (6 x T, loaded transposed)')
    p.add_argument('--k_bin_info', type=float, nargs=3,
                   help='kmin, kmax and bin width dk in h/Mpc')
    p.add_argument('--high_k_bin_info', type=float, nargs=2,
                   help='number of bins with smaller dk, and dk for higher bins')
    p.add_argument('-pu', '--pos_units_mpcoverh', type=float, help='position units in Mpc/h')
    p.add_argument('-cic', '--cic_correction', action='store_true',
                   help='apply the CIC window correction to input grids')
    p.add_argument('--snap_prefix2', help='input for mesh 2')
    p.add_argument('--snap_prefix3', help='input for mesh 3')
    p.add_argument('--triangle_type', choices=['all', 'equilateral', 'squeezed', 'isosceles'])
    p.add_argument('--squeezed_bin_index', type=int)
    p.add_argument('--isos_mult', type=float)
    p.add_argument('--isos_tol', type=float)
    p.add_argument('--meas_type', choices=['grid_info', 'unnorm_b_value'],
                   default='unnorm_b_value')
    p.add_argument('--resume', action='store_true',
                   help='continue after the last complete row of out_file')
    p.add_argument('--block', type=int, default=100, help='triangles per appended block')
    p.add_argument('--precision', choices=['bf16x3', 'fp32', 'fp64'], default=None)
    p.add_argument('--stop-after', type=int, default=0, help=argparse.SUPPRESS)
    return p


def load_mesh(kind, path, nmesh, box, pos_fac, cic, nmesh_cic):
    if kind in ('bigfile_grid', 'gadget', 'illustris'):
        raise NotImplementedError('%s input needs the nbodykit readers; save the grid or the '
                                  'particle positions with numpy.save and use npy_grid / '
                                  'npy_particles' % kind)
    if kind == 'npy_grid':
        arr = np.load(path, mmap_mode='r', allow_pickle=False)
        if arr.shape != (nmesh,) * 3:
            raise ValueError('grid %s has shape %s, expected Nmesh=%d cubed'
                             % (path, arr.shape, nmesh))
        m = ArrayMesh(np.ascontiguousarray(arr), BoxSize=box)
        if cic:
            m.compensation_window = int(nmesh_cic)
        return m
    pos = np.load(path, allow_pickle=False)
    return ParticleMesh(np.asarray(pos, dtype=np.float32), Nmesh=nmesh, BoxSize=box / pos_fac,
                        compensated=True)


def main(argv=None):
    args = parser().parse_args(argv)
    t0 = time.time()
    _status(t0, 'Starting script')
    if bool(args.k_bin_file) == bool(args.k_bin_info):
        raise RuntimeError('Must specify either a k-bin file or kmin/kmax/dk!')
    pos_fac = args.pos_units_mpcoverh if args.pos_units_mpcoverh is not None else 1.0
    triangle_type = args.triangle_type or 'all'
    if triangle_type == 'isosceles' and args.isos_mult is None:
        raise ValueError('Need to set isos_mult for isosceles triangles!')
    if args.snap_prefix3 and not args.snap_prefix2:
        raise ValueError('need second mesh if third mesh is input!')
    meshes = [load_mesh(args.snap_type, pth, args.nmesh, args.boxsize, pos_fac,
                        args.cic_correction, args.nmesh_cic)
              for pth in (args.snap_prefix1, args.snap_prefix2, args.snap_prefix3) if pth]
    _status(t0, 'Finished input (%d mesh%s)' % (len(meshes), 'es' if len(meshes) > 1 else ''))
    grid_only = args.meas_type == 'grid_info'
    common = dict(Nmesh=args.nmesh, BoxSize=np.ones(3) * args.boxsize,
                  pos_units_mpcoverh=pos_fac,
                  second=meshes[1] if len(meshes) > 1 else None,
                  third=meshes[2] if len(meshes) > 2 else None,
                  for_grid_info_only=grid_only, precision=args.precision)
    if args.k_bin_file:
        fb = FFTBispectrum(meshes[0], k_edges=np.loadtxt(args.k_bin_file).T, **common)
    else:
        kmin, kmax, dk = args.k_bin_info
        nlow, dkh = 0, -1.0
        if args.high_k_bin_info is not None:
            nlow, dkh = int(args.high_k_bin_info[0]), args.high_k_bin_info[1]
        fb = FFTBispectrum(meshes[0], dk=dk, kmin=kmin, kmax=kmax, num_lowk_bins=nlow,
                           dk_high=dkh, triangle_type=triangle_type,
                           squeezed_bin_index=args.squeezed_bin_index or 0,
                           isos_mult=args.isos_mult or 0,
                           isos_tol=args.isos_tol if args.isos_tol is not None else 0.1, **common)
    ntri = len(fb.k_edges)
    start = max(0, args.start_bin_index)
    end = min(ntri, args.end_bin_index)
    if args.resume and os.path.exists(args.out_file):
        bio.truncate_partial(args.out_file)
        nxt = bio.resume_index(args.out_file)
        if nxt > start:
            _status(t0, 'Resuming after triangle %d' % (nxt - 1))
            start = nxt
    _status(t0, 'Measuring triangles [%d, %d) of %d' % (start, end, ntri))
    done = 0
    for i0 in range(start, end, max(1, args.block)):
        i1 = min(end, i0 + max(1, args.block))
        if grid_only:
            fb.measure_gridinfo_faster(imin=i0, imax=i1, out_file=args.out_file)
        else:
            fb.measure_bispectrum_faster(imin=i0, imax=i1, out_file=args.out_file)
        done += 1
        if args.stop_after and done >= args.stop_after:
            _status(t0, 'Stopping after %d blocks (simulated interruption)' % done)
            sys.stdout.flush()
            os._exit(3)
    _status(t0, 'Finished FFT bispectrum')
    return 0


if __name__ == '__main__':
    sys.exit(main())
