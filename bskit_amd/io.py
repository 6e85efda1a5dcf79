"""Text output records of the bispectrum path and the grid-info/B combiner.

Line formats are those of the reference (``%e`` = 7 significant digits):

* fast unnormalised B (bskit/main.py:1913-1917):  ``%d`` + 7 x ``%e``
  index, k1_low, k1_high, k2_low, k2_high, k3_low, k3_high, B_unnorm
* fast grid info (main.py:2098-2106):            ``%d`` + 10 x ``%e``
  index, k1_mean, k2_mean, k3_mean, 6 edges, N_tri
* slow full (main.py:1760-1764):                 ``%d`` + 11 x ``%e``
  index, 3 k-means, 6 edges, B, N_tri
* slow grid info (main.py:1766-1770):            same as fast grid info
* slow unnormalised (main.py:1772):              ``%d %e``
* combined (scripts/process/process_fast_bs_measurement.py:81-82):
  ``%d`` + 11 x ``%.6e`` with a header line.

Files are appended to (the reference opens them in append mode); rows are
written whole and flushed, so an interrupted run leaves only complete rows
and can be resumed from the next index.
"""
import numpy as np

FAST_B = 'fast_b'
FAST_GRIDINFO = 'fast_gridinfo'
SLOW_FULL = 'slow_full'
SLOW_GRIDINFO = 'slow_gridinfo'
SLOW_UNNORM = 'slow_unnorm'

COMBINED_HEADER = ('triangle index [dimensionless] k1_mean [h Mpc^-1] k2_mean [h Mpc^-1] '
                   'k3_mean [h Mpc^-1] k1_min [h Mpc^-1] k1_max [h Mpc^-1] k2_min [h Mpc^-1] '
                   'k2_max [h Mpc^-1] k3_min [h Mpc^-1] k3_max [h Mpc^-1] B_bin [h^-6 Mpc^6] '
                   'N_modes')
SAVE_HEADER = ('k1_mean k2_mean k3_mean k1_low k1_high k2_low k2_high k3_low k3_high '
               '[all h Mpc^-1] B [h^-6 Mpc^6] N_tri')


def format_row(kind, index, edges, kmean=None, B=None, ntri=None):
    e = [float(x) for x in edges]
    if kind == FAST_B:
        return '%d %e %e %e %e %e %e %e\n' % (index, *e, B)
    if kind in (FAST_GRIDINFO, SLOW_GRIDINFO):
        return '%d %e %e %e %e %e %e %e %e %e %e\n' % (index, *kmean, *e, ntri)
    if kind == SLOW_FULL:
        return '%d %e %e %e %e %e %e %e %e %e %e %e\n' % (index, *kmean, *e, B, ntri)
    if kind == SLOW_UNNORM:
        return '%d %e\n' % (index, B)
    raise ValueError(kind)


def append_rows(path, kind, index, edges, kmean=None, B=None, ntri=None, chunk=100):
    """Append rows in chunks of `chunk`, flushing after each chunk."""
    n = len(index)
    with open(path, 'a') as f:
        for c0 in range(0, n, chunk):
            lines = []
            for t in range(c0, min(n, c0 + chunk)):
                lines.append(format_row(kind, index[t], edges[t],
                                        None if kmean is None else kmean[t],
                                        None if B is None else B[t],
                                        None if ntri is None else ntri[t]))
            f.write(''.join(lines))
            f.flush()


def save_bispectrum(out_file, b):
    """Write the b dict as 11 columns (k-means, edges, B, N_tri) with a header.

    bskit/main.py:2135-2161 (the reference hands a zip iterator to savetxt;
    here the columns are stacked into an array first).
    """
    if b is None:
        return
    km = np.asarray(b['k_mean'], dtype=np.float64)
    if km.ndim == 1:
        km = np.repeat(km[:, None], 3, axis=1)
    cols = np.column_stack([km[:, 0], km[:, 1], km[:, 2], np.asarray(b['k_edge'])[:, :6],
                            np.asarray(b['B']), np.asarray(b['N_tri'])])
    np.savetxt(out_file, cols, header=SAVE_HEADER)


def combine(bin_info_file, b_file, k_max, out_file, tol=0.01):
    """Join a grid-info file and an unnormalised-B file into the 12-column record.

    scripts/process/process_fast_bs_measurement.py:36-82: rows up to the
    k_max cut (by k1_low), index and edge checks (1% tolerance),
    B = B_unnorm / N_tri.  Returns the combined array.
    """
    info = np.atleast_2d(np.loadtxt(bin_info_file))
    bv = np.atleast_2d(np.loadtxt(b_file))
    k1min = info[:, 4]
    if k1min[-1] > k_max:
        first_over = k1min[k1min > k_max][0]
        i_from_kmax = np.where(k1min == first_over)[0][0] - 1
    else:
        i_from_kmax = len(k1min)
    i_max = int(min(len(info), len(bv), i_from_kmax))
    out = np.zeros((i_max, 12))
    for i in range(i_max):
        if info[i, 0] != bv[i, 0]:
            raise ValueError('Bin index mismatch between bin_info and B files: line %d: %d vs %d'
                             % (i, info[i, 0], bv[i, 0]))
        for j in range(6):
            if np.abs((info[i, 4 + j] - bv[i, 1 + j]) / info[i, 4 + j]) > tol:
                raise ValueError('Bin bound mismatch between bin_info and B files: line %d: %e vs %e'
                                 % (i, info[i, 4 + j], bv[i, 1 + j]))
        out[i, 0:10] = info[i, 0:10]
        with np.errstate(divide='ignore', invalid='ignore'):
            out[i, 10] = bv[i, 7] / info[i, 10]
        out[i, 11] = info[i, 10]
    np.savetxt(out_file, out, fmt='%d %.6e %.6e %.6e %.6e %.6e %.6e %.6e %.6e %.6e %.6e %.6e',
               header=COMBINED_HEADER)
    return out


def resume_index(path):
    """Next triangle index after the last complete row of an output file."""
    try:
        with open(path) as f:
            last = None
            for line in f:
                if line.endswith('\n') and line.strip() and not line.startswith('#'):
                    last = line
    except FileNotFoundError:
        return 0
    if last is None:
        return 0
    return int(last.split()[0]) + 1


def truncate_partial(path):
    """Drop an unterminated last line (a write cut short); returns bytes kept."""
    try:
        with open(path, 'rb+') as f:
            data = f.read()
            keep = data.rfind(b'\n') + 1
            if keep != len(data):
                f.truncate(keep)
            return keep
    except FileNotFoundError:
        return 0
