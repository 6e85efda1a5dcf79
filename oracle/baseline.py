"""Reference-structured CPU path timed as bench.py's cpu_baseline ("port").

Test/benchmark infrastructure only.  It follows the reference's fast path
(bskit/main.py:1784-1938) with the FFTs on the host cores:

1. forward r2c of the field, normalised 1/N^3 (main.py:1608-1621);
2. per k bin, the closed-shell mask and an unnormalised c2r
   (main.py:1846-1861), with scipy.fft using P worker threads;
3. per triangle, sum_x d_i d_j d_k in the input precision, computed by P
   threads that each own an x-slab (the MPI ranks of the reference) and
   summed afterwards (comm.gather + bcast, main.py:1875-1879).

Only a bounded sample runs (all of step 1, a few shells, a few triangles);
the per-grid time is extrapolated linearly in the shell and triangle counts.
"""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.fft as sfft


def cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _n2_half(N):
    f = np.fft.fftfreq(N, 1.0 / N).astype(np.int64)
    fz = np.arange(N // 2 + 1, dtype=np.int64)
    return f[:, None, None] ** 2 + f[None, :, None] ** 2 + fz[None, None, :] ** 2


def time_sample(field, n2lo, n2hi, tri, n_shell_sample=2, n_tri_sample=8, workers=None):
    """Seconds per grid, extrapolated from a timed sample; returns a dict."""
    P = workers or cores()
    N = field.shape[0]
    nb, T = len(n2lo), len(tri)
    t0 = time.perf_counter()
    dk = sfft.rfftn(field, workers=P) / float(N) ** 3
    t_fwd = time.perf_counter() - t0
    n2 = _n2_half(N)
    # shells needed by the sampled triangles
    # the first triangles of the list use the lowest (cheapest to cover) bins
    tri_s = np.asarray(tri)[:min(n_tri_sample, T)]
    need = sorted(set(int(x) for x in tri_s.ravel()))
    shells = {}
    t_sh = 0.0
    for b in need:
        t0 = time.perf_counter()
        m = (n2 >= n2lo[b]) & (n2 <= n2hi[b])
        shells[b] = (sfft.irfftn(np.where(m, dk, 0), s=field.shape, workers=P) *
                     float(N) ** 3).astype(field.dtype)
        t_sh += time.perf_counter() - t0
    t_shell = t_sh / len(need)
    slabs = np.array_split(np.arange(N), P)

    def part(sl, i, j, k):
        a, b_, c = shells[i][sl[0]:sl[-1] + 1], shells[j][sl[0]:sl[-1] + 1], shells[k][sl[0]:sl[-1] + 1]
        return float(np.sum(a * b_ * c))

    t0 = time.perf_counter()
    with ThreadPoolExecutor(P) as ex:
        for (i, j, k) in tri_s:
            sum(ex.map(lambda sl: part(sl, i, j, k), slabs))
    t_tri = (time.perf_counter() - t0) / len(tri_s)
    total = t_fwd + nb * t_shell + T * t_tri
    return {'seconds_per_grid': total, 'forward_s': t_fwd, 'shell_s': t_shell,
            'triangle_s': t_tri, 'cores': P, 'nshell_sampled': len(need),
            'ntri_sampled': len(tri_s)}
