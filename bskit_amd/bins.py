"""k-bin and triangle-bin generators of the FFTBispectrum interface.

Same names, arguments, outputs and errors as the reference module functions
(bskit/main.py:408-428, :1008-1389).  The triangle lists are produced
directly in their final order: because lower bin edges increase with the bin
index, sorting rows by (k1_low, k2_low, k3_low) as the reference does
(main.py:1380-1381) is lexicographic order of the bin-index triples, so the
index triples are enumerated in that order and the edges gathered from them.
"""
import numpy as np

__all__ = ['geometric_k_mean', 'generate_bin_edge_list', 'generate_equilateral_triangle_bin_list',
           'generate_squeezed_triangle_bin_list', 'generate_isosceles_triangle_bin_list',
           'generate_triangle_bin_list', 'closed_triangle_indices']


def geometric_k_mean(kmin, kmax):
    """Mean |k| over a spherical shell, (3/4)(kmax^4-kmin^4)/(kmax^3-kmin^3).

    bskit/main.py:408-428.
    """
    return 0.75 * (kmax ** 4 - kmin ** 4) / (kmax ** 3 - kmin ** 3)


def generate_bin_edge_list(kmin=-1., kmax=-1., dk=-1., num_lowk_bins=0, dk_high=-1.):
    """(nbin, 2) array of [k_low, k_high] bin edges; bskit/main.py:1008-1071.

    The lower edges are np.arange(kmin, kmax-dk, dk), so kmax = kmin + n*dk
    yields n-1 bins (the reference's behaviour).  With num_lowk_bins > 0 the
    bins above the first num_lowk_bins have width dk_high.
    """
    if kmin <= 0.:
        raise ValueError('kmin must be > 0!')
    if kmax <= 0.:
        raise ValueError('kmax must be > 0!')
    if dk <= 0.:
        raise ValueError('dk must be > 0!')
    if num_lowk_bins > 0 and dk_high < 0.:
        raise ValueError('Must specify dk_high if num_lowk_bins > 0!')
    low = np.arange(kmin, kmax - dk, dk)
    high = low + np.ones_like(low) * dk
    if 0 < num_lowk_bins < len(low):
        low2 = np.arange(high[num_lowk_bins - 1], kmax - dk_high, dk_high)
        high2 = low2 + np.ones_like(low2) * dk_high
        low = np.concatenate([low[:num_lowk_bins], low2])
        high = np.concatenate([high[:num_lowk_bins], high2])
    return np.stack([low, high], axis=1)


def _edges_of(kLU, idx):
    idx = np.asarray(idx, dtype=np.int64).reshape(-1, 3)
    return kLU[idx].reshape(len(idx), 6)


def generate_equilateral_triangle_bin_list(kmin=-1., kmax=-1., dk=-1., num_lowk_bins=0,
                                           dk_high=-1., return_indices=False):
    """Triangles (b, b, b) for every bin; bskit/main.py:1074-1118."""
    kLU = generate_bin_edge_list(kmin, kmax, dk, num_lowk_bins, dk_high)
    idx = np.repeat(np.arange(len(kLU))[:, None], 3, axis=1)
    return idx if return_indices else _edges_of(kLU, idx)


def generate_squeezed_triangle_bin_list(kmin=-1., kmax=-1., dk=-1., squeezed_bin_index=0,
                                        num_lowk_bins=0, dk_high=-1., return_indices=False):
    """Triangles (q, j, j) for all j > q = squeezed_bin_index; bskit/main.py:1121-1178."""
    kLU = generate_bin_edge_list(kmin, kmax, dk, num_lowk_bins, dk_high)
    j = np.arange(squeezed_bin_index + 1, len(kLU))
    idx = np.stack([np.full(len(j), squeezed_bin_index), j, j], axis=1)
    return idx if return_indices else _edges_of(kLU, idx)


def generate_isosceles_triangle_bin_list(kmin=-1., kmax=-1., dk=-1., isos_mult=0, isos_tol=0.1,
                                         num_lowk_bins=0, dk_high=-1., return_indices=False):
    """Triangles (L, S, S) with kbar_L ~ kbar_S / isos_mult; bskit/main.py:1181-1269.

    For each bin S the bin L with geometric k-mean nearest kbar_S/isos_mult
    (first on ties) is kept when |kbar_L - target| / target < isos_tol.
    """
    if isos_mult < 1.:
        raise ValueError('isos_mult must be greater than 1! Come on man...')
    kLU = generate_bin_edge_list(kmin, kmax, dk, num_lowk_bins, dk_high)
    kbar = geometric_k_mean(kLU[:, 0], kLU[:, 1])
    target = kbar / isos_mult
    dist = np.abs(kbar[None, :] - target[:, None])
    near = dist.argmin(axis=1)
    dev = dist.min(axis=1) / target
    keep = dev < isos_tol
    s = np.arange(len(kLU))[keep]
    idx = np.stack([near[keep], s, s], axis=1)
    return idx if return_indices else _edges_of(kLU, idx)


def closed_triangle_indices(kLU, num_fields=1):
    """Index triples of bins that can close a triangle, in the reference order.

    1 field: i >= j >= k; 2 fields (<AAB>): i >= j, any k; 3 fields: any.
    A triple is kept when every side's lower edge is reachable by the sum of
    the other two upper edges (bskit/main.py:1331-1340).
    """
    n = len(kLU)
    lo, hi = kLU[:, 0], kLU[:, 1]
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing='ij')
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    ok = (hi[i] + hi[j] >= lo[k]) & (hi[i] + hi[k] >= lo[j]) & (hi[j] + hi[k] >= lo[i])
    if num_fields == 1:
        ok &= (i >= j) & (j >= k)
    elif num_fields == 2:
        ok &= i >= j
    return np.stack([i[ok], j[ok], k[ok]], axis=1)


def generate_triangle_bin_list(kmin=-1., kmax=-1., dk=-1., mu_min=None, mu_max=None, dmu=None,
                               num_fields=1, num_lowk_bins=0, dk_high=-1., return_indices=False):
    """All closed triangle bins; bskit/main.py:1272-1389.

    Rows are [k1_low, k1_high, k2_low, k2_high, k3_low, k3_high] sorted by
    (k1_low, k2_low, k3_low); with one field k1 >= k2 >= k3, with two fields
    k1 >= k2, with three fields every ordering.
    """
    if num_fields not in [1, 2, 3]:
        raise ValueError('num_fields must be 1, 2, or 3!')
    if mu_min is not None or mu_max is not None or dmu is not None:
        raise NotImplementedError('Mu binning not implemented yet!')
    kLU = generate_bin_edge_list(kmin, kmax, dk, num_lowk_bins, dk_high)
    idx = closed_triangle_indices(kLU, num_fields)
    return idx if return_indices else _edges_of(kLU, idx)
