"""Oracle restatement of the reference's k-bin and triangle-bin generators.

Test infrastructure only (see ``oracle/__init__``).  Each function follows
the numbered steps of the cited reference function, written out with plain
numpy/itertools so the output order and the float arithmetic match.
"""
import itertools

import numpy as np


def geometric_k_mean(kmin, kmax):
    """Volume-weighted mean |k| of a shell: bskit/main.py:408-428."""
    return 0.75 * (kmax ** 4 - kmin ** 4) / (kmax ** 3 - kmin ** 3)


def bin_edges(kmin=-1., kmax=-1., dk=-1., num_lowk_bins=0, dk_high=-1.):
    """(nbin, 2) array of [kL, kU]: bskit/main.py:1008-1071.

    Steps: validate (:1042-1048); kL = arange(kmin, kmax-dk, dk) (:1052);
    kU = kL + dk (:1054); optional high-k segment starting at
    kU[num_lowk_bins-1] with step dk_high (:1059-1066).
    """
    if kmin <= 0.:
        raise ValueError('kmin must be > 0!')
    if kmax <= 0.:
        raise ValueError('kmax must be > 0!')
    if dk <= 0.:
        raise ValueError('dk must be > 0!')
    if num_lowk_bins > 0 and dk_high < 0.:
        raise ValueError('Must specify dk_high if num_lowk_bins > 0!')
    lows = np.arange(kmin, kmax - dk, dk)
    widths = np.ones_like(lows) * dk
    highs = lows + widths
    if 0 < num_lowk_bins < len(lows):
        lows_hi = np.arange(highs[num_lowk_bins - 1], kmax - dk_high, dk_high)
        highs_hi = lows_hi + np.ones_like(lows_hi) * dk_high
        lows = np.hstack((lows[:num_lowk_bins], lows_hi))
        highs = np.hstack((highs[:num_lowk_bins], highs_hi))
    return np.vstack((lows, highs)).T


def equilateral(kmin, kmax, dk, num_lowk_bins=0, dk_high=-1., return_indices=False):
    """bskit/main.py:1074-1118: one (b, b, b) triangle per bin."""
    e = bin_edges(kmin, kmax, dk, num_lowk_bins, dk_high)
    if return_indices:
        r = np.arange(len(e))
        return np.vstack((r, r, r)).T
    return np.hstack((e, e, e))


def squeezed(kmin, kmax, dk, squeezed_bin_index=0, num_lowk_bins=0, dk_high=-1.,
             return_indices=False):
    """bskit/main.py:1121-1178: (q, j, j) for every j > q."""
    e = bin_edges(kmin, kmax, dk, num_lowk_bins, dk_high)
    q = squeezed_bin_index
    if return_indices:
        js = np.arange(q + 1, len(e))
        return np.vstack((np.full(len(js), q), js, js)).T
    rest = e[q + 1:]
    first = np.array([e[q] for _ in range(len(rest))])
    return np.hstack((first, rest, rest))


def isosceles(kmin, kmax, dk, isos_mult=0, isos_tol=0.1, num_lowk_bins=0, dk_high=-1.,
              return_indices=False):
    """bskit/main.py:1181-1269.

    For every bin S find the bin L whose geometric k-mean is nearest to
    kmean_S / isos_mult (first minimum wins, as np.argmin); keep S when the
    fractional deviation is below isos_tol; emit (L, S, S).
    """
    if isos_mult < 1.:
        raise ValueError('isos_mult must be greater than 1! Come on man...')
    e = bin_edges(kmin, kmax, dk, num_lowk_bins, dk_high)
    kmean = np.array([geometric_k_mean(a, b) for a, b in e])
    targets = kmean / isos_mult
    nearest = []
    for t in targets:
        d = np.abs(kmean - t)
        nearest.append((d.argmin(), d.min() / t))
    nearest = np.array(nearest)
    keep = [row[1] < isos_tol for row in nearest]
    long_idx = nearest[keep][:, 0].astype(int)
    if return_indices:
        s_idx = np.arange(len(e))[keep]
        l_idx = np.arange(len(e))[long_idx]
        return np.vstack((l_idx, s_idx, s_idx)).T
    return np.hstack((e[long_idx], e[keep], e[keep]))


def _closes(t):
    """Triangle filter of bskit/main.py:1331-1340 on ((lo,hi),(lo,hi),(lo,hi))."""
    return (t[0][1] + t[1][1] >= t[2][0] and t[0][1] + t[2][1] >= t[1][0]
            and t[1][1] + t[2][1] >= t[0][0])


def all_triangles(kmin, kmax, dk, num_fields=1, num_lowk_bins=0, dk_high=-1.,
                  return_indices=False, mu_min=None, mu_max=None, dmu=None):
    """bskit/main.py:1272-1389.

    1 field: combinations_with_replacement of bins, filter, swap sides 0<->2 so
    k1>=k2>=k3.  2 fields: pairs with replacement for (k1,k2) times every k3,
    filter, swap sides 0<->1.  3 fields: full product, filter.  Then sort rows
    by (k1_low, k2_low, k3_low) and map lower edges back to bin indices by
    exact equality.
    """
    if num_fields not in (1, 2, 3):
        raise ValueError('num_fields must be 1, 2, or 3!')
    if mu_min is not None or mu_max is not None or dmu is not None:
        raise NotImplementedError('Mu binning not implemented yet!')
    e = bin_edges(kmin, kmax, dk, num_lowk_bins, dk_high)
    if num_fields == 1:
        cand = list(itertools.combinations_with_replacement(e, 3))
        tri = np.array([t for t in cand if _closes(t)])
        tri[:, [0, 2]] = tri[:, [2, 0]]
    elif num_fields == 2:
        pairs = list(itertools.combinations_with_replacement(e, 2))
        cand = [[p[0], p[1], c] for p in pairs for c in e]
        tri = np.array([t for t in cand if _closes(t)])
        tri[:, [0, 1]] = tri[:, [1, 0]]
    else:
        cand = list(itertools.product(e, repeat=3))
        tri = np.array([t for t in cand if _closes(t)])
    tri = tri.reshape(len(tri), 6)
    # stable lexicographic order on (k1_low, k2_low, k3_low) then the rest
    order = np.lexsort((tri[:, 5], tri[:, 3], tri[:, 1], tri[:, 4], tri[:, 2], tri[:, 0]))
    tri = tri[order]
    if not return_indices:
        return tri
    return np.array([[np.where(e[:, 0] == row[c])[0][0] for c in (0, 2, 4)] for row in tri])
