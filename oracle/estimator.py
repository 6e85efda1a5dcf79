"""FP64 numpy restatement of the reference's FFT bispectrum estimator.

Test infrastructure only (see ``oracle/__init__``).

Conventions restated from the reference:

* forward transform of the input mesh, normalised by 1/N^3
  (``bskit/main.py:1608-1621`` -> pmesh ``paint(mode='complex')``);
* per k-bin closed-interval shell mask kL <= |k| <= kU
  (``main.py:1850-1852``; unit shells ``:314-320``; |k| shells ``:264-268``)
  followed by an *unnormalised* c2r (``paint(mode='real')``);
* fast-path unnormalised B = L^6/N^3 * sum_x d_i d_j d_k * u^6
  (``main.py:1875-1882``);
* grid info N_tri = sum_x n_i n_j n_k / N^3 and
  kbar_1 = sum_x kk_i n_j n_k / N^3 / N_tri (and cyclically), divided by u
  (``main.py:2024-2064``);
* slow path B = L^6/N^3 * sum_x d_i d_j d_k / N_tri * u^6 with the field roles
  <AAA>, <AAB>, <ABC> of ``main.py:624-640``.

Shell membership is decided on integer |n|^2 against per-edge thresholds
(``shell_limits``): a mode on an edge within 1e-9 relative belongs to the bin
(closed interval, ties in both neighbouring bins).  The reference's float
predicate is ambiguous exactly on such ties; away from ties both agree.
"""
import numpy as np

TIE_RTOL = 1e-9


def kf_of(box):
    return 2.0 * np.pi / float(box)


def shell_limits(kL, kU, kf):
    """Integer |n|^2 range [lo, hi] of the closed shell kL <= |k| <= kU."""
    kL = np.atleast_1d(np.asarray(kL, dtype=np.float64))
    kU = np.atleast_1d(np.asarray(kU, dtype=np.float64))
    lo = np.empty(len(kL), dtype=np.int64)
    hi = np.empty(len(kU), dtype=np.int64)
    for b in range(len(kL)):
        a = (kL[b] / kf) ** 2 if kL[b] > 0 else 0.0
        c = (kU[b] / kf) ** 2 if kU[b] > 0 else -1.0
        lo[b] = int(np.ceil(a * (1.0 - TIE_RTOL))) if a > 0 else 0
        hi[b] = int(np.floor(c * (1.0 + TIE_RTOL))) if c >= 0 else -1
    return lo, hi


def n2_half(N):
    """Integer |n|^2 on the r2c grid (N, N, N//2+1), components in [-N/2, N/2]."""
    f = np.fft.fftfreq(N, 1.0 / N).astype(np.int64)
    fz = np.arange(N // 2 + 1, dtype=np.int64)
    return f[:, None, None] ** 2 + f[None, :, None] ** 2 + fz[None, None, :] ** 2


def forward(field, workers=None):
    """delta(k) = (1/N^3) sum_x delta(x) e^{-ik.x} on the half grid, FP64
    (workers: host threads of scipy's FFT for large meshes)."""
    field = np.asarray(field, dtype=np.float64)
    if workers:
        import scipy.fft as sfft
        return sfft.rfftn(field, workers=workers) / field.size
    return np.fft.rfftn(field) / field.size


def shell_field(dk_half, n2, lo, hi, N, kind='delta', kf=1.0, power=1.0, workers=None):
    """Unnormalised c2r of the masked half spectrum.

    kind 'delta': mask * delta(k); 'unit': mask; 'kmag': |k| * mask;
    'kpow': |k|**power * mask (k_field, bskit/main.py:227-277, weight
    kk**p at :264-268).
    """
    m = (n2 >= lo) & (n2 <= hi)
    if kind == 'delta':
        spec = np.where(m, dk_half, 0.0)
    elif kind == 'unit':
        spec = m.astype(np.complex128)
    elif kind == 'kmag':
        spec = np.where(m, kf * np.sqrt(n2.astype(np.float64)), 0.0).astype(np.complex128)
    elif kind == 'kpow':
        kk = kf * np.sqrt(n2.astype(np.float64))
        with np.errstate(divide='ignore'):
            spec = np.where(m, kk ** power, 0.0).astype(np.complex128)
    else:
        raise ValueError(kind)
    if workers:
        import scipy.fft as sfft
        return sfft.irfftn(spec, s=(N, N, N), axes=(0, 1, 2), workers=workers) * float(N) ** 3
    return np.fft.irfftn(spec, s=(N, N, N), axes=(0, 1, 2)) * float(N) ** 3


def shells(dk_half, N, lo, hi, kind='delta', kf=1.0, workers=None):
    n2 = n2_half(N)
    return [shell_field(dk_half, n2, lo[b], hi[b], N, kind, kf, workers=workers)
            for b in range(len(lo))]


def shells_on_mesh(dk_half, N, S, lo, hi, kind='delta', kf=1.0, workers=None):
    """The same masked spectra as ``shells`` evaluated on an S^3 mesh.

    The modes of the N-mesh half spectrum with signed components
    representable on S (|n_a| < S/2) are placed on the S-mesh half spectrum
    and transformed there; modes outside that range must not be in any shell
    (true whenever S > 2m).  For S > 3m the cell sums of triple products on
    the S mesh times (N/S)^3 equal those on the N mesh (zero-mode identity).
    At S == N this is ``shells`` itself.
    """
    if S == N:
        return shells(dk_half, N, lo, hi, kind, kf, workers=workers)
    f = np.fft.fftfreq(S, 1.0 / S).astype(np.int64)
    n2 = n2_half(S)
    for b in range(len(lo)):
        if hi[b] >= (S // 2) ** 2:
            raise ValueError('shell %d reaches the Nyquist plane of S=%d' % (b, S))
    sub = None
    if dk_half is not None:
        ix = np.mod(f, N)
        sub = dk_half[ix][:, ix][:, :, :S // 2 + 1]
    return [shell_field(sub, n2, lo[b], hi[b], S, kind, kf, workers=workers)
            for b in range(len(lo))]


def triple_sums(fa, fb, fc, tri):
    """sum_x A_i B_j C_k for every (i, j, k) in tri, FP64."""
    out = np.empty(len(tri), dtype=np.float64)
    for t, (i, j, k) in enumerate(tri):
        out[t] = np.sum(fa[i] * fb[j] * fc[k])
    return out


def grid_info(N, box, lo, hi, tri, units=1.0):
    """(N_tri float64[T], kbar float64[T,3]) by the FFT method in FP64."""
    kf = kf_of(box)
    n = shells(None, N, lo, hi, 'unit')
    kk = shells(None, N, lo, hi, 'kmag', kf)
    ncell = float(N) ** 3
    ntri = triple_sums(n, n, n, tri) / ncell
    kbar = np.empty((len(tri), 3))
    with np.errstate(divide='ignore', invalid='ignore'):
        kbar[:, 0] = triple_sums(kk, n, n, tri) / ncell / ntri
        kbar[:, 1] = triple_sums(n, kk, n, tri) / ncell / ntri
        kbar[:, 2] = triple_sums(n, n, kk, tri) / ncell / ntri
    return ntri, kbar / units


def bispectrum_unnorm(fields, box, lo, hi, tri, units=1.0):
    """Unnormalised B for 1, 2 or 3 real input meshes (roles of main.py:624-640).

    fields: list of 1-3 real N^3 arrays.  Returns float64[T].
    """
    N = fields[0].shape[0]
    ncell = float(N) ** 3
    dks = [forward(f) for f in fields]
    sh = [shells(dk, N, lo, hi, 'delta') for dk in dks]
    if len(sh) == 1:
        a, b, c = sh[0], sh[0], sh[0]
    elif len(sh) == 2:
        a, b, c = sh[0], sh[0], sh[1]
    else:
        a, b, c = sh[0], sh[1], sh[2]
    return triple_sums(a, b, c, tri) * float(box) ** 6 / ncell * units ** 6


def full_dk(field):
    """Full complex delta(k) (FFT index order, 1/N^3) for the direct enumerator."""
    field = np.asarray(field, dtype=np.float64)
    return np.fft.fftn(field) / field.size


def pk_fft(field, box, kmin, kmax):
    """(P, Nbin, k_mean) exactly as pk_FFT forms them (bskit/main.py:333-405):
    Nbin = sum_x n^2 / N^3 with n the unnormalised c2r of the unit mask,
    k_mean = sum_x kappa^2 / N^3 / Nbin with kappa the c2r of |k|^(1/2) on
    the mask, P = sum_x d^2 L^3 / N^3 / Nbin with d the c2r of the masked
    forward transform.  Literal real-space sums, FP64."""
    field = np.asarray(field, dtype=np.float64)
    N = field.shape[0]
    kf = kf_of(box)
    lo, hi = shell_limits([kmin], [kmax], kf)
    n2 = n2_half(N)
    m = (n2 >= lo[0]) & (n2 <= hi[0])
    s = float(N) ** 3
    n = np.fft.irfftn(m.astype(np.complex128), s=(N, N, N), axes=(0, 1, 2)) * s
    nbin = np.sum(n * n) / s
    kap = np.fft.irfftn(np.where(m, np.sqrt(kf * np.sqrt(n2.astype(np.float64))), 0.0)
                        .astype(np.complex128), s=(N, N, N), axes=(0, 1, 2)) * s
    kmean = np.sum(kap * kap) / s / nbin
    d = np.fft.irfftn(np.where(m, forward(field), 0.0), s=(N, N, N), axes=(0, 1, 2)) * s
    p = np.sum(d * d) * float(box) ** 3 / s / nbin
    return p, nbin, kmean


def cic_paint(pos, N, box):
    """Cloud-in-cell density (particles per node) in FP64: node i at
    i*box/N, a particle at mesh coordinate u = x N / box adds (1 - f) to
    node floor(u) and f to floor(u)+1 per axis (f = u - floor(u)), indices
    wrapped mod N.  nbodykit's CIC window as used through
    to_mesh(window='cic') by scripts/measure/measure_bs_fast.py:209-217."""
    u = np.asarray(pos, dtype=np.float64) * (N / box)
    fl = np.floor(u)
    f = u - fl
    i0 = np.mod(fl.astype(np.int64), N)
    i1 = (i0 + 1) % N
    grid = np.zeros(N * N * N)
    for c in range(8):
        ix = i1[:, 0] if c & 4 else i0[:, 0]
        iy = i1[:, 1] if c & 2 else i0[:, 1]
        iz = i1[:, 2] if c & 1 else i0[:, 2]
        w = ((f[:, 0] if c & 4 else 1 - f[:, 0]) * (f[:, 1] if c & 2 else 1 - f[:, 1]) *
             (f[:, 2] if c & 1 else 1 - f[:, 2]))
        grid += np.bincount((ix * N + iy) * N + iz, weights=w, minlength=N ** 3)
    return grid.reshape(N, N, N)


def cic_compensation(n, nmesh, nmesh_cic):
    """Per-axis factor CompensateCICShotnoiseNgrid multiplies delta(k) by
    (scripts/measure/measure_bs_fast.py:45-57): with the circular frequency
    w = 2 pi n / nmesh of integer frequency n, 1 / (1 - 2/3 sin^2(w nmesh /
    (2 nmesh_cic)))^(1/2); the factor of a mode is the product over its
    three axes."""
    w = 2.0 * np.pi * np.asarray(n, dtype=np.float64) / nmesh
    return 1.0 / (1.0 - 2.0 / 3.0 * np.sin(0.5 * w * nmesh / nmesh_cic) ** 2) ** 0.5


def binned_direct(field, box, bins):
    """bk_binned (bskit/main.py:834-1004) restated: (B complex, Nbin, k1, k2,
    k3 means).

    Modes of bin b are collected from the half grid of the 1/N^3 forward
    transform (:877-880) with np.digitize on |k|^2 against [kmin^2, kmax^2]
    (half-open, :912-916); a mode of Hermitian weight 2 (0 < k_z < Nyquist;
    the k_z = 0 and Nyquist planes have weight 1) also contributes its
    conjugate -k with conj(delta) (:930-941).  A triple is closed when
    k1 + k2 + k3 = 0 (:966-979: |k_tot|^2 < 1e-4 k_f holds only for the zero
    vector of the lattice), found here by looking -(k1 + k2) up among bin 2's
    modes.  B = L^6 sum delta1 delta2 delta3 / Nbin, k means / Nbin (:993-1001).
    """
    field = np.asarray(field, dtype=np.float64)
    N = field.shape[0]
    kf = kf_of(box)
    dk = np.fft.rfftn(field) / field.size
    f = np.fft.fftfreq(N, 1.0 / N).astype(np.int64)
    fz = np.arange(N // 2 + 1, dtype=np.int64)
    nx, ny, nz = np.meshgrid(f, f, fz, indexing='ij')
    k2 = (kf * nx) ** 2 + (kf * ny) ** 2 + (kf * nz) ** 2
    weight = np.where((nz == 0) | (nz == N // 2), 1, 2)
    sets = []
    for kmin, kmax in bins:
        inb = np.digitize(k2.ravel(), [kmin ** 2, kmax ** 2]).reshape(k2.shape) == 1
        vec = np.stack([nx[inb], ny[inb], nz[inb]], axis=1)
        val = dk[inb]
        two = weight[inb] == 2
        vec = np.concatenate([vec, -vec[two]])
        val = np.concatenate([val, np.conj(val[two])])
        sets.append((vec, val))
    (v0, d0), (v1, d1), (v2, d2) = sets
    span = 4 * N + 1  # keys of integer vectors with components in (-2N, 2N]

    def key(v):
        return ((v[..., 0] + 2 * N) * span + (v[..., 1] + 2 * N)) * span + (v[..., 2] + 2 * N)

    k2key = key(v2)
    order = np.argsort(k2key)
    skey = k2key[order]
    nbin = 0
    bsum = 0j
    ksum = np.zeros(3)
    m0 = np.sqrt((v0 * v0).sum(1)) * kf
    m1 = np.sqrt((v1 * v1).sum(1)) * kf
    m2 = np.sqrt((v2 * v2).sum(1)) * kf
    for i in range(len(v0)):
        need = key(-(v0[i][None, :] + v1))
        pos = np.searchsorted(skey, need)
        pos = np.minimum(pos, len(skey) - 1) if len(skey) else pos
        hit = (skey[pos] == need) if len(skey) else np.zeros(len(need), bool)
        if not hit.any():
            continue
        j = np.nonzero(hit)[0]
        t = order[pos[hit]]
        nbin += len(j)
        bsum += np.sum(d0[i] * d1[j] * d2[t])
        ksum += [m0[i] * len(j), m1[j].sum(), m2[t].sum()]
    B = bsum * float(box) ** 6
    if nbin > 0:
        B /= nbin
        ksum /= nbin
    return B, nbin, ksum[0], ksum[1], ksum[2]
