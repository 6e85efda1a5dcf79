"""Kernel-level parity of the HIP path against the FP64 oracle (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

from oracle import estimator as est  # noqa: E402
from oracle import direct  # noqa: E402


def _engine(N, box, k_low, k_high, shell_mesh=None):
    from bskit_amd.engine import GridEngine
    return GridEngine(N, box, k_low, k_high, shell_mesh=shell_mesh)


def _bins(nb, width, box):
    kf = 2 * np.pi / box
    lo = kf * (1 + width * np.arange(nb))
    return lo, lo + width * kf


def _field(N, seed, dtype):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((N, N, N))
    return (f + 0.3 * (f * f - 1)).astype(dtype)


def _compact(dk_half, eng):
    """Gather numpy's half spectrum onto the engine's compact cube [ry][rx][kz]."""
    N = eng.N
    n = eng.nlo + np.arange(eng.MX)
    idx = np.mod(n, N)
    return dk_half[idx][:, idx][:, :, :eng.MZ].transpose(1, 0, 2)


@pytest.mark.parametrize('N,dtype', [(32, np.float32), (48, np.float32), (64, np.float64),
                                     (96, np.float32)])
def test_forward_matches_numpy(N, dtype):
    box = 500.0
    lo, hi = _bins(4, 3, box)
    eng = _engine(N, box, lo, hi)
    f = _field(N, 1, dtype)
    dk = eng.forward(torch.from_numpy(f).cuda()).cpu().numpy()
    ref = _compact(est.forward(f), eng)
    err = np.abs(dk - ref).max() / np.abs(ref).max()
    print('forward', N, dtype.__name__, 'max rel err', err)
    assert err < (2e-6 if dtype == np.float32 else 1e-13)


@pytest.mark.parametrize('N', [16, 32, 48, 64])
def test_forward_full_cube(N):
    # bins reaching past Nyquist keep the whole axis (MX == N)
    box = 100.0
    kf = 2 * np.pi / box
    eng = _engine(N, box, [kf], [kf * N])
    assert eng.MX == N and eng.MZ == N // 2 + 1
    f = _field(N, 2, np.float64)
    dk = eng.forward(torch.from_numpy(f).cuda()).cpu().numpy()
    ref = _compact(est.forward(f), eng)
    assert np.abs(dk - ref).max() / np.abs(ref).max() < 1e-13


@pytest.mark.parametrize('N,nb,width,S', [(32, 5, 2, 32), (48, 6, 3, 48), (64, 8, 2, 64),
                                          (64, 4, 2, 32), (96, 5, 2, 40), (64, 4, 2, 64),
                                          (128, 12, 2, 80), (256, 4, 16, 200)])
def test_delta_shells(N, nb, width, S):
    """Shell fields on the automatic shell mesh S (and on N when pinned)."""
    from bskit_amd import _lib as L
    box = 700.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi, shell_mesh=S if S == N else None)
    assert eng.S == S
    f64 = _field(N, 3, np.float64)
    dk_ref = est.forward(f64)
    ref = np.stack(est.shells_on_mesh(dk_ref, N, S, eng.n2lo, eng.n2hi, 'delta')).reshape(nb, -1)
    scale = np.abs(ref).max()
    # FP64 pipeline
    d64 = eng.forward(torch.from_numpy(f64).cuda())
    s64 = eng.shells(d64, fmt=L.FMT_F64, dtype=torch.float64).data.cpu().numpy()
    assert np.abs(s64 - ref).max() / scale < 1e-12
    # FP32 pipeline, F32 and split-BF16 storage
    d32 = eng.forward(torch.from_numpy(f64.astype(np.float32)).cuda())
    s32 = eng.shells(d32, fmt=L.FMT_F32).data.cpu().numpy()
    sbf = eng.shells(d32, fmt=L.FMT_BF16X2).as_float().cpu().numpy()
    e32 = np.abs(s32 - ref).max() / scale
    ebf = np.abs(sbf - ref).max() / scale
    print('shells N=%d fp32 %.2e bf16x2 %.2e' % (N, e32, ebf))
    assert e32 < 3e-6
    assert ebf < 3e-5


@pytest.mark.parametrize('N,nb,width,S', [(256, 4, 16, 200), (96, 5, 2, 40), (128, 12, 2, 80)])
def test_radix_variants_match_oracle(N, nb, width, S):
    """Every FFT pass on its second radix plan (the autotuner's other choice,
    e.g. the two radix-10 stages of the 100-point z transform at S = 200)
    gives the oracle's forward spectrum and shells."""
    from bskit_amd import _lib as L
    box = 700.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    assert eng.S == S
    for p in range(6):
        eng.set_variant(p, 1)
    f64 = _field(N, 5, np.float64)
    dk_ref = est.forward(f64)
    ref = np.stack(est.shells_on_mesh(dk_ref, N, S, eng.n2lo, eng.n2hi, 'delta')).reshape(nb, -1)
    d32 = eng.forward(torch.from_numpy(f64.astype(np.float32)).cuda())
    cref = _compact(dk_ref, eng)
    efw = np.abs(d32.cpu().numpy() - cref).max() / np.abs(cref).max()
    s32 = eng.shells(d32, fmt=L.FMT_F32).data.cpu().numpy()
    e32 = np.abs(s32 - ref).max() / np.abs(ref).max()
    print('variant 1 N=%d S=%d forward %.2e shells %.2e' % (N, S, efw, e32))
    assert efw < 5e-6 and e32 < 5e-6


@pytest.mark.parametrize('N,nb,width', [(32, 6, 2), (48, 5, 3), (64, 4, 2), (80, 6, 2)])
def test_unit_and_kmag_shells(N, nb, width):
    from bskit_amd import _lib as L
    box = 300.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    kf = eng.kf
    for kind, name in ((L.KIND_UNIT, 'unit'), (L.KIND_KMAG, 'kmag')):
        ref = np.stack(est.shells_on_mesh(None, N, eng.S, eng.n2lo, eng.n2hi, name,
                                          kf)).reshape(nb, -1)
        got = eng.shells(None, kind=kind, fmt=L.FMT_F64, dtype=torch.float64).data.cpu().numpy()
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12


def _bcheck(got, ref, absum, tol):
    # error relative to the sum of |terms| (cancellation-aware) and plain relative
    err = np.abs(got - ref) / absum
    return err.max() < tol, err.max()


@pytest.mark.parametrize('N,nb,width', [(32, 6, 2), (48, 7, 2), (64, 12, 2), (64, 20, 1),
                                        (32, 33, 0.5), (96, 8, 1), (80, 6, 2)])
def test_triples_all_paths(N, nb, width):
    """Sums per cell on the engine's shell mesh against the oracle's N^3 mesh."""
    from bskit_amd import _lib as L
    from bskit_amd import bins
    box = 1000.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    kLU = np.stack([lo, hi], axis=1)
    tri = bins.closed_triangle_indices(kLU, 1)
    f64 = _field(N, 5, np.float64)
    shells_ref = est.shells(est.forward(f64), N, eng.n2lo, eng.n2hi, 'delta')
    ref = est.triple_sums(shells_ref, shells_ref, shells_ref, tri) / N ** 3
    absum = np.array([np.sum(np.abs(shells_ref[i] * shells_ref[j] * shells_ref[k]))
                      for i, j, k in tri]) / N ** 3
    d32 = eng.forward(torch.from_numpy(f64.astype(np.float32)).cuda())
    s32 = eng.shells(d32, fmt=L.FMT_F32)
    sbf = eng.shells(d32, fmt=L.FMT_BF16X2)
    sc = 1.0 / eng.ncell
    plain = eng.triples([s32], tri, path='plain', scale=sc).cpu().numpy()
    tiled = eng.triples([s32], tri, path='tiled', scale=sc).cpu().numpy()
    mfma = eng.triples([sbf], tri, path='mfma', scale=sc).cpu().numpy()
    ok_p, ep = _bcheck(plain, ref, absum, 2e-6)
    ok_t, et = _bcheck(tiled, ref, absum, 2e-6)
    ok_m, em = _bcheck(mfma, ref, absum, 5e-5)
    print('N=%d S=%d nb=%d T=%d plain %.2e tiled %.2e mfma %.2e'
          % (N, eng.S, nb, len(tri), ep, et, em))
    assert ok_p and ok_t and ok_m


@pytest.mark.parametrize('N,S', [(32, 32), (64, 40)])
def test_triples_fp64_plain(N, S):
    from bskit_amd import _lib as L
    from bskit_amd import bins
    nb, box = 6, 800.0
    lo, hi = _bins(nb, 2, box)
    eng = _engine(N, box, lo, hi)
    assert eng.S == S
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1)
    f64 = _field(N, 6, np.float64)
    sh = est.shells(est.forward(f64), N, eng.n2lo, eng.n2hi, 'delta')
    ref = est.triple_sums(sh, sh, sh, tri) / N ** 3
    d = eng.forward(torch.from_numpy(f64).cuda())
    s = eng.shells(d, fmt=L.FMT_F64, dtype=torch.float64)
    for path in ('plain', 'tiled'):
        got = eng.triples([s], tri, path=path, scale=1.0 / eng.ncell).cpu().numpy()
        assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-11, path


@pytest.mark.parametrize('N,nb,width', [(32, 6, 2), (48, 8, 2), (96, 8, 1)])
def test_cross_modes(N, nb, width):
    from bskit_amd import _lib as L
    from bskit_amd import bins
    box = 1000.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    kLU = np.stack([lo, hi], axis=1)
    fa, fb, fc = (_field(N, s, np.float64) for s in (11, 12, 13))
    sh = [est.shells(est.forward(f), N, eng.n2lo, eng.n2hi, 'delta') for f in (fa, fb, fc)]
    sets_bf = [eng.shells(eng.forward(torch.from_numpy(f.astype(np.float32)).cuda()),
                          fmt=L.FMT_BF16X2) for f in (fa, fb, fc)]
    sets_32 = [eng.shells(eng.forward(torch.from_numpy(f.astype(np.float32)).cuda()),
                          fmt=L.FMT_F32) for f in (fa, fb, fc)]
    for nf in (2, 3):
        tri = bins.closed_triangle_indices(kLU, nf)
        sides = (sh[0], sh[0], sh[1]) if nf == 2 else (sh[0], sh[1], sh[2])
        ref = est.triple_sums(*sides, tri) / N ** 3
        absum = np.array([np.sum(np.abs(sides[0][i] * sides[1][j] * sides[2][k]))
                          for i, j, k in tri]) / N ** 3
        sc = 1.0 / eng.ncell
        got_m = eng.triples(sets_bf[:nf], tri, path='mfma', scale=sc).cpu().numpy()
        got_p = eng.triples(sets_32[:nf], tri, path='plain', scale=sc).cpu().numpy()
        ok_m, em = _bcheck(got_m, ref, absum, 5e-5)
        ok_p, ep = _bcheck(got_p, ref, absum, 2e-6)
        print('cross nf=%d T=%d mfma %.2e plain %.2e' % (nf, len(tri), em, ep))
        assert ok_m and ok_p


@pytest.mark.parametrize('N,nb,width', [(32, 6, 2), (48, 8, 2), (96, 8, 1), (80, 6, 2)])
def test_grid_info_exact_counts(N, nb, width):
    from bskit_amd import bins
    box = 1000.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1)
    ntri, kbar = eng.grid_info(tri)
    cnt, ksum, _ = direct.enumerate_triangles(N, eng.n2lo, eng.n2hi, tri)
    assert np.array_equal(ntri.astype(np.int64), cnt)
    nz = cnt > 0
    kref = ksum[nz] / cnt[nz, None] * eng.kf
    assert np.abs(kbar[nz] / kref - 1).max() < 1e-10
    assert np.isnan(kbar[~nz]).all()


def test_grid_info_slabs_match_resident():
    from bskit_amd import bins
    N, nb, box = 48, 6, 1000.0
    lo, hi = _bins(nb, 2, box)
    eng = _engine(N, box, lo, hi)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1)
    a = eng.grid_info(tri)
    b = eng.grid_info(tri, max_bytes=N * N * 8 * 2 * nb * 5)  # forces 4-plane slabs
    assert np.array_equal(a[0], b[0])
    assert np.allclose(a[1], b[1], rtol=1e-12, equal_nan=True)


def _level_mask(S, level):
    """Cells of sublattice level 0..3 of the S^3 mesh (C order)."""
    i = np.arange(S)
    x, y, z = i[:, None, None], i[None, :, None], i[None, None, :]
    if level == 0:
        m = np.ones((S, S, S), bool)
    elif level == 1:
        m = (x + y + z) % 2 == 0
    elif level == 2:
        m = ((x % 2) == (y % 2)) & ((y % 2) == (z % 2))
    else:
        m = (x % 2 == 0) & (y % 2 == 0) & (z % 2 == 0)
    return np.broadcast_to(m, (S, S, S)).reshape(-1)


# Error bound of a split-BF16 MFMA sum relative to A_t = 2^L sum over the
# triangle's level cells of |n_i n_j n_k| / S^3 (every term's own size):
# each stored shell value hi + lo is within 2^-18 of the FP32 field (three
# sides: 3 x 2^-18), the pair product's split adds 2^-18 and the dropped
# lo*lo product 2^-18, the FP32 product 2^-24; 24 FP32 MFMA roundings per
# chain of 8 K steps at most 2 ulp each (48 x 2^-23) and the FP32 partial
# store 2^-24; 2^-15 leaves room for the FP32 transform of the shells.
MFMA_COUNT_BOUND = 2.0 ** -15


@pytest.mark.parametrize('N,nb,width', [(16, 4, 1), (16, 5, 1), (16, 7, 1), (32, 4, 1),
                                        (32, 2, 2), (32, 3, 2), (32, 6, 2), (64, 4, 1)])
def test_mfma_unit_shell_counts_exact(N, nb, width):
    """N_tri from the split-BF16 MFMA path on unit shells: within
    MFMA_COUNT_BOUND * A_t of the direct integer count for every triangle,
    and rounded equal to it wherever that bound is below 1/2.

    The unit shells peak at x = 0 (n_b(0) = number of modes), so a few cells
    near the peak carry most of the sum and their representation error does
    not average out: the error grows with the count (about 2e-6 x N_tri
    measured), not with sqrt(cells).  Counts beyond the bound come from the
    FP64 grid-info path, which is exact at every size."""
    from bskit_amd import _lib as L
    from bskit_amd import bins
    box = 1000.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1)
    unit = eng.shells(None, kind=L.KIND_UNIT, fmt=L.FMT_BF16X2, dtype=torch.float32)
    s = eng.triples([unit], tri, path='mfma', scale=1.0 / eng.ncell).cpu().numpy()
    cnt, _, _ = direct.enumerate_triangles(N, eng.n2lo, eng.n2hi, tri)
    S = eng.S
    n64 = [x.reshape(-1) for x in est.shells_on_mesh(None, N, S, eng.n2lo, eng.n2hi, 'unit')]
    levels = eng.triangle_levels(tri) if unit.layout is not None else np.zeros(len(tri), int)
    masks = {lv: _level_mask(S, lv) for lv in set(levels.tolist())}
    absum = np.array([2.0 ** lv * np.sum(np.abs(n64[i] * n64[j] * n64[k])[masks[lv]])
                      for (i, j, k), lv in zip(tri, levels)]) / S ** 3
    bound = MFMA_COUNT_BOUND * absum
    err = np.abs(s - cnt)
    sure = bound < 0.5
    print('N=%d nb=%d: max |count - mfma| %.3f (max count %d), max err/A_t %.2e, '
          '%d of %d triangles inside the rounding bound (A_t < %d)'
          % (N, nb, err.max(), cnt.max(), (err / absum).max(), sure.sum(), len(tri),
             0.5 / MFMA_COUNT_BOUND))
    assert np.all(err <= bound)
    assert sure.sum() > 0
    assert np.array_equal(np.rint(s[sure]).astype(np.int64), cnt[sure])


@pytest.mark.parametrize('N,nb', [(32, 33), (64, 17)])
def test_mfma_tail_shells(N, nb):
    """Shells past the last full 16-block (the bench's [k_f, k_f] squeezed
    shell) in the pair tiles."""
    from bskit_amd import _lib as L
    from bskit_amd import bins
    box = 1000.0
    kf = 2 * np.pi / box
    width = 0.5 if nb > 32 else 1.0
    lo = kf * (1 + width * np.arange(nb - 1))
    hi = lo + width * kf
    lo = np.concatenate([lo, [kf]])
    hi = np.concatenate([hi, [kf]])
    eng = _engine(N, box, lo, hi)
    base = bins.closed_triangle_indices(np.stack([lo[:-1], hi[:-1]], 1), 1)
    extra = np.array([[nb - 1, j, j] for j in range(1, nb - 1)] + [[nb - 1, 5, 2]])
    tri = np.concatenate([base, extra])
    plan = eng.pair_plan(tri, nb, 0)
    assert 0 < plan.npair <= len(tri)
    f64 = _field(N, 21, np.float64)
    sh = est.shells(est.forward(f64), N, eng.n2lo, eng.n2hi, 'delta')
    ref = est.triple_sums(sh, sh, sh, tri) / N ** 3
    absum = np.array([np.sum(np.abs(sh[i] * sh[j] * sh[k])) for i, j, k in tri]) / N ** 3
    sbf = eng.shells(eng.forward(torch.from_numpy(f64.astype(np.float32)).cuda()),
                     fmt=L.FMT_BF16X2)
    got = eng.triples([sbf], tri, path='mfma', scale=1.0 / eng.ncell).cpu().numpy()
    ok, err = _bcheck(got, ref, absum, 5e-5)
    print('tail shells N=%d nb=%d pairs %d tiles %d items %d err %.2e'
          % (N, nb, plan.npair, plan.ntile, plan.nitem, err))
    assert ok


@pytest.mark.parametrize('mode,nb,ntri', [(0, 40, 2000), (1, 24, 800), (2, 20, 900),
                                          (0, 70, 3000), (2, 70, 0), (1, 100, 0)])
def test_mfma_random_lists_match_fp64(mode, nb, ntri):
    """Random triangle lists (any side order, repeats, up to 70 shells) in every
    role mode: split-BF16 MFMA against the FP64 plain path on the same fields.
    ntri = 0: a dense list whose pair tiles need more kb blocks than a wave
    may hold at these shell counts, so tiles are split over waves."""
    from bskit_amd import _lib as L
    N, box = 64, 1000.0
    kf = 2 * np.pi / box
    step, width = (0.4, 0.4) if nb <= 70 else (0.25, 0.5)  # overlapping: no empty shell
    lo = kf * (1 + step * np.arange(nb))
    hi = lo + width * kf
    eng = _engine(N, box, lo, hi)
    rng = np.random.default_rng(7 + mode)
    if ntri:
        tri = rng.integers(0, nb, size=(ntri, 3)).astype(np.int32)
    else:
        tri = np.array([(i, j, k) for i in range(5) for j in range(4) for k in range(0, nb, 3)],
                       dtype=np.int32)
    nf = mode + 1
    fs = [_field(N, 30 + i, np.float64) for i in range(nf)]
    bf = [eng.shells(eng.forward(torch.from_numpy(f.astype(np.float32)).cuda()),
                     fmt=L.FMT_BF16X2) for f in fs]
    s64 = [eng.shells(eng.forward(torch.from_numpy(f).cuda()), fmt=L.FMT_F64,
                      dtype=torch.float64) for f in fs]
    sc = 1.0 / eng.ncell
    got = eng.triples(bf, tri, path='mfma', scale=sc).cpu().numpy()
    ref = eng.triples(s64, tri, path='plain', scale=sc).cpu().numpy()
    # |terms| scale from the FP64 shells
    roles = {1: (0, 0, 0), 2: (0, 0, 1), 3: (0, 1, 2)}[nf]
    a, b, c = (s64[r].data for r in roles)
    absum = torch.stack([(a[i].abs() * b[j].abs() * c[k].abs()).sum() for i, j, k in tri])
    absum = absum.cpu().numpy() * sc
    assert absum.min() > 0
    err = np.abs(got - ref) / absum
    plan = eng.pair_plan(tri, nb, mode)
    print('mode %d nb %d T %d pairs %d tiles %d items %d ni %d passes %d err %.2e'
          % (mode, nb, ntri, plan.npair, plan.ntile, plan.nitem, plan.items_per_wave,
             plan.npass, err.max()))
    assert err.max() < 5e-5


@pytest.mark.parametrize('N,nb,width,S', [(64, 8, 2, 64), (128, 12, 2, 80), (256, 8, 8, 200),
                                          (256, 32, 2, 200), (64, 1, 22, 64)])
def test_zfused_matches_separate_passes_and_fp64(N, nb, width, S):
    """The fused z c2r + split-BF16 sums against the separate z pass + MFMA
    kernel (same shells up to FP32 summation order) and the FP64 plain path;
    also with one shell (a runtime pencil count of one in its transforms)."""
    from bskit_amd import _lib as L
    from bskit_amd import bins
    box = 1000.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    assert eng.S == S
    kLU = np.stack([lo, hi], axis=1)
    tri = bins.closed_triangle_indices(kLU, 1)
    assert eng.zfused_supported(tri)
    f = _field(N, 11, np.float64)
    d32 = eng.forward(torch.from_numpy(f.astype(np.float32)).cuda())
    sc = 1.0 / eng.ncell
    got = eng.bispectrum_zfused(d32, tri, scale=sc).cpu().numpy()
    sbf = eng.shells(d32, fmt=L.FMT_BF16X2)
    sep = eng.triples([sbf], tri, path='mfma', scale=sc).cpu().numpy()
    s64 = eng.shells(eng.forward(torch.from_numpy(f).cuda()), fmt=L.FMT_F64, dtype=torch.float64)
    ref = eng.triples([s64], tri, path='plain', scale=sc).cpu().numpy()
    a = s64.data
    absum = torch.stack([(a[i].abs() * a[j].abs() * a[k].abs()).sum() for i, j, k in tri])
    absum = absum.cpu().numpy() * sc
    e_sep = (np.abs(got - sep) / absum).max()
    e_ref = (np.abs(got - ref) / absum).max()
    print('zfused N=%d S=%d nb=%d T=%d vs separate %.2e vs fp64 %.2e'
          % (N, S, nb, len(tri), e_sep, e_ref))
    assert e_sep < 2e-6
    assert e_ref < 5e-5


def test_autotune_picks_valid_config_and_keeps_results(tmp_path, monkeypatch):
    """The tuner's choice is cached, re-applied, and changes only rounding."""
    from bskit_amd import _lib as L
    from bskit_amd import bins, tune
    monkeypatch.setenv('BSK_TUNE_CACHE', str(tmp_path / 'tune.json'))
    monkeypatch.setattr(tune, '_CACHE', None)
    N, nb, box = 96, 8, 1000.0
    lo, hi = _bins(nb, 1, box)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1)
    f = torch.from_numpy(_field(N, 41, np.float32)).cuda()
    eng = _engine(N, box, lo, hi)
    ref = eng.triples([eng.shells(eng.forward(f), fmt=L.FMT_BF16X2)], tri, path='mfma',
                      scale=1.0 / eng.ncell).cpu().numpy()
    cfg = eng.autotune(tri=tri, field=f)
    assert len(cfg['variants']) == 6 and set(cfg['variants']) <= {0, 1}
    assert (tmp_path / 'tune.json').exists()
    got = eng.triples([eng.shells(eng.forward(f), fmt=L.FMT_BF16X2)], tri, path='mfma',
                      scale=1.0 / eng.ncell).cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()
    eng2 = _engine(N, box, lo, hi)
    assert tune.lookup(eng2, len(tri), tri=tri) == cfg
    assert tune.lookup(eng2, len(tri), tri=tri[::-1].copy()) is None  # another list
    assert eng2.autotune(tri=tri) == cfg         # cached: no re-timing
    assert eng2.mfma_nwg() == cfg.get('mfma_nwg', eng2.mfma_nwg())
    # cross <AAB> lists are tuned with their own pair plans
    tri2 = bins.closed_triangle_indices(np.stack([lo, hi], 1), 2)
    cfg2 = tune.tune(eng2, tri=tri2, field=f, nfield=2)
    assert cfg2.get('mfma_nwg') and 'zfused' not in cfg2
    # one tuning per mesh class, applied to the next engine of the class
    monkeypatch.setattr(tune, '_CLASS', {})
    e3, e4 = _engine(N, box, lo, hi), _engine(N, box, lo, hi)
    c3 = tune.tune_class(e3, tri=tri, field=f)
    c4 = tune.tune_class(e4, tri=tri2, nfield=2)
    assert c3 is c4 and e4.tuned_nwg == c3.get('mfma_nwg')
