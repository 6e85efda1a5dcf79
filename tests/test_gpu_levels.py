"""Decimated shell copies and the two-level MFMA triple sums (GPU).

The cells of the S^3 shell mesh with even x, y and z form an (S/2)^3 mesh
on which each shell field is the same sum of modes.  A triangle whose
shells' largest |n| components add up to less than S/2 has no product
frequency that aliases to zero on that mesh, so its cell sum there is
exactly 1/8 of the full one.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def _engine(N, box, k_low, k_high):
    from bskit_amd.engine import GridEngine
    return GridEngine(N, box, k_low, k_high)


def _bins(nb, width, box):
    kf = 2 * np.pi / box
    lo = kf * (1 + width * np.arange(nb))
    return lo, lo + width * kf


def _field(N, seed, dtype):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((N, N, N))
    return (f + 0.3 * (f * f - 1)).astype(dtype)


@pytest.mark.parametrize('N,nb,width,fmt', [(64, 8, 2, 'bf16x2'), (128, 12, 2, 'f32'),
                                            (64, 6, 3, 'f64'), (256, 4, 16, 'bf16x2')])
def test_decimated_copy_is_the_even_cells(N, nb, width, fmt):
    """The z pass's decimated copy holds exactly the even cells (bit for bit)
    and zeros in its padding, whatever the buffer held before."""
    from bskit_amd import _lib as L
    box = 800.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    S, H = eng.S, eng.S // 2
    code = {'bf16x2': L.FMT_BF16X2, 'f32': L.FMT_F32, 'f64': L.FMT_F64}[fmt]
    dt = np.float64 if fmt == 'f64' else np.float32
    dk = eng.forward(torch.from_numpy(_field(N, 5, dt)).cuda())
    out = eng.alloc_shells(code, nb, decimated=True)
    off, cells = out.dec
    assert off >= S ** 3 and off % 128 == 0 and cells == -(-H ** 3 // 128) * 128
    planes = [out.data] + ([out.lo] if out.lo is not None else [])
    for p in planes:
        p.fill_(7)
    sh = eng.shells(dk, fmt=code, dtype=torch.float64 if fmt == 'f64' else torch.float32,
                    out=out)
    for p in planes:
        full = p[:, :S ** 3].reshape(nb, S, S, S)
        dec = p[:, off:off + H ** 3].reshape(nb, H, H, H)
        assert torch.equal(dec, full[:, ::2, ::2, ::2])
        assert bool((p[:, off + H ** 3:off + cells] == 0).all())
    print('decimated copy N=%d S=%d nb=%d %s: %d + %d cells' % (N, S, nb, fmt, S ** 3, cells))
    assert sh is out


@pytest.mark.parametrize('N,nb,width', [(64, 10, 1), (128, 12, 2), (512, 32, 2)])
def test_triangle_levels_follow_the_mode_bound(N, nb, width):
    from bskit_amd import bins
    box = 1000.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1)
    lv = eng.triangle_levels(tri)
    m = np.array([math.isqrt(int(v)) if v >= 0 else 0 for v in eng.n2hi])
    ref = (2 * m[tri].sum(axis=1) < eng.S).astype(np.int32)
    assert np.array_equal(lv, ref)
    print('levels N=%d S=%d: %d of %d triangles on the decimated mesh'
          % (N, eng.S, lv.sum(), len(tri)))
    if N == 512:
        assert 0 < lv.sum() < len(tri)


def test_decimated_sums_are_exact_in_fp64():
    """FP64 shells: for level-1 triangles 8 x (sum over the decimated copy)
    equals the full-mesh sum to rounding; for triangles past the bound the
    decimated sum picks up aliased terms."""
    from bskit_amd import _lib as L
    from bskit_amd import bins
    N, nb, box = 96, 10, 1000.0
    lo, hi = _bins(nb, 2, box)
    eng = _engine(N, box, lo, hi)
    S, H = eng.S, eng.S // 2
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1)
    lv = eng.triangle_levels(tri)
    assert lv.any() and not lv.all()
    dk = eng.forward(torch.from_numpy(_field(N, 9, np.float64)).cuda())
    sh = eng.shells(dk, fmt=L.FMT_F64, dtype=torch.float64,
                    out=eng.alloc_shells(L.FMT_F64, nb, decimated=True))
    off = sh.dec[0]
    a = sh.data[:, :S ** 3]
    d = sh.data[:, off:off + H ** 3]
    full = torch.stack([(a[i] * a[j] * a[k]).sum() for i, j, k in tri]).cpu().numpy()
    dec = 8 * torch.stack([(d[i] * d[j] * d[k]).sum() for i, j, k in tri]).cpu().numpy()
    absum = torch.stack([(a[i].abs() * a[j].abs() * a[k].abs()).sum()
                         for i, j, k in tri]).cpu().numpy()
    err = np.abs(full - dec) / absum
    print('fp64 decimated: level-1 max %.2e, level-0 max %.2e'
          % (err[lv == 1].max(), err[lv == 0].max()))
    assert err[lv == 1].max() < 1e-12
    assert err[lv == 0].max() > 1e-6  # the bound is not vacuous here


@pytest.mark.parametrize('N,nb,width,mode', [(128, 12, 2, 0), (512, 32, 2, 0), (128, 12, 2, 1),
                                             (96, 10, 2, 2)])
def test_two_level_mfma_matches_one_level_and_fp64(N, nb, width, mode):
    """Split-BF16 sums with the two-level plan against the one-level plan on
    the same shells and against the FP64 plain path."""
    from bskit_amd import _lib as L
    from bskit_amd import bins
    box = 1000.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1)
    if mode:
        rng = np.random.default_rng(3)
        tri = np.concatenate([tri, rng.integers(0, nb, size=(200, 3))]).astype(np.int32)
    sc = 1.0 / eng.ncell
    nf = mode + 1
    fs = [_field(N, 40 + i, np.float64) for i in range(nf)]
    d32 = [eng.forward(torch.from_numpy(f.astype(np.float32)).cuda()) for f in fs]
    two = [eng.shells(d, fmt=L.FMT_BF16X2, out=eng.alloc_shells(L.FMT_BF16X2, nb, True))
           for d in d32]
    one = [eng.shells(d, fmt=L.FMT_BF16X2, out=eng.alloc_shells(L.FMT_BF16X2, nb, False))
           for d in d32]
    plan = eng.pair_plan(tri, nb, mode, levels=True)
    assert plan.npass_dec > 0
    got2 = eng.triples(two, tri, path='mfma', scale=sc).cpu().numpy()
    got1 = eng.triples(one, tri, path='mfma', scale=sc).cpu().numpy()
    s64 = [eng.shells(eng.forward(torch.from_numpy(f).cuda()), fmt=L.FMT_F64,
                      dtype=torch.float64) for f in fs]
    ref = eng.triples(s64, tri, path='plain', scale=sc).cpu().numpy()
    a = [x.data for x in s64]
    ra, rb, rc = {0: (0, 0, 0), 1: (0, 0, 1), 2: (0, 1, 2)}[mode]
    absum = torch.stack([(a[ra][i].abs() * a[rb][j].abs() * a[rc][k].abs()).sum()
                         for i, j, k in tri]).cpu().numpy() * sc
    e21 = (np.abs(got2 - got1) / absum).max()
    e2r = (np.abs(got2 - ref) / absum).max()
    print('two-level N=%d nb=%d mode %d: %d full + %d decimated passes, vs one-level %.2e, '
          'vs fp64 %.2e' % (N, nb, mode, plan.npass_full, plan.npass_dec, e21, e2r))
    assert e21 < 5e-5
    assert e2r < 5e-5
