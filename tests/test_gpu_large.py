"""The reference's production mesh size (1024^3) and shell sets streamed in
x slabs (for shell sets larger than HBM).

The 1024^3 input is a band-limited field: the spectrum of a seeded random
32^3 field cut to |n| <= 12, synthesised on the 1024^3 mesh by torch's
inverse FFT (a comparator outside the product path).  Its exact 1/N^3
forward transform is those coefficients, so the forward pass is checked on
every retained mode without a host 1024^3 FFT, and the oracle's shells and
sums run on a 64^3 embedding of the same coefficients (exact by the shell
mesh identity: 64 > 3 x 11).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

import bskit_amd  # noqa: E402
from bskit_amd import bins as B  # noqa: E402
from bskit_amd import synth  # noqa: E402
from bskit_amd.engine import GridEngine  # noqa: E402
from oracle import estimator as est  # noqa: E402

BOX = 1000.0
KF = 2 * np.pi / BOX
NCUT = 12


def _embed(coef_small, Ns, M):
    """Half spectrum of an M^3 mesh holding the |n| <= NCUT coefficients of
    an Ns^3 half spectrum (signed frequencies kept)."""
    f = np.fft.fftfreq(Ns, 1.0 / Ns).astype(int)
    out = np.zeros((M, M, M // 2 + 1), dtype=np.complex128)
    idx = np.mod(f, M)
    out[np.ix_(idx, idx, np.arange(Ns // 2 + 1))] = coef_small
    return out


@pytest.fixture(scope='module')
def band_limited():
    Ns = 32
    g = synth.quadratic_field(Ns, BOX, seed=17, dtype=np.float64)
    c = est.forward(g)
    n2 = est.n2_half(Ns)
    c[n2 > NCUT * NCUT] = 0.0
    N = 1024
    big = torch.zeros((N, N, N // 2 + 1), dtype=torch.complex64, device='cuda')
    f = np.fft.fftfreq(Ns, 1.0 / Ns).astype(int)
    idx = torch.from_numpy(np.mod(f, N)).cuda()
    sub = torch.from_numpy(c.astype(np.complex64)).cuda()
    big[idx[:, None, None], idx[None, :, None], torch.arange(Ns // 2 + 1, device='cuda')] = sub
    field = torch.fft.irfftn(big, s=(N, N, N)) * float(N) ** 3
    del big
    torch.cuda.synchronize()
    return dict(N=N, field=field.contiguous(), coef=c, Ns=Ns)


def test_forward_1024_band_limited(band_limited):
    """The 1024^3 forward transform equals the synthesised coefficients on
    every retained mode (FP32)."""
    N, field, c, Ns = band_limited['N'], band_limited['field'], band_limited['coef'], \
        band_limited['Ns']
    eng = GridEngine(N, BOX, [KF], [(NCUT + 2) * KF])
    dk = eng.forward(field).cpu().numpy().astype(np.complex128)  # [ry][rx][kz]
    ref = np.zeros_like(dk)
    f = np.fft.fftfreq(Ns, 1.0 / Ns).astype(int)
    for ix, nx in enumerate(f):
        for iy, ny in enumerate(f):
            rx, ry = nx - eng.nlo, ny - eng.nlo
            if 0 <= rx < eng.MX and 0 <= ry < eng.MY:
                kz = min(eng.MZ, Ns // 2 + 1)
                ref[ry, rx, :kz] = c[ix, iy, :kz]
    err = np.abs(dk - ref).max() / np.abs(ref).max()
    print('1024^3 forward: max |delta_k - exact| / max = %.2e over %d modes' % (err, ref.size))
    assert err < 2e-6


@pytest.mark.parametrize('prec,tol', [('fp32', 1e-4), ('bf16x3', 1e-3)])
def test_bispectrum_1024_against_oracle(band_limited, prec, tol):
    """FFTBispectrum on the 1024^3 mesh (5 bins of 2k_f from k_f) against the
    FP64 oracle on a 64^3 embedding of the same spectrum."""
    N, field, c, Ns = band_limited['N'], band_limited['field'], band_limited['coef'], \
        band_limited['Ns']
    fb = bskit_amd.FFTBispectrum(field, Nmesh=N, BoxSize=BOX, kmin=KF, kmax=KF + 10 * KF + KF,
                                 dk=2 * KF, precision=prec)
    got = fb.measure_bispectrum_faster()['B']
    kLU, tri = fb.bin_edges, fb.k_indices
    M = 64
    lo, hi = est.shell_limits(kLU[:, 0], kLU[:, 1], KF)
    sh = est.shells(_embed(c, Ns, M), M, lo, hi)
    ref = est.triple_sums(sh, sh, sh, tri) * BOX ** 6 / float(M) ** 3
    err = np.abs(got / ref - 1).max()
    print('1024^3 %s: T=%d max rel B err vs FP64 oracle %.2e' % (prec, len(tri), err))
    assert err < tol


@pytest.mark.parametrize('path,tol', [('mfma', 1e-3), ('plain', 1e-4)])
def test_slab_streaming_on_the_1024_mesh(band_limited, path, tol):
    """Shells on the full 1024^3 mesh (the 1024-point inverse passes) made and
    summed in x slabs of 64 planes, never whole, against the oracle."""
    N, field, c, Ns = band_limited['N'], band_limited['field'], band_limited['coef'], \
        band_limited['Ns']
    kLU = B.generate_bin_edge_list(KF, KF + 10 * KF + KF, 2 * KF)
    tri = B.closed_triangle_indices(kLU, 1)
    eng = GridEngine(N, BOX, kLU[:, 0], kLU[:, 1], shell_mesh=N)
    assert eng.S == 1024
    dk = eng.forward(field)
    got = eng.bispectrum_slabs([dk], tri, path=path, scale=BOX ** 6 / float(eng.ncell),
                               xslab=64).cpu().numpy()
    M = 64
    lo, hi = est.shell_limits(kLU[:, 0], kLU[:, 1], KF)
    sh = est.shells(_embed(c, Ns, M), M, lo, hi)
    ref = est.triple_sums(sh, sh, sh, tri) * BOX ** 6 / float(M) ** 3
    err = np.abs(got / ref - 1).max()
    print('1024^3 shell mesh, %s in 16 slabs: T=%d max rel B err %.2e' % (path, len(tri), err))
    assert err < tol


@pytest.mark.parametrize('nfield', [1, 2])
def test_slab_streaming_equals_resident(nfield):
    """Auto and <AAB> sums made slab by slab equal the resident shell sets
    (MFMA: within the split-BF16 rounding of other pair plans; FP64 plain:
    to FP64 rounding), also through FFTBispectrum when the shell sets exceed
    its HBM threshold."""
    N = 128
    kLU = B.generate_bin_edge_list(KF, KF + 12 * 2 * KF + KF, 2 * KF)
    tri = B.closed_triangle_indices(kLU, nfield)
    eng = GridEngine(N, BOX, kLU[:, 0], kLU[:, 1])
    a, b = synth.correlated_pair_device(N, BOX, 31, 32, 'cuda')
    fields = [a, b][:nfield]
    scale = BOX ** 6 / float(eng.ncell)
    dks = [eng.forward(f) for f in fields]
    sets = [eng.shells(d, fmt=bskit_amd._lib.FMT_BF16X2) for d in dks]
    res = eng.triples(sets, tri, path='mfma', scale=scale).cpu().numpy()
    xs = eng.slab_planes(nfield, bskit_amd._lib.FMT_BF16X2, max_bytes=12 << 20)
    assert xs < eng.S
    got = eng.bispectrum_slabs(dks, tri, scale=scale, max_bytes=12 << 20).cpu().numpy()
    rel = np.abs(got - res).max() / np.abs(res).max()
    dk64 = [eng.forward(f.double()) for f in fields]
    s64 = [eng.shells(d, fmt=bskit_amd._lib.FMT_F64, dtype=torch.float64) for d in dk64]
    r64 = eng.triples(s64, tri, path='plain', scale=scale).cpu().numpy()
    g64 = eng.bispectrum_slabs(dk64, tri, path='plain', scale=scale,
                               max_bytes=24 << 20).cpu().numpy()
    rel64 = np.abs(g64 - r64).max() / np.abs(r64).max()
    print('N=%d nfield=%d slabs of %d planes: max |diff| / max |B|: mfma %.2e, fp64 %.2e'
          % (N, nfield, xs, rel, rel64))
    assert rel < 1e-4 and rel64 < 1e-13
    # the drop-in switches to slabs above its threshold
    kw = dict(kmin=KF, kmax=KF + 12 * 2 * KF + KF, dk=2 * KF, precision='bf16x3')
    if nfield == 2:
        kw['second'] = b
    fb = bskit_amd.FFTBispectrum(a, Nmesh=N, BoxSize=BOX, **kw)
    r_res = fb.measure_bispectrum_faster()['B']
    fb2 = bskit_amd.FFTBispectrum(a, Nmesh=N, BoxSize=BOX, **kw)
    fb2.slab_threshold = 0.0
    r_slab = fb2.measure_bispectrum_faster()['B']
    assert np.abs(r_slab - r_res).max() < 1e-4 * np.abs(r_res).max()
