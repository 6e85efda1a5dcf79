"""The FFTBispectrum drop-in on the GPU: fixtures, output files, paths, cross spectra."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

import bskit_amd  # noqa: E402
from bskit_amd import synth  # noqa: E402
from oracle import direct  # noqa: E402
from oracle import estimator as est  # noqa: E402

FIX_BINS = dict(kmin=0.00314, kmax=0.1, dk=0.00628, num_lowk_bins=3, dk_high=0.01884)


def _fixture(golden_dir, name):
    return np.loadtxt(os.path.join(golden_dir, name))


@pytest.mark.parametrize('N,L,name', [
    (512, 1000., 'Lbox1000_512_kf_3kf_3lowkbins.dat'),
    (256, 500., 'Lbox1000_512_kf_3kf_3lowkbins_subbox0_block1.dat')])
def test_gridinfo_fixture(tmp_path, golden_dir, N, L, name):
    """measure_gridinfo_faster reproduces the reference's grid-info files."""
    ref = _fixture(golden_dir, name)
    dummy = np.zeros((N, N, N), dtype=np.float32)
    fb = bskit_amd.FFTBispectrum(dummy, Nmesh=N, BoxSize=L, for_grid_info_only=True, **FIX_BINS)
    out = tmp_path / 'gi.dat'
    r = fb.measure_gridinfo_faster(imin=0, imax=200, out_file=str(out))
    assert np.array_equal(r['index'], ref[:, 0].astype(int))
    assert np.abs(r['k_edge'] - ref[:, 4:10]).max() < 1e-12
    # N_tri exact (the reference's -9.06e-14 in row 29 of the subbox file is 0)
    assert np.array_equal(r['N_tri'], np.rint(ref[:, 10]))
    ok = r['N_tri'] > 0
    assert np.abs(r['k_mean'][ok] / ref[ok, 1:4] - 1).max() < 5e-7
    assert np.isnan(r['k_mean'][~ok]).all()
    # the written file matches the fixture to print precision
    got = np.loadtxt(out)
    assert got.shape == ref.shape
    assert np.array_equal(got[:, 0], ref[:, 0])
    assert np.allclose(got[ok][:, 1:], ref[ok][:, 1:], rtol=6e-7, atol=0)


def test_fast_equals_slow_and_combiner(tmp_path):
    N, L = 64, 1000.
    field = synth.quadratic_field(N, L, seed=3, dtype=np.float32)
    kf = 2 * np.pi / L
    args = dict(Nmesh=N, BoxSize=L, kmin=kf, kmax=kf + 8 * 2 * kf + kf, dk=2 * kf,
                precision='fp32')
    fb = bskit_amd.FFTBispectrum(field, **args)
    gi = fb.measure_gridinfo_faster(out_file=str(tmp_path / 'gi.dat'))
    bu = fb.measure_bispectrum_faster(out_file=str(tmp_path / 'b.dat'))
    fs = bskit_amd.FFTBispectrum(field, **args)
    full = fs.measure_bispectrum(meas_type='full', out_file=str(tmp_path / 'slow.dat'))
    assert np.array_equal(gi['N_tri'], full['N_tri'])
    assert np.allclose(gi['k_mean'], full['k_mean'], rtol=1e-12)
    ok = full['N_tri'] > 0
    assert np.allclose(bu['B'][ok] / gi['N_tri'][ok], full['B'][ok], rtol=1e-6)
    comb = bskit_amd.io.combine(str(tmp_path / 'gi.dat'), str(tmp_path / 'b.dat'), 1.0,
                                str(tmp_path / 'comb.dat'))
    slow = np.loadtxt(tmp_path / 'slow.dat')
    # combined B (from two %e-rounded files) vs slow B, as the reference's fixtures
    assert np.allclose(comb[ok, 10], slow[ok, 10], rtol=2e-6)
    assert np.array_equal(comb[:, 11], slow[:, 11])


@pytest.mark.parametrize('precision,tol', [('bf16x3', 1e-3), ('fp32', 1e-4)])
def test_b_against_fp64_oracle(precision, tol):
    """B (= B_unnorm / N_tri) against the FP64 oracle at 64^3, 8 bins of 2 k_f."""
    N, L = 64, 1000.
    kf = 2 * np.pi / L
    field = synth.quadratic_field(N, L, seed=1, dtype=np.float32)
    fb = bskit_amd.FFTBispectrum(field, Nmesh=N, BoxSize=L, kmin=kf, kmax=kf + 17 * kf, dk=2 * kf,
                                 precision=precision)
    r = fb.measure_bispectrum(meas_type='full')
    lo, hi = est.shell_limits(fb.bin_edges[:, 0], fb.bin_edges[:, 1], kf)
    tri = fb.k_indices
    ref = est.bispectrum_unnorm([field.astype(np.float64)], L, lo, hi, tri)
    cnt, _, _ = direct.enumerate_triangles(N, lo, hi, tri)
    assert np.array_equal(r['N_tri'].astype(np.int64), cnt)
    rel = np.abs(r['B'] * cnt / ref - 1)
    print(precision, 'T=%d max rel B err %.2e median %.2e' % (len(tri), rel.max(), np.median(rel)))
    assert rel.max() < tol


def test_fp64_input_config1():
    """64^3 FP64 input, 8 bins of 2 k_f: FP64 pipeline against the oracle."""
    N, L = 64, 1000.
    kf = 2 * np.pi / L
    field = synth.quadratic_field(N, L, seed=2, dtype=np.float64)
    fb = bskit_amd.FFTBispectrum(field, Nmesh=N, BoxSize=L, kmin=kf, kmax=kf + 17 * kf, dk=2 * kf)
    assert fb.precision == 'fp64'
    r = fb.measure_bispectrum_faster()
    lo, hi = est.shell_limits(fb.bin_edges[:, 0], fb.bin_edges[:, 1], kf)
    ref = est.bispectrum_unnorm([field], L, lo, hi, fb.k_indices)
    assert np.abs(r['B'] / ref - 1).max() < 1e-10


def test_cross_aab_fast_and_slow():
    N, L = 64, 1000.
    kf = 2 * np.pi / L
    a, b = synth.correlated_quadratic_pair(N, L, 4, 5, rho=0.8)
    args = dict(Nmesh=N, BoxSize=L, kmin=kf, kmax=kf + 6 * 2 * kf + kf, dk=2 * kf)
    fb = bskit_amd.FFTBispectrum(a, second=b, **args)
    assert fb.num_fields == 2
    fast = fb.measure_bispectrum_faster()
    fs = bskit_amd.FFTBispectrum(a, second=b, **args)
    slow = fs.measure_bispectrum(meas_type='unnorm_b_value')
    assert np.allclose(fast['B'], slow['B'], rtol=1e-12)
    lo, hi = est.shell_limits(fb.bin_edges[:, 0], fb.bin_edges[:, 1], kf)
    ref = est.bispectrum_unnorm([a.astype(np.float64), b.astype(np.float64)], L, lo, hi,
                                fb.k_indices)
    cnt, _, _ = direct.enumerate_triangles(N, lo, hi, fb.k_indices)
    err = np.abs(fast['B'] - ref) / np.abs(ref)
    print('cross <AAB> max rel err %.2e' % err.max())
    assert err.max() < 1e-3


@pytest.mark.parametrize('precision,rtol', [('fp32', 1e-7), ('bf16x3', 1e-5)])
def test_equilateral_isosceles_squeezed_consistent_with_all(precision, rtol):
    """A triangle's B does not depend on the list it is measured in, up to
    summation order (the FP32 sweeps' partial sums are laid out per list) and,
    on the MFMA path, split-BF16 rounding (the pair sides, the two shells whose
    product is split, are chosen per list)."""
    N, L = 64, 1000.
    kf = 2 * np.pi / L
    field = synth.quadratic_field(N, L, seed=8, dtype=np.float32)
    base = dict(Nmesh=N, BoxSize=L, kmin=kf, kmax=kf + 10 * 2 * kf + kf, dk=2 * kf,
                precision=precision)
    fa = bskit_amd.FFTBispectrum(field, **base)
    ra = fa.measure_bispectrum_faster()
    table = {tuple(t): b for t, b in zip(fa.k_indices.tolist(), ra['B'])}
    worst = 0.0
    for tt, extra in (('equilateral', {}), ('isosceles', dict(isos_mult=2)),
                      ('squeezed', dict(squeezed_bin_index=0))):
        f = bskit_amd.FFTBispectrum(field, triangle_type=tt, **extra, **base)
        r = f.measure_bispectrum_faster()
        for t, b in zip(f.k_indices.tolist(), r['B']):
            key = tuple(sorted(t, reverse=True))
            worst = max(worst, abs(b / table[key] - 1))
            assert np.isclose(b, table[key], rtol=rtol), (tt, t)
    print('%s: largest relative difference between lists %.2e' % (precision, worst))


def test_k_edges_input_and_units():
    """k_edges input works in both fast methods; pos_units rescale B and k."""
    N, L = 32, 500.
    kf = 2 * np.pi / L
    field = synth.quadratic_field(N, L, seed=9, dtype=np.float32)
    gen = bskit_amd.generate_triangle_bin_list(kf, kf + 5 * 2 * kf + kf, 2 * kf)
    f1 = bskit_amd.FFTBispectrum(field, Nmesh=N, BoxSize=L, k_edges=gen)
    r1 = f1.measure_bispectrum_faster()
    g1 = f1.measure_gridinfo_faster()
    f2 = bskit_amd.FFTBispectrum(field, Nmesh=N, BoxSize=L, kmin=kf, kmax=kf + 5 * 2 * kf + kf,
                                 dk=2 * kf)
    r2 = f2.measure_bispectrum_faster()
    assert np.allclose(r1['B'], r2['B'], rtol=1e-12)
    assert g1['k_mean'].shape == (len(gen), 3)
    u = 2.0
    f3 = bskit_amd.FFTBispectrum(field, Nmesh=N, BoxSize=L, kmin=kf, kmax=kf + 5 * 2 * kf + kf,
                                 dk=2 * kf, pos_units_mpcoverh=u)
    r3 = f3.measure_bispectrum_faster()
    assert np.allclose(r3['B'], r2['B'] * u ** 6, rtol=1e-12)


def test_bk_binned_matches_fft_on_low_bins():
    """Direct enumeration equals the FFT estimator where periodic images cannot close."""
    N, L = 32, 1000.
    kf = 2 * np.pi / L
    field = synth.quadratic_field(N, L, seed=6, dtype=np.float64)
    bins = [(2.5 * kf, 4.5 * kf), (2.5 * kf, 4.5 * kf), (2.5 * kf, 4.5 * kf)]
    B, nt, k1, k2, k3 = bskit_amd.bk_binned(field_mesh := bskit_amd.ArrayMesh(field, L), *bins)
    Bf, ntf, kf1, kf2, kf3 = bskit_amd.bk_FFT_full(field_mesh, *bins)
    assert nt == ntf
    assert abs(B.real / Bf - 1) < 1e-10 and abs(B.imag) < 1e-8 * abs(B.real)
    assert abs(k1 / kf1 - 1) < 1e-10


@pytest.mark.parametrize('N,bins_n2', [
    (16, [(21.5, 40.5), (27.5, 50.5), (37.5, 68.5)]),   # above N/3, reaching the Nyquist plane
    (16, [(24.5, 48.5)] * 3),
    (24, [(60.5, 140.5), (80.5, 150.5), (100.5, 170.5)])])
def test_bk_binned_open_closure_matches_oracle(N, bins_n2):
    """bk_binned (open closure, the reference's half-grid mode sets with
    weight-2 conjugates, half-open bins) against the oracle's restatement of
    bskit/main.py:877-1001 on bins where open and periodic closure differ and
    modes sit on the Nyquist planes."""
    L = 1000.
    kf = 2 * np.pi / L
    field = synth.quadratic_field(N, L, seed=8, dtype=np.float64)
    bins = [(np.sqrt(a) * kf, np.sqrt(b) * kf) for a, b in bins_n2]
    B, nt, k1, k2, k3 = bskit_amd.bk_binned(bskit_amd.ArrayMesh(field, L), *bins)
    rB, rn, r1, r2, r3 = est.binned_direct(field, L, bins)
    lo, hi = est.shell_limits([b[0] for b in bins], [b[1] for b in bins], kf)
    per = direct.enumerate_triangles(N, lo, hi, [[0, 1, 2]])[0][0]
    print('N=%d open Nbin %d (periodic %d), |dB|/|B| %.2e'
          % (N, nt, per, abs(B - rB) / abs(rB)))
    assert nt == rn and nt < per
    assert abs(B - rB) <= 1e-10 * abs(rB)
    assert np.allclose([k1, k2, k3], [r1, r2, r3], rtol=1e-12)
    # the periodic form counts what the FFT estimator counts
    _, ntp, _, _, _ = bskit_amd.bk_binned(bskit_amd.ArrayMesh(field, L), *bins, periodic=True)
    assert ntp == per
