"""Power spectrum and subbox measurements (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

from oracle import estimator as est  # noqa: E402


def _field(N, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((N, N, N))
    return f + 0.3 * (f * f - 1)


@pytest.mark.parametrize('N,box', [(32, 300.0), (48, 700.0)])
def test_power_spectrum_matches_pk_fft_oracle(N, box):
    import bskit_amd as bsk
    f64 = _field(N, 5)
    kf = 2 * np.pi / box
    edges = np.array([[kf * (1 + 2 * i), kf * (3 + 2 * i)] for i in range(6)])
    p64, n64, k64 = bsk.power_spectrum(bsk.ArrayMesh(f64, BoxSize=box), edges)
    p32, n32, k32 = bsk.power_spectrum(bsk.ArrayMesh(f64.astype(np.float32), BoxSize=box), edges)
    for b, (lo, hi) in enumerate(edges):
        p, nb, km = est.pk_fft(f64, box, lo, hi)
        assert abs(n64[b] - nb) < 1e-6 and n32[b] == n64[b]
        assert np.isclose(k64[b], km, rtol=1e-12)
        assert np.isclose(p64[b], p, rtol=1e-10)
        assert np.isclose(p32[b], p, rtol=2e-5)
    P, Nb, K = bsk.pk_FFT(bsk.ArrayMesh(f64, BoxSize=box), edges[2, 0], edges[2, 1])
    assert np.isclose(P, p64[2], rtol=1e-12) and Nb == n64[2] and np.isclose(K, k64[2])


def test_subbox_measurements_equal_explicit_cuts(tmp_path):
    import bskit_amd as bsk
    from bskit_amd import subbox
    N, box, s = 64, 1000.0, 2
    f = _field(N, 9).astype(np.float32)
    ft = torch.from_numpy(f).cuda()
    sub_box = box / s
    kf = 2 * np.pi / sub_box
    kw = dict(kmin=kf, kmax=kf + 6 * kf + kf / 2, dk=kf)
    res = subbox.measure_subboxes(bsk.ArrayMesh(ft, BoxSize=box), s,
                                  out_file_prefix=str(tmp_path / 'bk'), start=0, end=3, **kw)
    for ind in range(4):
        a, b, c = (int(v) * (N // s) for v in subbox.subbox_index_to_multiindex(ind, s))
        cut = f[a:a + N // s, b:b + N // s, c:c + N // s]
        ref = bsk.FFTBispectrum(bsk.ArrayMesh(np.ascontiguousarray(cut), BoxSize=sub_box), **kw)
        rb = ref.measure_bispectrum_faster()
        assert np.array_equal(res[ind]['B'], rb['B'])
        assert (tmp_path / ('bk_subbox%d.dat' % ind)).exists()
