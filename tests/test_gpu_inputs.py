"""Inputs of the estimator on the GPU: CIC particle painting (deterministic),
the CIC window compensation against its closed form, per-mesh compensation
in cross spectra, k_field at any power, and stream ordering of the
throughput and serving paths."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

import bskit_amd  # noqa: E402
from bskit_amd import _lib as L  # noqa: E402
from bskit_amd import synth  # noqa: E402
from bskit_amd.engine import GridEngine, cic_paint  # noqa: E402
from oracle import estimator as est  # noqa: E402


def _positions(npart, box, seed, margin=0.0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-margin, box + margin, size=(npart, 3)).astype(np.float32)
    pos[:4] = 0.0                       # nodes exactly
    pos[4:8] = np.float32(box) * 0.999999  # just below the box edge
    return pos


@pytest.mark.parametrize('N,npart,contrast', [(32, 5000, False), (48, 20000, True),
                                              (64, 40000, False), (24, 3000, True),
                                              (40, 3000, False)])
def test_cic_matches_oracle(N, npart, contrast):
    """Tiled (N % 16 == 0) and atomic (other N) paths against the FP64 paint,
    positions outside the box wrapped."""
    box = 250.0
    pos = _positions(npart, box, seed=N, margin=7.0)
    got = cic_paint(torch.from_numpy(pos).cuda(), N, box, to_contrast=contrast)
    got = got.cpu().numpy().astype(np.float64)
    ref = est.cic_paint(pos, N, box)
    if contrast:
        ref = ref * (N ** 3 / npart) - 1.0
        scale = N ** 3 / npart
    else:
        scale = 1.0
    # FP32 mesh coordinates: fractions carry ~N * 2^-24 absolute error
    err = np.abs(got - ref).max() / scale
    print('N=%d npart=%d max err %.2e (particles)' % (N, npart, err))
    assert err < 8 * N * 2.0 ** -24
    if not contrast:
        assert abs(got.sum() - npart) < 1e-3 * npart ** 0.5


@pytest.mark.parametrize('N', [64, 48])
def test_cic_is_deterministic(N):
    """The same particles in another order, painted twice and concurrently on
    two streams, give the same bits (integer node sums)."""
    box = 100.0
    pos = synth.zeldovich_particles(N // 2, box, seed=5, sigma_disp=3.0)
    perm = np.random.default_rng(1).permutation(len(pos))
    a = torch.from_numpy(pos).cuda()
    b = torch.from_numpy(pos[perm].copy()).cuda()
    ref = cic_paint(a, N, box).clone()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = []
    for rep in range(3):
        with torch.cuda.stream(s1):
            o1 = cic_paint(a, N, box, stream=s1)
        with torch.cuda.stream(s2):
            o2 = cic_paint(b, N, box, stream=s2)
        outs += [(o1, s1), (o2, s2)]
    torch.cuda.synchronize()
    for o, _ in outs:
        assert torch.equal(o, ref)


def test_cic_bench_size_is_deterministic():
    """256^3 Zel'dovich particles on 512^3 nodes (the particle-input
    workload): two paints in different particle orders are equal bit for bit
    and conserve the particle count."""
    N, box = 512, 1000.0
    pos = torch.from_numpy(synth.zeldovich_particles(256, box, seed=11, sigma_disp=5.0)).cuda()
    perm = torch.randperm(pos.shape[0], device='cuda', generator=torch.Generator('cuda').manual_seed(3))
    a = cic_paint(pos, N, box, to_contrast=False)
    b = cic_paint(pos[perm].contiguous(), N, box, to_contrast=False)
    assert torch.equal(a, b)
    total = a.double().sum().item()
    assert abs(total - pos.shape[0]) < 1.0


@pytest.mark.parametrize('N,window', [(32, 32), (48, 96), (64, 32)])
def test_compensation_matches_closed_form(N, window):
    """delta_comp(k) / delta(k) on every retained mode equals the product over
    the axes of 1/(1 - 2/3 sin^2(w N / (2 N_cic)))^(1/2), w the circular
    frequency (CompensateCICShotnoiseNgrid, measure_bs_fast.py:45-57)."""
    box = 500.0
    kf = 2 * np.pi / box
    eng = GridEngine(N, box, [kf], [kf * (N // 2 - 1)])
    rng = np.random.default_rng(N)
    for dt, tol in ((np.float32, 2e-6), (np.float64, 1e-13)):
        f = torch.from_numpy(rng.standard_normal((N, N, N)).astype(dt)).cuda()
        plain = eng.forward(f, window=0).cpu().numpy()
        comp = eng.forward(f, window=window).cpu().numpy()
        n = eng.nlo + np.arange(eng.MX)
        cx = est.cic_compensation(n, N, window)
        cz = est.cic_compensation(np.arange(eng.MZ), N, window)
        factor = cx[:, None, None] * cx[None, :, None] * cz[None, None, :]  # [ry][rx][kz]
        ref = plain.astype(np.complex128) * factor
        err = np.abs(comp - ref).max() / np.abs(ref).max()
        print('N=%d window=%d %s max rel err %.2e' % (N, window, dt.__name__, err))
        assert err < tol
        assert eng.cic_window == window


def test_cross_spectrum_compensates_each_mesh_alone():
    """<AAB> of a compensated particle mesh and an uncompensated array: only
    the particle mesh is divided by the CIC window (each mesh is compensated
    on its own, measure_bs_fast.py:209-223)."""
    N, box = 48, 600.0
    kf = 2 * np.pi / box
    pos = synth.zeldovich_particles(24, box, seed=2, sigma_disp=8.0)
    arr = synth.quadratic_field(N, box, seed=4, dtype=np.float64)
    pm = bskit_amd.ParticleMesh(pos, N, box, compensated=True)
    fb = bskit_amd.FFTBispectrum(pm, Nmesh=N, BoxSize=box, kmin=kf, kmax=kf + 5 * 2 * kf + kf,
                                 dk=2 * kf, second=bskit_amd.ArrayMesh(arr, box),
                                 precision='fp64')
    r = fb.measure_bispectrum_faster()
    # oracle: the same painted particle mesh (GPU paint, pinned above),
    # compensated in FP64 by the closed form; the array as is
    painted = cic_paint(torch.from_numpy(pos).cuda(), N, box).cpu().numpy().astype(np.float64)
    dka = est.forward(painted)
    n = np.fft.fftfreq(N, 1.0 / N)
    c = est.cic_compensation(n, N, N)
    cz = est.cic_compensation(np.arange(N // 2 + 1), N, N)
    dka = dka * c[:, None, None] * c[None, :, None] * cz[None, None, :]
    dkb = est.forward(arr)
    kLU = fb.bin_edges
    lo, hi = est.shell_limits(kLU[:, 0], kLU[:, 1], kf)
    sa = est.shells(dka, N, lo, hi)
    sb = est.shells(dkb, N, lo, hi)
    ref = est.triple_sums(sa, sa, sb, fb.k_indices) * box ** 6 / N ** 3
    absum = np.array([np.sum(np.abs(sa[i] * sa[j] * sb[k])) for i, j, k in fb.k_indices])
    err = np.abs(r['B'] - ref).max() / (absum.max() * box ** 6 / N ** 3)
    print('cross <AAB> particle x array: max err %.2e of sum|terms|' % err)
    assert err < 1e-12
    # and it differs from compensating both (the old behaviour) by far more
    sb2 = est.shells(dkb * c[:, None, None] * c[None, :, None] * cz[None, None, :], N, lo, hi)
    wrong = est.triple_sums(sa, sa, sb2, fb.k_indices) * box ** 6 / N ** 3
    assert np.abs(wrong - ref).max() > 1e3 * np.abs(r['B'] - ref).max()


@pytest.mark.parametrize('p', [0.5, 2.0, -1.0, 1.5])
def test_k_field_any_power(p):
    """k_field(p) is the unnormalised c2r of |k|^p on the closed shell
    (bskit/main.py:227-277) for any power (p = 0.5 is what pk_FFT uses)."""
    N, box = 32, 400.0
    kf = 2 * np.pi / box
    kmin, kmax = 2.5 * kf, 6.0 * kf
    got = bskit_amd.k_field(np.full(3, box), np.full(3, N), kmin, kmax, p).cpu().numpy()
    lo, hi = est.shell_limits([kmin], [kmax], kf)
    ref = est.shell_field(None, est.n2_half(N), lo[0], hi[0], N, 'kpow', kf, power=p)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print('k_field p=%g max rel err %.2e' % (p, err))
    assert err < 1e-12
    # p = 0 is the number field, p = 1 the k field
    n0 = bskit_amd.k_field(np.full(3, box), np.full(3, N), kmin, kmax, 0).cpu().numpy()
    assert np.abs(n0 - est.shell_field(None, est.n2_half(N), lo[0], hi[0], N, 'unit')).max() < 1e-9


def test_pipeline_orders_after_the_callers_stream():
    """A field written on the caller's stream right before submit() (no
    synchronize in between) is what the grid reads."""
    from bskit_amd.pipeline import GridPipeline
    from bskit_amd import bins
    N, box = 64, 1000.0
    kf = 2 * np.pi / box
    kLU = bins.generate_bin_edge_list(kf, kf + 6 * 2 * kf + kf, 2 * kf)
    eng = GridEngine(N, box, kLU[:, 0], kLU[:, 1])
    tri = bins.closed_triangle_indices(kLU, 1)
    src = torch.from_numpy(synth.quadratic_field(N, box, seed=9)).cuda()
    ref_pipe = GridPipeline(eng, tri, inflight=2, scale=1.0 / eng.ncell)
    ref = ref_pipe.submit(src.clone())
    ref_pipe.synchronize()
    ref = ref.clone()
    pipe = GridPipeline(eng, tri, inflight=2, scale=1.0 / eng.ncell)
    outs = []
    for rep in range(4):
        # a slow producer on the default stream: many in-place passes that
        # end with the field equal to src
        f = torch.zeros_like(src)
        for _ in range(50):
            f.mul_(0.5)
        f.add_(src)
        outs.append(pipe.submit(f))
        del f  # freed right away: the pipeline stream must keep it alive
    pipe.synchronize()
    for o in outs:
        assert torch.equal(o, ref)


def test_server_orders_after_the_callers_stream():
    from bskit_amd import serve
    N, box = 64, 1000.0
    kf = 2 * np.pi / box
    kLU = serve.serving_bins(N, 10, kf)
    src = torch.from_numpy(synth.quadratic_field(N, box, seed=10)).cuda()
    srv = serve.Server(nstreams=2)
    alone = serve.Job(jid=0, nmesh=N, box=box, kLU=kLU, nfield=1, fields=[src.clone()])
    ref = srv.run_one(alone)
    jobs = []
    for j in range(4):
        f = torch.zeros_like(src)
        for _ in range(50):
            f.mul_(0.5)
        f.add_(src)
        jobs.append(serve.Job(jid=j + 1, nmesh=N, box=box, kLU=kLU, nfield=1, fields=[f]))
        del f
    srv.run(jobs)
    for j in jobs:
        assert torch.equal(j.result, ref)
