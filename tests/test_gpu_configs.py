"""The benchmark workloads at their own sizes, against the FP64 oracle.

1. 64^3 FP64 local-quadratic field, 8 bins of 2k_f from k_f, all triangles.
2. 256^3 FP32 local-quadratic field, 16 bins of 3k_f from k_f, all
   triangles: FP32 plain and split-BF16 MFMA each against FP64.
3. 256^3 Zel'dovich particles CIC-painted (compensated) on a 512^3 mesh, 32
   bins of 2k_f: all, equilateral, isosceles m=2 and squeezed with
   |k1| = k_f.
4. Batched cross <AAB> pairs at 128^3, 256^3 and 384^3 in one graph.
5. A serving mix over N in {128 .. 768}: alone equals in the mix.

Every max relative error is printed with its configuration.  The oracle
runs on the host in FP64 (numpy / scipy FFTs, the shell-mesh identity of
the estimator pinned in test_oracle_fixtures.py); triangle counts come from
the direct integer pair counter (no FFT).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

import bskit_amd  # noqa: E402
from bskit_amd import bins as B  # noqa: E402
from bskit_amd import synth  # noqa: E402
from oracle import direct  # noqa: E402
from oracle import estimator as est  # noqa: E402

BOX = 1000.0
KF = 2 * np.pi / BOX
WORKERS = min(16, len(os.sched_getaffinity(0)))


def _rel(got, ref):
    return np.abs(np.asarray(got) - ref) / np.abs(ref)


def _oracle_sums(fields, N, S, kLU, tri, comp=None):
    """L^6/S^3 sum over the S^3 mesh of the FP64 shells of 1 or 2 host
    fields (<AAA> or <AAB>) for the triangles tri; comp: per-axis CIC
    compensation of the first field (closed form)."""
    lo, hi = est.shell_limits(kLU[:, 0], kLU[:, 1], KF)
    sets = []
    for fi, f in enumerate(fields):
        dk = est.forward(f, workers=WORKERS)
        if comp is not None and fi == 0:
            n = np.fft.fftfreq(N, 1.0 / N)
            c = est.cic_compensation(n, N, comp)
            cz = est.cic_compensation(np.arange(N // 2 + 1), N, comp)
            dk *= c[:, None, None] * c[None, :, None] * cz[None, None, :]
        need = sorted(set(np.asarray(tri).reshape(-1).tolist()))
        sh = est.shells_on_mesh(dk, N, S, lo[need], hi[need], workers=WORKERS)
        full = [None] * len(kLU)
        for b, s in zip(need, sh):
            full[b] = s
        sets.append(full)
        del dk
    a, c = sets[0], sets[-1]
    return est.triple_sums(a, a, c, tri) * BOX ** 6 / float(S) ** 3


def _counts(N, kLU, tri):
    lo, hi = est.shell_limits(kLU[:, 0], kLU[:, 1], KF)
    return direct.count_triangles(N, lo, hi, tri)


def _fb(src, N, precision, **kw):
    return bskit_amd.FFTBispectrum(src, Nmesh=N, BoxSize=BOX, precision=precision, **kw)


def test_config1_64cubed_fp64():
    """64^3 FP64 field, 8 bins of 2k_f from k_f (kmax = 18 k_f), all triangles
    (T = 98): the slow path's B and N_tri in FP64, FP32 and split-BF16."""
    N = 64
    field = synth.quadratic_field(N, BOX, seed=1, dtype=np.float64)
    kw = dict(kmin=KF, kmax=18 * KF, dk=2 * KF)
    fb = _fb(field, N, None, **kw)
    assert fb.precision == 'fp64'
    r = fb.measure_bispectrum()
    tri = fb.k_indices
    kLU = fb.bin_edges
    assert len(tri) == 98 and len(kLU) == 8
    cnt = _counts(N, kLU, tri)
    assert np.array_equal(r['N_tri'], cnt.astype(np.float64))
    ref = _oracle_sums([field], N, N, kLU, tri) / cnt
    e64 = _rel(r['B'], ref).max()
    errs = {'fp64': e64}
    f32 = field.astype(np.float32)
    ref32 = _oracle_sums([f32], N, N, kLU, tri) / cnt
    for prec in ('fp32', 'bf16x3'):
        rr = _fb(f32, N, prec, **kw).measure_bispectrum_faster()
        errs[prec] = _rel(rr['B'] / cnt, ref32).max()
    print('config 1 (64^3, 8 bins, T=98): max rel B err vs FP64 oracle: fp64 %.2e, fp32 %.2e, '
          'split-bf16 %.2e; N_tri exact' % (errs['fp64'], errs['fp32'], errs['bf16x3']))
    assert errs['fp64'] < 1e-10
    assert errs['fp32'] < 1e-4
    assert errs['bf16x3'] < 1e-3


def test_config2_256cubed_fp32():
    """256^3 FP32 field, 16 bins of 3k_f from k_f (kmax = 50.5 k_f: the last
    bin ends at 49 k_f), all triangles (T = 564): FP32 plain <= 1e-4 and
    split-BF16 <= 1e-3 against FP64; N_tri equal to the direct count."""
    N = 256
    field = synth.quadratic_field(N, BOX, seed=1, dtype=np.float32)
    kw = dict(kmin=KF, kmax=50.5 * KF, dk=3 * KF)
    res = {}
    for prec in ('fp32', 'bf16x3'):
        fb = _fb(field, N, prec, **kw)
        res[prec] = fb.measure_bispectrum_faster()['B']
    tri, kLU = fb.k_indices, fb.bin_edges
    assert len(tri) == 564 and len(kLU) == 16
    S = bskit_amd.main.get_engine(N, BOX, kLU).S
    ref = _oracle_sums([field], N, S, kLU, tri)
    e32 = _rel(res['fp32'], ref)
    ebf = _rel(res['bf16x3'], ref)
    gi = _fb(field, N, 'fp32', for_grid_info_only=True, **kw).measure_gridinfo_faster()
    cnt = _counts(N, kLU, tri)
    print('config 2 (256^3 FP32, 16 bins of 3k_f, T=564, shell mesh %d): max rel B err vs FP64 '
          'oracle: fp32 plain %.2e (median %.1e), split-bf16 %.2e (median %.1e); N_tri exact '
          '(max %d)' % (S, e32.max(), np.median(e32), ebf.max(), np.median(ebf), cnt.max()))
    assert np.array_equal(gi['N_tri'], cnt.astype(np.float64))
    assert e32.max() < 1e-4
    assert ebf.max() < 1e-3


def _config3_lists(nb):
    kmax = KF + nb * 2 * KF + KF
    return dict(kmin=KF, kmax=kmax, dk=2 * KF)


def test_config3_512cubed_particles_all_bin_sets():
    """256^3 Zel'dovich particles CIC-painted to 512^3 with the window
    compensation, 32 bins of 2k_f from k_f.  Split-BF16 against the FP64 path
    on the same mesh for all, equilateral, isosceles(m=2) and squeezed
    (|k1| = k_f through k_edges); the equilateral and isosceles rows equal
    the same rows of the 'all' list; N_tri of every list equals the direct
    count; 160 triangles of the lists against the host FP64 oracle."""
    N, nb = 512, 32
    pos = synth.zeldovich_particles(256, BOX, seed=3, sigma_disp=5.0)
    pm = bskit_amd.ParticleMesh(pos, N, BOX, compensated=True)
    kw = _config3_lists(nb)
    kLU = B.generate_bin_edge_list(**kw)
    assert len(kLU) == nb
    lists = {'all': dict(triangle_type='all'), 'equilateral': dict(triangle_type='equilateral'),
             'isosceles': dict(triangle_type='isosceles', isos_mult=2)}
    sq_edges = np.array([[KF, KF, a, b, a, b] for a, b in kLU[1:]])
    out = {}
    for name, extra in list(lists.items()) + [('squeezed_kf', None)]:
        for prec in ('bf16x3', 'fp64'):
            if extra is None:
                fb = _fb(pm, N, prec, k_edges=sq_edges)
            else:
                fb = _fb(pm, N, prec, **kw, **extra)
            r = fb.measure_bispectrum_faster()
            out[(name, prec)] = (r['B'], fb)
        e = _rel(out[(name, 'bf16x3')][0], out[(name, 'fp64')][0])
        print('config 3 %-12s T=%4d: split-bf16 vs FP64 path max rel B err %.2e (median %.1e)'
              % (name, len(out[(name, 'fp64')][0]), e.max(), np.median(e)))
        assert e.max() < 1e-3
    # rows of the special lists equal the same rows of 'all'
    fb_all = out[('all', 'fp64')][1]
    row = {tuple(t): i for i, t in enumerate(np.asarray(fb_all.k_indices).tolist())}
    for name in ('equilateral', 'isosceles'):
        fb = out[(name, 'fp64')][1]
        for p in ('fp64', 'bf16x3'):
            sel = [row[tuple(sorted(t, reverse=True))] for t in np.asarray(fb.k_indices).tolist()]
            a = out[(name, p)][0]
            b = out[('all', p)][0][sel]
            if p == 'fp64' and name == 'equilateral':
                # one FP64 sweep per triangle, same side order: the same bits
                assert np.array_equal(a, b)
            elif p == 'fp64':  # isosceles rows list the sides in another order
                assert _rel(a, b).max() < 1e-12
            else:  # other pair plans, other FP32 chains
                assert _rel(a, b).max() < 2e-4
    # triangle counts of every list against the direct count
    for name in ('all', 'equilateral', 'isosceles'):
        fb = out[(name, 'fp64')][1]
        gi = _fb(pm, N, 'fp64', for_grid_info_only=True, **kw, **lists[name])
        gi = gi.measure_gridinfo_faster()
        assert np.array_equal(gi['N_tri'], _counts(N, kLU, fb.k_indices).astype(np.float64))
    sq_fb = out[('squeezed_kf', 'fp64')][1]
    sq_bins, sq_tri = sq_fb._fast_bins()
    gi = bskit_amd.FFTBispectrum(pm, Nmesh=N, BoxSize=BOX, k_edges=sq_edges,
                                 for_grid_info_only=True).measure_gridinfo_faster()
    sq_cnt = _counts(N, sq_bins, sq_tri)
    assert np.array_equal(gi['N_tri'], sq_cnt.astype(np.float64))
    # host FP64 oracle on a subset: every 40th 'all' triangle and the
    # equilateral / squeezed rows
    field = pm.device_field('cuda').cpu().numpy().astype(np.float64)
    S = bskit_amd.main.get_engine(N, BOX, kLU).S
    tri_all = np.asarray(fb_all.k_indices)
    pick = np.arange(0, len(tri_all), 40)
    ref = _oracle_sums([field], N, S, kLU, tri_all[pick], comp=N)
    eq_fb = out[('equilateral', 'fp64')][1]
    ref_eq = _oracle_sums([field], N, S, kLU, np.asarray(eq_fb.k_indices), comp=N)
    S_sq = bskit_amd.main.get_engine(N, BOX, sq_bins).S
    ref_sq = _oracle_sums([field], N, S_sq, sq_bins, sq_tri, comp=N)
    errs = []
    for p in ('fp64', 'bf16x3'):
        e = max(_rel(out[('all', p)][0][pick], ref).max(),
                _rel(out[('equilateral', p)][0], ref_eq).max(),
                _rel(out[('squeezed_kf', p)][0], ref_sq).max())
        errs.append(e)
    print('config 3 vs host FP64 oracle (%d triangles): FP64 path %.2e, split-bf16 %.2e'
          % (len(pick) + len(ref_eq) + len(ref_sq), errs[0], errs[1]))
    assert errs[0] < 1e-9
    assert errs[1] < 1e-3


def test_config4_batched_cross_pairs():
    """One <AAB> pair at each of 128^3, 256^3 and 384^3 (correlation 0.8,
    8 / 16 / 24 bins) in one captured graph: every replay equals the eager
    run, and each pair's sums match the FP64 oracle (a subset of triangles
    at 384^3)."""
    from bskit_amd import batch, serve
    pairs = []
    for n, nb, seed in ((128, 8, 11), (256, 16, 13), (384, 24, 15)):
        a, b = synth.correlated_pair_device(n, BOX, seed, seed + 1, 'cuda')
        pairs.append((a, b, BOX, serve.serving_bins(n, nb, KF)))
    cb = batch.CrossBatch(pairs, nbranch=3)
    eager = [o.clone() for o in cb.run_eager()]
    torch.cuda.synchronize()
    cb.capture()
    for rep in range(3):
        outs = cb.replay()
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert torch.equal(o, e)
        # the program's own eager work between replays
        torch.cuda.synchronize()
    for (a, b, _, kLU), p, e in zip(pairs, cb.pairs, eager):
        n = a.shape[0]
        tri = p.tri
        pick = np.arange(len(tri)) if len(tri) <= 400 else np.linspace(
            0, len(tri) - 1, 400).astype(int)
        ref = _oracle_sums([a.cpu().numpy(), b.cpu().numpy()], n, p.eng.S, kLU, tri[pick])
        err = _rel(e.cpu().numpy()[pick], ref).max()
        print('config 4 pair %d^3 nb=%d T=%d (%d checked): max rel B err vs FP64 oracle %.2e'
              % (n, len(kLU), len(tri), len(pick), err))
        assert err < 1e-3


def test_config5_serving_mix_alone_equals_in_mix():
    """Jobs over N in {128, 192, 256, 384, 512, 768} with 8-40 bins, auto and
    cross: each result is bit-identical alone and in the mix (released
    together, overlapping on 4 streams); two jobs against the FP64 oracle."""
    from bskit_amd import serve
    rng = np.random.default_rng(21)
    jobs = []
    pool = {}
    for i, n in enumerate(serve.MESH_SIZES):
        pool[n] = synth.correlated_pair_device(n, BOX, 700 + 2 * i, 701 + 2 * i, 'cuda')
    jid = 0
    for n in serve.MESH_SIZES:
        for nf in (1, 2):
            nb = int(rng.integers(8, 41))
            jobs.append(serve.Job(jid=jid, nmesh=n, box=BOX, kLU=serve.serving_bins(n, nb, KF),
                                  nfield=nf, fields=list(pool[n])))
            jid += 1
    srv = serve.Server(nstreams=4)
    alone = {}
    for j in jobs:
        alone[j.jid] = srv.run_one(j).numpy().copy()
    serve.reset(jobs)
    for j in jobs:
        j.arrival = 0.0
    srv.run(jobs)
    for j in jobs:
        assert np.array_equal(j.result.numpy(), alone[j.jid]), 'job %d differs in the mix' % j.jid
    for j in (jobs[3], jobs[8]):  # 192^3 cross and 512^3 auto
        n = j.nmesh
        eng = srv.engine(n, BOX, j.kLU)
        host = [f.cpu().numpy() for f in j.fields[:j.nfield]]
        pick = np.arange(len(j.tri)) if len(j.tri) <= 300 else np.linspace(
            0, len(j.tri) - 1, 300).astype(int)
        ref = _oracle_sums(host, n, eng.S, j.kLU, j.tri[pick])
        err = _rel(alone[j.jid][pick], ref).max()
        print('config 5 job %d (%d^3, %s, nb=%d, T=%d): max rel B err vs FP64 oracle %.2e'
              % (j.jid, n, 'auto' if j.nfield == 1 else 'cross', len(j.kLU), len(j.tri), err))
        assert err < 1e-3
    print('config 5: %d jobs over N in %s identical alone and in the mix'
          % (len(jobs), serve.MESH_SIZES))
