"""Serving loop and graph-captured batches (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

from oracle import estimator as est  # noqa: E402


def _small_pool(sizes, box):
    from bskit_amd import synth
    return {n: [synth.correlated_pair_device(n, box, 10 + n + 2 * p, 11 + n + 2 * p)
                for p in range(2)] for n in sizes}


def _oracle_unnorm(job):
    lo, hi = est.shell_limits(job.kLU[:, 0], job.kLU[:, 1], est.kf_of(job.box))
    fields = [f.double().cpu().numpy() for f in job.fields[:job.nfield]]
    ref = est.bispectrum_unnorm(fields, job.box, lo, hi, job.tri)
    # cancellation-aware scale: sum of |terms| per triangle
    N = job.nmesh
    sh = [est.shells(est.forward(f), N, lo, hi) for f in fields]
    a, b, c = (sh[0], sh[0], sh[0]) if job.nfield == 1 else (sh[0], sh[0], sh[1])
    absum = np.array([np.sum(np.abs(a[i] * b[j] * c[k])) for i, j, k in job.tri])
    return ref, absum * job.box ** 6 / float(N) ** 3


def test_serve_results_match_alone_and_oracle():
    from bskit_amd import serve
    box = 1000.0
    sizes = (32, 48, 64)
    pool = _small_pool(sizes, box)
    jobs = serve.make_jobs(24, seed=3, box=box, sizes=sizes, nbin_range=(4, 12), pool=pool)
    srv = serve.Server(nstreams=4)
    alone = []
    for j in jobs:
        srv.prepare(j)
        alone.append(srv.run_one(j))
    serve.reset(jobs)
    for j in jobs:
        j.arrival = 0.0          # all at once: maximal overlap across streams
    stats = srv.run(jobs)
    assert stats['peak_footprint_bytes'] > 0
    assert len({j.stream for j in jobs}) > 1
    for j, ref in zip(jobs, alone):
        assert j.done is not None and j.latency >= 0
        assert torch.equal(j.result, ref), 'job %d differs when served in the mix' % j.jid
    # two jobs (one auto, one cross) against the FP64 oracle on the N^3 mesh
    for nf in (1, 2):
        j = next(x for x in jobs if x.nfield == nf and len(x.tri))
        ref, absum = _oracle_unnorm(j)
        err = np.abs(j.result.numpy() - ref) / absum
        print('job %d N=%d nb=%d nf=%d T=%d err %.2e' % (j.jid, j.nmesh, len(j.kLU), nf,
                                                        len(j.tri), err.max()))
        assert err.max() < 5e-5


def test_serve_admission_respects_budget():
    from bskit_amd import serve
    box = 1000.0
    pool = _small_pool((32, 48), box)
    jobs = serve.make_jobs(12, seed=5, box=box, sizes=(32, 48), nbin_range=(4, 8), pool=pool)
    srv = serve.Server(nstreams=4)
    for j in jobs:
        srv.prepare(j)
    biggest = max(j.footprint for j in jobs)
    srv.budget = biggest                  # forces most jobs to wait for memory
    for j in jobs:
        j.arrival = 0.0
    stats = srv.run(jobs)
    assert stats['peak_footprint_bytes'] <= biggest
    # at every issue instant the jobs in flight fit in the budget
    for b in jobs:
        live = [a for a in jobs if a.issued <= b.issued and a.done > b.issued]
        assert b in live
        assert sum(a.footprint for a in live) <= srv.budget
    # and memory did serialise some of them
    order = sorted(jobs, key=lambda j: j.issued)
    assert any(b.issued >= a.done for a, b in zip(order, order[1:]))


def test_serve_poisson_latency_stats():
    from bskit_amd import serve
    box = 1000.0
    pool = _small_pool((32,), box)
    jobs = serve.make_jobs(20, seed=1, box=box, sizes=(32,), nbin_range=(4, 6), pool=pool)
    serve.poisson_arrivals(jobs, rate=2000.0, seed=1)
    srv = serve.Server(nstreams=2)
    srv.run(jobs)
    st = serve.latency_stats(jobs)
    assert 0 < st['p50_ms'] <= st['p99_ms'] <= st['max_ms']
    assert all(j.issued >= j.arrival - 1e-6 for j in jobs)


def test_cross_batch_graph_replays_match_eager():
    from bskit_amd import batch
    pairs = batch.config_pairs(seed=2, layout=((32, 3), (48, 3), (64, 2)), nbin_range=(4, 10))
    cb = batch.CrossBatch(pairs, nbranch=3)
    eager = [o.clone() for o in cb.run_eager()]
    torch.cuda.synchronize()
    cb.capture()
    for r in range(3):
        # eager work and a synchronisation between replays
        x = torch.randn(1 << 20, device='cuda')
        (x * 2).sum().item()
        outs = cb.replay()
        torch.cuda.synchronize()
        for i, (o, e) in enumerate(zip(outs, eager)):
            assert torch.equal(o, e), 'replay %d pair %d differs from the eager run' % (r, i)
    # one pair against the FP64 oracle (cross <AAB>, N^3 mesh)
    p = cb.pairs[0]
    a, b = (f.double().cpu().numpy() for f in p.fields)
    kLU = pairs[0][3]
    lo, hi = est.shell_limits(kLU[:, 0], kLU[:, 1], est.kf_of(1000.0))
    ref = est.bispectrum_unnorm([a, b], 1000.0, lo, hi, p.tri)
    got = eager[0].cpu().numpy()
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() / scale < 1e-4
