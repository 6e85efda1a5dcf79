#!/usr/bin/env python
"""Benchmark of the FFT bispectrum path on one MI355X.

Workload (the metric's configuration): 256^3 lattice particles displaced by
the Zel'dovich field of a seeded Gaussian (P(k) = 2e4 (k/0.1)^-1.5,
sigma_disp ~ 5 Mpc/h), CIC-painted on the GPU to a 512^3 mesh in
L = 1000 Mpc/h with CIC window compensation; 32 k bins of width 2 k_f from
k_f, measured on the union of the bin sets all triangles (3624),
equilateral (32), isosceles m=2 and squeezed with |k1| = k_f (a 33rd shell
[k_f, k_f]; 31 triangles): T = 3655 triangles over 33 shells.

One step ("one grid") = forward FFT of the HBM-resident painted field +
33 shell fields + all T triangle sums + normalisation by the cached N_tri.
The value is the throughput with --inflight grids (default 3) in flight on
as many HIP streams through bskit_amd.pipeline.GridPipeline (consecutive
grids alternate between the streams, each with its own buffers, so one
grid's FFT passes overlap another's triple sums; the pair kernel's
workgroup count for this mode is picked by timing 256/512/768/1024 in
flight); the same path is then timed one grid at a time with the count
tuned alone (stage times, the roofline's kernel time, `one_grid_at_a_time`).  Grid info (N_tri, k-means)
is computed once (FP64, exact counts) and reported separately.  Two paths
are timed: split-BF16 MFMA (the value) and the reference-structured FP32
sweep per triangle; both are compared with the FP64 path for the relative
B error.

Latency (the metric's p50/p99): a serving run of seeded fast-path jobs
(N uniform over {128,192,256,384,512,768}, nb uniform over [8, 40], auto or
cross <AAB> with probability 1/2) with Poisson arrivals at 80% of the
capacity measured by a saturated run of the same jobs.  Batched cross: 64
<AAB> pairs (32 at 128^3, 24 at 256^3, 8 at 384^3, nb in [8, 24]) replayed
as one HIP graph.

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ('grids/sec at 512^3, 32 bins (split-BF16 MFMA vs FP32 plain) + p50/p99 job latency')
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(args, torch, np):
    from bskit_amd import bins as B
    from bskit_amd import synth
    from bskit_amd.engine import cic_paint
    N, L, nb = args.nmesh, args.box, args.nbin
    kf = 2 * math.pi / L
    kLU = B.generate_bin_edge_list(kf, kf + nb * 2 * kf + kf, 2 * kf)
    assert len(kLU) == nb
    k_low = np.concatenate([kLU[:, 0], [kf]])
    k_high = np.concatenate([kLU[:, 1], [kf]])
    special = nb  # shell [k_f, k_f]: |k1| = k_f exactly
    sets = {
        'all': B.closed_triangle_indices(kLU, 1),
        'equilateral': B.generate_equilateral_triangle_bin_list(kf, kf + nb * 2 * kf + kf, 2 * kf,
                                                                return_indices=True),
        'isosceles_m2': B.generate_isosceles_triangle_bin_list(kf, kf + nb * 2 * kf + kf, 2 * kf,
                                                               isos_mult=2, return_indices=True),
        'squeezed_kf': np.array([[special, j, j] for j in range(1, nb)]),
    }
    canon = set()
    tri = []
    for name, t in sets.items():
        for row in np.asarray(t).tolist():
            key = tuple(sorted(row, reverse=True))
            if key not in canon:
                canon.add(key)
                tri.append(key)
    tri = np.array(tri, dtype=np.int32)
    t0 = time.perf_counter()
    pos = synth.zeldovich_particles(args.npart_side, L, seed=args.seed, sigma_disp=5.0)
    t_pos = time.perf_counter() - t0
    pos_d = torch.from_numpy(pos).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    field = cic_paint(pos_d, N, L)
    torch.cuda.synchronize()
    t_cic = time.perf_counter() - t0
    del pos_d
    return dict(N=N, L=L, kf=kf, k_low=k_low, k_high=k_high, tri=tri, sets=sets, field=field,
                field_host=None, t_pos=t_pos, t_cic=t_cic, pos=pos)


class StageTimer:
    """HIP-event timing of the stages of one step on torch's current stream."""

    def __init__(self, torch):
        self.torch = torch
        self.ev = {}

    def mark(self, name):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record()
        self.ev.setdefault(name, []).append(e)

    def durations(self, a, b):
        return [x.elapsed_time(y) for x, y in zip(self.ev[a], self.ev[b])]

    def pair(self, name):
        """Two fresh events (start, end) recorded by the callee, kept under name."""
        e = (self.torch.cuda.Event(enable_timing=True), self.torch.cuda.Event(enable_timing=True))
        self.ev.setdefault(name, []).append(e)
        return e

    def pair_durations(self, name):
        return [a.elapsed_time(b) for a, b in self.ev.get(name, [])]


def run_path(eng, wl, path, fmt, steps, warmup, torch, ntri_dev, timer=None, stages=True,
             inflight=1):
    """Time `steps` grids of one path.  With inflight > 1 consecutive grids
    go to `inflight` HIP streams in turn, each with its own spectrum, shells,
    output and workspace, so one grid's FFT passes overlap another's triple
    sums; every grid still runs forward + shells + all triangle sums +
    normalisation.  timer: the kernel events of the triple sums are recorded
    in every step, the stage marks only with stages=True."""
    from bskit_amd import _lib as L
    from bskit_amd.engine import Workspace
    # L^6/N^3 sum over the N^3 mesh == L^6/S^3 sum over the S^3 shell mesh
    scale = wl['L'] ** 6 / float(eng.ncell)
    dtype = torch.float64 if fmt == L.FMT_F64 else torch.float32
    ctype = torch.complex128 if fmt == L.FMT_F64 else torch.complex64
    field = wl['field'].to(dtype)
    # level layout: write only the parts of each shell the sums read
    use_parts = path == 'mfma' and eng.uses_levels(fmt)
    parts = eng.shell_parts(wl['tri']) if use_parts else None
    streams = [torch.cuda.current_stream()] + [torch.cuda.Stream() for _ in range(inflight - 1)]
    ctx = []
    for s in streams:
        with torch.cuda.stream(s):
            ctx.append(dict(
                shells=eng.alloc_shells(fmt, eng.nb) if path != 'zfused' else None,
                out=torch.empty(len(wl['tri']), dtype=torch.float64, device='cuda'),
                dk=torch.empty((eng.MY, eng.MX, eng.MZ), dtype=ctype, device='cuda'),
                work=eng.work if s is streams[0] else Workspace(eng.device)))
    torch.cuda.synchronize()

    def step(i, t):
        s, c = streams[i % inflight], ctx[i % inflight]
        mark = (lambda name: t.mark(name)) if (t is not None and stages) else (lambda name: None)
        with torch.cuda.stream(s):
            mark('start')
            dk = eng.forward(field, stream=s, work=c['work'], out=c['dk'])
            mark('fwd')
            kev = t.pair('kernel') if (t is not None and path in ('mfma', 'zfused')) else None
            if path == 'zfused':  # shells x + y passes, then z c2r fused into the sums
                eng.bispectrum_zfused(dk, wl['tri'], scale=scale, out=c['out'], stream=s,
                                      work=c['work'], kernel_events=kev,
                                      after_shells=lambda: mark('shells'))
            else:
                eng.shells(dk, fmt=fmt, dtype=dtype, out=c['shells'], parts=parts, stream=s,
                           work=c['work'])
                mark('shells')
                eng.triples([c['shells']], wl['tri'], path=path, scale=scale, out=c['out'],
                            stream=s, work=c['work'], kernel_events=kev)
            mark('triples')
            b = c['out'] / ntri_dev
            mark('end')
        return b

    for i in range(warmup):
        step(i, None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        b = step(i, timer)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    return b, elapsed


def config_accuracy(torch, np):
    """Max relative B error of the FP32 and split-BF16 paths against the FP64
    path (all on the GPU; the FP64 path is pinned to the host FP64 oracle by
    the GPU tests) for the two smaller workloads: a 64^3 FP64 field with 8
    bins of 2k_f (all triangles, T = 98) and a 256^3 FP32 field with 16 bins
    of 3k_f (T = 564), both the seeded local-quadratic field."""
    import bskit_amd
    from bskit_amd import synth
    out = {}
    kf = 2 * math.pi / 1000.0
    for name, N, dt, kw in (('64^3 fp64, 8 bins of 2k_f, T=98', 64, np.float64,
                             dict(kmin=kf, kmax=18 * kf, dk=2 * kf)),
                            ('256^3 fp32, 16 bins of 3k_f, T=564', 256, np.float32,
                             dict(kmin=kf, kmax=50.5 * kf, dk=3 * kf))):
        field = synth.quadratic_field(N, 1000.0, seed=1, dtype=dt)
        res = {}
        for prec in ('fp64', 'fp32', 'bf16x3'):
            src = field if prec == 'fp64' else field.astype(np.float32)
            fb = bskit_amd.FFTBispectrum(src, Nmesh=N, BoxSize=1000.0, precision=prec, **kw)
            res[prec] = fb.measure_bispectrum_faster()['B']
        ref = res['fp64']
        out[name] = {'T': int(len(ref)),
                     'fp32_max_rel_err': float(np.max(np.abs(res['fp32'] / ref - 1))),
                     'bf16x3_max_rel_err': float(np.max(np.abs(res['bf16x3'] / ref - 1)))}
    return out


def cpu_host_info():
    """CPU model string and NUMA layout of the host (Linux sysfs/procfs)."""
    model = None
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    nodes = {}
    base = '/sys/devices/system/node'
    try:
        for d in sorted(os.listdir(base)):
            if d.startswith('node') and d[4:].isdigit():
                with open(os.path.join(base, d, 'cpulist')) as f:
                    nodes[d] = f.read().strip()
    except OSError:
        pass
    try:
        allowed = sorted(os.sched_getaffinity(0))
    except AttributeError:
        allowed = []
    return {'cpu_model': model, 'numa_nodes': nodes, 'threads_allowed': len(allowed),
            'logical_cpus': os.cpu_count()}


def profile_counters(name):
    """{counter: per-launch mean} from a committed rocprofv3 counter summary
    table (profiles/<name>), or {} when absent."""
    path = os.path.join(ROOT, 'profiles', name)
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as f:
        for line in f:
            w = [x.strip() for x in line.strip().strip('|').split('|')]
            if len(w) == 2:
                try:
                    out[w[0]] = float(w[1])
                except ValueError:
                    pass
    return out


def profiled_traffic(kernel, nmesh, shell_mesh):
    """HBM bytes per launch of kernel from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel calibration): the
    row named by the bare kernel name (the mean over a grid's launches of
    all its instances) if there is one, else its only instance; None when
    no summary covers this workload."""
    path = os.path.join(ROOT, 'profiles', 'hbm_traffic_%d_s%d.txt' % (nmesh, shell_mesh))
    if not os.path.exists(path):
        return None
    exact, inst = None, []
    with open(path) as f:
        for line in f:
            w = line.split()
            if not w or line.startswith('#'):
                continue
            if w[0] == kernel:
                exact = float(w[-1])
            elif w[0].split('<')[0] == kernel:
                inst.append(float(w[-1]))
    if exact is not None:
        return exact
    return inst[0] if len(inst) == 1 else None


def run_serving(args, torch, np):
    """p50/p99 latency of the job mix at args.load x measured capacity."""
    from bskit_amd import serve
    t0 = time.perf_counter()
    pool = serve.make_pool(box=args.box, seed=args.seed + 100)
    jobs = serve.make_jobs(args.jobs, seed=args.seed, box=args.box, pool=pool)
    srv = serve.Server(nstreams=4)
    for j in jobs:
        srv.prepare(j)
    log('[serve] %d jobs prepared (%d engines) in %.1fs' % (len(jobs), len(srv.engines),
                                                            time.perf_counter() - t0))
    # capacity: every job released at t = 0 (also warms workspaces and caches)
    for j in jobs:
        j.arrival = 0.0
    sat = srv.run(jobs)
    capacity = len(jobs) / sat['makespan_s']
    log('[serve] saturated: %.2fs, capacity %.1f jobs/s' % (sat['makespan_s'], capacity))
    serve.reset(jobs)
    rate = args.load * capacity
    serve.poisson_arrivals(jobs, rate, seed=args.seed)
    res = srv.run(jobs, log=log)
    st = serve.latency_stats(jobs)
    gpu_s = sum(2 if j.nfield == 2 else 1 for j in jobs)
    log('[serve] poisson %.1f jobs/s: p50 %.2f ms p99 %.2f ms' % (rate, st['p50_ms'],
                                                                 st['p99_ms']))
    out = dict(st)
    out.update({'jobs': len(jobs), 'arrival_rate_jobs_per_s': rate,
                'capacity_jobs_per_s': capacity, 'load': args.load,
                'makespan_s': res['makespan_s'],
                'throughput_jobs_per_s': len(jobs) / res['makespan_s'],
                'peak_footprint_GB': res['peak_footprint_bytes'] / 1e9,
                'hbm_budget_GB': srv.budget / 1e9, 'streams': len(srv.streams),
                'tuned_classes': len(srv.tuned),
                'fields': gpu_s,
                'mix': 'N uniform {128,192,256,384,512,768}, nb uniform [8,40], '
                       'auto/cross 1:1, fast-path B_unnorm per job copied to host; '
                       'inputs: 2 seeded correlated field pairs per N resident in HBM'})
    for j in jobs:
        j.result = None
    del pool, jobs, srv
    torch.cuda.empty_cache()
    return out


def run_batch(args, torch, np):
    """64 cross <AAB> pairs as one HIP graph: pairs/s over several replays."""
    from bskit_amd import batch
    pairs = batch.config_pairs(seed=args.seed, box=args.box)
    cb = batch.CrossBatch(pairs, nbranch=4)
    cb.capture()
    cb.replay()
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        cb.replay()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    # the same batch launched eagerly on the branch streams (host-issued)
    t0 = time.perf_counter()
    for _ in range(reps):
        cb.run_eager()
    torch.cuda.synchronize()
    el_eager = (time.perf_counter() - t0) / reps
    # algorithmic HBM bytes of one batch: per pair, the two forward
    # transforms (field read, compact spectrum written), the two shell sets
    # (spectrum read, the layout parts written) and the pair kernel's staged
    # shell reads (each segment's window once); the whole batch over one
    # replay against the HBM peak
    alg = 0.0
    for p in cb.pairs:
        eng = p.eng
        H = eng.MX * eng.MY * eng.MZ
        alg += 2 * (4.0 * eng.N ** 3 + 8.0 * H)
        if p.path == 'mfma':
            plan = eng.pair_plan(p.tri, eng.nb, 1, levels=p.sets[0].layout is not None)
            if p.sets[0].layout is not None:
                parts = eng.layout.part_cells()
                written = sum(4.0 * sum(parts[q] for q in range(4) if pm >> q & 1)
                              for pm in p.parts.host)
                alg += sum(4.0 * wn * sum(parts[q] for q in range(4) if pm >> q & 1)
                           for pm, w0, wn, *_ in plan.segments)
            else:
                written = 4.0 * eng.nb * eng.ncell
                alg += 4.0 * eng.nb * eng.ncell
            alg += 2 * (8.0 * H + written)
    roof = {'kernel': 'whole batch (forward, shells and pair kernels of every pair)',
            'bound': 'hbm', 'achieved': alg / el / 1e9, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
            'frac': alg / el / 1e9 / HBM_PEAK_GBS, 'traffic': None,
            'algorithmic_bytes': alg}
    res = {'pairs': len(pairs), 'triangles': cb.ntri, 'graph_ms': el * 1e3,
           'eager_ms': el_eager * 1e3, 'roofline': roof,
           'pairs_per_s': len(pairs) / el, 'replays': reps, 'branches': cb.nbranch,
           'tuned_classes': {'%d^3/S%d' % k: v for k, v in cb.tuned.items()},
           'layout': '32 @128^3, 24 @256^3, 8 @384^3, nb uniform [8,24], <AAB>'}
    log('[batch] %d pairs in %.2f ms per replay of the branch graphs (%.2f ms eager)'
        % (len(pairs), el * 1e3, el_eager * 1e3))
    del cb, pairs
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=50)
    ap.add_argument('--plain-steps', type=int, default=2)
    ap.add_argument('--inflight', type=int, default=3,
                    help='grids in flight on as many HIP streams in the timed run')
    ap.add_argument('--nmesh', type=int, default=512)
    ap.add_argument('--nbin', type=int, default=32)
    ap.add_argument('--npart-side', type=int, default=256)
    ap.add_argument('--box', type=float, default=1000.0)
    ap.add_argument('--seed', type=int, default=1)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline sample')
    ap.add_argument('--no-fp64', action='store_true', help='skip the FP64 accuracy reference')
    ap.add_argument('--no-plain', action='store_true', help='skip the FP32 plain path')
    ap.add_argument('--no-serve', action='store_true', help='skip the serving (latency) run')
    ap.add_argument('--no-batch', action='store_true', help='skip the graph-batched cross run')
    ap.add_argument('--no-tune', action='store_true', help='skip the autotuner (default variants)')
    ap.add_argument('--no-zfused', action='store_true',
                    help='time the separate shell and triple-sum passes only')
    ap.add_argument('--no-levels', action='store_true',
                    help='sum every triangle over the whole shell mesh (no sublattice levels)')
    ap.add_argument('--jobs', type=int, default=2000, help='jobs in the serving run')
    ap.add_argument('--load', type=float, default=0.8, help='arrival rate / measured capacity')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='BYTES',
                    help='only run 5 device-to-device copies of BYTES (HBM counter calibration)')
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU: the FFT bispectrum path is sharded by job across GPUs '
              '(no cross-GPU traffic), so run one process per GPU; --gpus %d is not supported.'
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from bskit_amd import _lib as L
    from bskit_amd.engine import GridEngine

    if not torch.cuda.is_available():
        print('bench.py needs a GPU', file=sys.stderr)
        sys.exit(1)
    torch.cuda.init()
    if args.copy_calibration:
        n = args.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes': 4 * n, 'copies': 5}), flush=True)
        return
    log('[bench] building workload N=%d nbin=%d' % (args.nmesh, args.nbin))
    wl = build_workload(args, torch, np)
    T = len(wl['tri'])
    eng = GridEngine(wl['N'], wl['L'], wl['k_low'], wl['k_high'], cic_window=wl['N'])
    eng.levels = not args.no_levels
    log('[bench] T=%d shells=%d compact m=%d shell mesh S=%d' % (T, eng.nb, eng.m, eng.S))

    tuned = None
    if not args.no_tune:
        t0 = time.perf_counter()
        tuned = eng.autotune(tri=wl['tri'], field=wl['field'], log=log)
        tuned = {'variants': tuned['variants'], 'mfma_nwg': tuned.get('mfma_nwg'),
                 'zfused_nwg': tuned.get('zfused_nwg'), 'zfused': tuned.get('zfused'),
                 'seconds': time.perf_counter() - t0}
    t0 = time.perf_counter()
    ntri, kbar = eng.grid_info(wl['tri'])
    torch.cuda.synchronize()
    t_gridinfo = time.perf_counter() - t0
    log('[bench] grid info %.2fs, N_tri range %d..%d' % (t_gridinfo, ntri.min(), ntri.max()))
    ntri_dev = torch.from_numpy(np.where(ntri > 0, ntri, np.nan)).cuda()

    # main path: the z c2r fused into the triple-sum kernel where it is
    # supported and the autotuner found it faster, else shells on the shell
    # mesh + the pair kernel
    zf_ok = not args.no_zfused and eng.zfused_supported(wl['tri'])
    main_path = 'zfused' if zf_ok and eng.prefer_zfused else 'mfma'
    # headline: args.inflight grids in flight on as many streams, through
    # bskit_amd.pipeline.GridPipeline (the library's throughput mode), no
    # events; its pair-kernel workgroup count is tuned in that mode (fewer
    # workgroups leave CUs to the other streams' FFT passes).  Then the same
    # path one grid at a time with stage marks and HIP events around the
    # pair-kernel launches of every step (stage times, the one-at-a-time rate
    # and the roofline's kernel time).
    nwg_in_flight = None
    if args.inflight > 1 and main_path == 'mfma':
        from bskit_amd.pipeline import GridPipeline
        pipe = GridPipeline(eng, wl['tri'], inflight=args.inflight,
                            scale=wl['L'] ** 6 / float(eng.ncell), norm=ntri_dev)
        if not args.no_tune:
            rates = pipe.tune(wl['field'])
            nwg_in_flight = pipe.nwg
            log('[bench] pair-kernel workgroups in flight: %s -> %d'
                % ({k: round(v, 1) for k, v in rates.items()}, nwg_in_flight))
            if tuned is not None:
                tuned['mfma_nwg_in_flight'] = nwg_in_flight
        for _ in range(args.warmup):
            pipe.submit(wl['field'])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b_mfma = pipe.submit(wl['field'])
        torch.cuda.synchronize()
        el_mfma = time.perf_counter() - t0
        del pipe
    else:
        b_mfma, el_mfma = run_path(eng, wl, main_path, L.FMT_BF16X2, args.steps, args.warmup,
                                   torch, ntri_dev, inflight=args.inflight)
    ms_step = el_mfma / args.steps * 1e3
    grids_mfma = args.steps / el_mfma
    log('[bench] %s, %d in flight: %.3f ms/grid, %.2f grids/s'
        % (main_path, args.inflight, ms_step, grids_mfma))
    timer = StageTimer(torch)
    b_seq, el_seq = run_path(eng, wl, main_path, L.FMT_BF16X2, args.steps, args.warmup, torch,
                             ntri_dev, timer)
    # the two runs may use different pair-kernel workgroup counts, which
    # split the cells into other FP32 chains: the sums then differ by rounding
    ok = torch.isfinite(b_seq) & torch.isfinite(b_mfma)
    sequential = {'grids_per_s': args.steps / el_seq, 'ms_per_grid': el_seq / args.steps * 1e3,
                  'max_diff_vs_in_flight': float((b_seq[ok] - b_mfma[ok]).abs().max()
                                                 / b_seq[ok].abs().max())}
    log('[bench] %s, one at a time: %.3f ms/grid, %.2f grids/s'
        % (main_path, sequential['ms_per_grid'], sequential['grids_per_s']))
    stage_ms = {
        'forward': float(np.mean(timer.durations('start', 'fwd'))),
        'shells': float(np.mean(timer.durations('fwd', 'shells'))),
        'triples_mfma': float(np.mean(timer.durations('shells', 'triples'))),
        'normalise': float(np.mean(timer.durations('triples', 'end'))),
    }
    log('[bench] stages (ms): %s' % stage_ms)
    b_mfma = b_mfma.cpu().numpy()
    alt = None
    if zf_ok:  # the other implementation of the same step, reported beside it
        alt_path = 'mfma' if main_path == 'zfused' else 'zfused'
        ts = StageTimer(torch)
        b_alt, el_alt = run_path(eng, wl, alt_path, L.FMT_BF16X2, args.steps, args.warmup,
                                 torch, ntri_dev, ts)
        alt = {'path': alt_path, 'ms_per_grid': el_alt / args.steps * 1e3,
               'grids_per_s': args.steps / el_alt,
               'shells_ms': float(np.mean(ts.durations('fwd', 'shells'))),
               'triples_ms': float(np.mean(ts.durations('shells', 'triples'))),
               'kernel_ms': float(np.mean(ts.pair_durations('kernel')))}
        log('[bench] %s: %.3f ms/grid' % (alt_path, alt['ms_per_grid']))

    res_plain = res_tiled = None
    if not args.no_plain:
        tp = StageTimer(torch)
        b_plain, el_plain = run_path(eng, wl, 'plain', L.FMT_F32, args.plain_steps, 1, torch,
                                     ntri_dev, tp)
        res_plain = {'grids_per_s': args.plain_steps / el_plain,
                     'ms_per_grid': el_plain / args.plain_steps * 1e3,
                     'triples_ms': float(np.mean(tp.durations('shells', 'triples')))}
        b_plain = b_plain.cpu().numpy()
        log('[bench] fp32 plain: %.1f ms/grid' % res_plain['ms_per_grid'])
        # the LDS-tiled FP32 vector path (every shell cell staged once per
        # 32-cell tile for all triangles): the strongest FP32 baseline
        tt = StageTimer(torch)
        b_tiled, el_tiled = run_path(eng, wl, 'tiled', L.FMT_F32, args.plain_steps, 1, torch,
                                     ntri_dev, tt)
        res_tiled = {'grids_per_s': args.plain_steps / el_tiled,
                     'ms_per_grid': el_tiled / args.plain_steps * 1e3,
                     'triples_ms': float(np.mean(tt.durations('shells', 'triples')))}
        b_tiled = b_tiled.cpu().numpy()
        log('[bench] fp32 tiled: %.1f ms/grid' % res_tiled['ms_per_grid'])

    acc = {}
    if not args.no_fp64:
        b64, _ = run_path(eng, wl, 'plain', L.FMT_F64, 1, 0, torch, ntri_dev)
        b64 = b64.cpu().numpy()
        ok = np.isfinite(b64) & (b64 != 0)
        rel_m = np.abs(b_mfma[ok] / b64[ok] - 1)
        acc['bf16x3_max_rel_err'] = float(rel_m.max())
        acc['bf16x3_median_rel_err'] = float(np.median(rel_m))
        acc['bf16x3_frac_bins_above_1e-3'] = float(np.mean(rel_m > 1e-3))
        if res_plain is not None:
            rel_p = np.abs(b_plain[ok] / b64[ok] - 1)
            acc['fp32_max_rel_err'] = float(rel_p.max())
            acc['fp32_tiled_max_rel_err'] = float(np.abs(b_tiled[ok] / b64[ok] - 1).max())
        log('[bench] accuracy vs FP64: %s' % acc)

    # roofline of the dominant kernel: the split-BF16 MFMA pair kernel, timed
    # by HIP events around its launches inside every timed step.  The
    # separate path sums each triangle over the sparsest sublattice of the
    # shell mesh on which its sum is exact (level L: S^3 / 2^L cells), one
    # launch per segment of the level layout (its item-list passes folded
    # into it); figures are per launch (grid totals / launches).
    N = wl['N']
    ncell = float(eng.ncell)                     # shell-mesh cells S^3
    nsh = eng.nb
    levels = eng.uses_levels()
    lv = eng.triangle_levels(wl['tri']) if levels else np.zeros(T, np.int32)
    lcount = np.bincount(lv, minlength=L.LEVELS)
    # one multiply-add per triangle per cell summed
    flops_sep = 2.0 * sum(lcount[k] * ncell / 2 ** k for k in range(L.LEVELS))
    flops_fused = 2.0 * T * ncell

    def roofline(kernel, grid_ms, alg_bytes, launches, flops, plan):
        t_k = grid_ms * 1e-3
        ai = flops / alg_bytes
        bound = 'hbm' if ai * HBM_PEAK_GBS * 1e9 < BF16_PEAK_TFLOPS * 1e12 else 'mfma'
        achieved = alg_bytes / t_k / 1e9
        ctr = profile_counters('%s_%d_counters.md' % (kernel, N)) if kernel == 'k_triple_pairs' \
            else {}
        # issued MFMA (16x16x32 bf16: 16384 flop each) per launch from the
        # committed counters, over this run's kernel time, against the dense
        # bf16 peak
        mfma_util = (ctr['SQ_INSTS_MFMA'] * 16384.0 / (t_k / launches) / (BF16_PEAK_TFLOPS * 1e12)
                     if 'SQ_INSTS_MFMA' in ctr else None)
        return {'kernel': kernel, 'bound': bound, 'achieved': achieved,
                'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                'traffic': profiled_traffic(kernel, N, eng.S),
                'traffic_source': ('profiles/hbm_traffic_%d_s%d.txt (rocprofv3 FETCH_SIZE + '
                                   'WRITE_SIZE per dispatch, copy-calibrated; the file names the '
                                   'commit it was measured at)' % (N, eng.S)),
                'mfma_issue_util': mfma_util,
                'counters_source': ('profiles/%s_%d_counters.md' % (kernel, N)) if ctr else None,
                'kernel_ms': grid_ms / launches, 'launches_per_grid': launches,
                'algorithmic_bytes': alg_bytes / launches,
                'useful_tflops': flops / t_k / 1e12, 'arith_intensity_flop_per_byte': ai,
                'pairs': plan.npair, 'tiles': plan.ntile, 'items': plan.nitem,
                # measured with counters: per K step a new tile costs each lane
                # 8 products and 4 bf16 splits, an item 3 MFMAs, so the vector
                # ALU binds before HBM or the matrix cores
                'limiter': ('VALU: forming the split-bf16 pair products (%d tiles for %d '
                            'items)' % (plan.ntile, plan.nitem)
                            if kernel == 'k_triple_pairs' else None)}

    # the fused kernel reads each shell's inverse y output once: kz < mz_b
    # complex values of every (x, y) row
    mz = np.minimum(eng.MZ, np.floor(np.sqrt(np.maximum(eng.n2hi, 0))).astype(int) + 1)
    d2_bytes = float(np.sum(mz)) * 8.0 * eng.S * eng.S

    def path_roofline(path, ms):
        if path == 'zfused':
            plan = eng.pair_plan(wl['tri'], nsh, 0)
            return roofline('k_triple_zrows', ms, d2_bytes, plan.npass, flops_fused, plan)
        plan = eng.pair_plan(wl['tri'], nsh, 0, levels=levels)
        # split-BF16 hi+lo of the shells each segment stages, read once
        if levels:
            parts = eng.layout.part_cells()
            alg = sum(4.0 * wn * sum(parts[q] for q in range(4) if pm >> q & 1)
                      for pm, w0, wn, *_ in plan.segments)
        else:
            alg = nsh * 4 * ncell
        # launches: a segment's item-list passes run as one launch when the
        # workgroup count is a multiple of 8 and its lists are not replicated
        nwg = eng.mfma_nwg()
        launches = sum(1 if (npass > 1 and rep == 1 and nwg % 8 == 0) else npass
                       for _, _, _, npass, _, rep in plan.segments)
        return roofline('k_triple_pairs', ms, alg, launches, flops_sep, plan)

    roof = path_roofline(main_path, float(np.mean(timer.pair_durations('kernel'))))
    if alt is not None:
        alt['roofline'] = path_roofline(alt['path'], alt['kernel_ms'])
    H = eng.MX * eng.MY * eng.MZ                 # compact delta(k) cube
    t_sh = (alt['shells_ms'] if main_path == 'zfused' else stage_ms['shells']) * 1e-3
    # delta(k) read + shells written (level layout: the cosets, then the
    # all-even and all-odd cells again)
    shells_alg = 8.0 * H + nsh * 4 * (float(eng.layout.stride) if levels else ncell)
    stage_gbs = {'shells_GBps': shells_alg / t_sh / 1e9,
                 'forward_GBps': (4.0 * N ** 3 + 8.0 * H) / (stage_ms['forward'] * 1e-3) / 1e9}
    # algorithmic bytes of one grid in the estimator's own form (shells and
    # triple reads on the N^3 input mesh, s = 4 B per shell value) beside the
    # form this path runs (shells on the S^3 shell mesh, each triangle over its
    # sublattice level): the N^3 figure over the step time exceeds the HBM
    # peak, which the exact shell-mesh and sublattice reductions explain
    Hn = N * N * (N // 2 + 1)
    bytes_n3 = (4.0 * N ** 3 + 8 * Hn) + (8 * Hn + nsh * 4.0 * N ** 3) + nsh * 4.0 * N ** 3
    bytes_s = (4.0 * N ** 3 + 8 * H) + shells_alg + roof['algorithmic_bytes'] * roof[
        'launches_per_grid'] if main_path == 'mfma' else None
    step_bytes = {'n3_mesh_form_GB': bytes_n3 / 1e9,
                  'n3_mesh_form_rate_GBps_per_step': bytes_n3 / (sequential['ms_per_grid'] * 1e-3)
                  / 1e9,
                  'shell_mesh_form_GB': None if bytes_s is None else bytes_s / 1e9,
                  'shell_mesh_form_rate_GBps_per_step': None if bytes_s is None else
                  bytes_s / (sequential['ms_per_grid'] * 1e-3) / 1e9}

    acc_cfg = None
    if not args.no_fp64:
        acc_cfg = config_accuracy(torch, np)
        log('[bench] accuracy by config: %s' % acc_cfg)

    latency = None
    if not args.no_serve:
        latency = run_serving(args, torch, np)
    batched = None
    if not args.no_batch:
        batched = run_batch(args, torch, np)

    cpu = None
    if not args.no_cpu:
        from oracle import baseline
        from oracle import estimator as est
        field_h = wl['field'].cpu().numpy()
        lo, hi = est.shell_limits(wl['k_low'], wl['k_high'], wl['kf'])
        log('[bench] cpu baseline sample...')
        r = baseline.time_sample(field_h, lo, hi, wl['tri'], n_shell_sample=4, n_tri_sample=24)
        cpu = {'value': 1.0 / r['seconds_per_grid'], 'unit': 'grids/s', 'cores': r['cores'],
               'kind': 'port', 'host': cpu_host_info(),
               'sample': ('reference-structured numpy/scipy path on %d host threads: full forward '
                          'r2c (%.1fs), %d shell c2r (%.2fs each), %d triangle sums (%.2fs each); '
                          'extrapolated to %d shells and %d triangles'
                          % (r['cores'], r['forward_s'], r['nshell_sampled'], r['shell_s'],
                             r['ntri_sampled'], r['triangle_s'], nsh, T))}
        log('[bench] cpu: %.1f s/grid' % r['seconds_per_grid'])

    line = {
        'metric': METRIC,
        'value': grids_mfma,
        'unit': 'grids/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'bf16',
        'data': 'synthetic',
        'config': {'workload': ('%d^3 CIC mesh of %d^3 Zel\'dovich particles, L=%g Mpc/h, %d bins '
                                'of 2k_f from k_f; all+equilateral+isosceles(m=2)+squeezed(|k1|=k_f)'
                                % (N, args.npart_side, wl['L'], args.nbin)),
                   'nmesh': N, 'shell_mesh': eng.S, 'nbin': args.nbin, 'nshell': nsh, 'ntri': T,
                   'path': ('split-BF16 MFMA (hi*hi+hi*lo+lo*hi), FP32 accumulate, FP64 '
                            'reduce; %s' % ('z c2r fused into the triple sums'
                                            if main_path == 'zfused' else
                                            'shells on the shell mesh, each triangle summed '
                                            'over the sparsest exact sublattice (triangles at '
                                            '1, 1/2, 1/4, 1/8 of the cells: %s)'
                                            % '/'.join(str(int(c)) for c in lcount))),
                   'grids_in_flight': args.inflight,
                   'parallelism': ('single GPU, %d grids in flight on %d HIP streams'
                                   % (args.inflight, args.inflight) if args.inflight > 1
                                   else 'single GPU')},
        'fp32_plain': res_plain,
        'fp32_tiled': res_tiled,
        # one grid at a time on both sides
        'speedup_vs_fp32_plain': ((sequential['grids_per_s'] / res_plain['grids_per_s'])
                                  if res_plain else None),
        'speedup_vs_fp32_tiled': ((sequential['grids_per_s'] / res_tiled['grids_per_s'])
                                  if res_tiled else None),
        'accuracy_by_config': acc_cfg,
        'one_grid_at_a_time': sequential,
        'stage_ms': stage_ms,
        'other_path': alt,
        'stage_GBps': stage_gbs,
        'step_algorithmic_bytes': step_bytes,
        'gridinfo_once_s': t_gridinfo,
        'autotune': tuned,
        'accuracy_vs_fp64': acc,
        'latency': latency,
        'batched_cross': batched,
        'roofline': roof,
        'cpu_baseline': cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
