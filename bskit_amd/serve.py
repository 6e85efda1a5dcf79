"""Request loop for fast-path bispectrum jobs on one GPU.

A job is one fast-path measurement (bskit/main.py:1784-1938 with
main.py:1875-1882): forward transform of one field (auto) or two fields
(cross <AAB>, the roles of main.py:624-640), every shell of its bin list
in split-BF16, and the unnormalised B of every closed triangle, copied back
to host memory.  The reference measures one grid per MPI job; here many
jobs share the GPU:

* admission: a job starts only when its HBM footprint (shells, spectra,
  workspaces) fits in the budget left by the jobs in flight (default
  min(256 GB, 85% of free HBM));
* continuous batching: admitted jobs are issued at once onto the stream
  with the fewest jobs in flight (default-priority streams), so small and
  large grids overlap on the device instead of waiting for a batch to drain;
* per-key caches: one GridEngine (plan, thresholds, MFMA layouts) per
  (Nmesh, BoxSize, bins) key, created before the timed run by ``prepare``.

Every reduction in the path runs in a fixed order, so a job's result does
not depend on which other jobs share the device.
"""
import collections
import ctypes
import dataclasses
import math
import time

import numpy as np
import torch

from . import _lib as L
from . import bins as B
from .engine import PATH_MFMA, PATH_PLAIN, GridEngine, Workspace, require_device

MESH_SIZES = (128, 192, 256, 384, 512, 768)


@dataclasses.dataclass
class Job:
    jid: int
    nmesh: int
    box: float
    kLU: np.ndarray
    nfield: int                  # 1: auto <AAA>; 2: cross <AAB>
    fields: list                 # device float32 N^3 tensors
    arrival: float = 0.0         # seconds after the start of the run
    # filled by the server
    tri: np.ndarray = None
    path: str = None
    footprint: int = 0
    result: torch.Tensor = None  # pinned host float64 [T]
    issued: float = None
    done: float = None
    stream: int = None

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def _bins_key(kLU):
    return tuple((float(a), float(b)) for a, b in np.asarray(kLU, dtype=np.float64))


class Server:
    """Admission + continuous batching of Jobs on one device."""

    def __init__(self, device=None, hbm_budget=None, nstreams=4, poll_s=2e-5, assign='work',
                 depth=None, autotune=True):
        """assign: the stream an arriving job is issued to -- 'work', the one
        with the least outstanding work (the footprints of its jobs in flight,
        which grow with the mesh and the shell set), so a small job does not
        queue behind a large one where another stream is nearly idle; or
        'count', the one with the fewest jobs in flight.  depth: at most this
        many jobs in flight per stream (None: no limit); arrivals beyond it wait
        in the host queue and go to the first stream with room.  autotune: tune
        the radix plans and the pair kernel's workgroup count once per mesh
        class (Nmesh, shell mesh) on the first job of the class, in prepare
        (bskit_amd.tune.tune_class), and apply it to every engine of the
        class."""
        if assign not in ('work', 'count'):
            raise ValueError("assign must be 'work' or 'count'")
        if depth is not None and depth < 1:
            raise ValueError('depth must be >= 1 or None')
        self.assign = assign
        self.depth = depth
        self.autotune = autotune
        self.tuned = {}  # mesh class -> configuration applied
        self.device = require_device(device)
        free, _ = torch.cuda.mem_get_info(self.device)
        self.budget = int(hbm_budget) if hbm_budget else int(min(256e9, 0.85 * free))
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(nstreams)]
        self.works = [Workspace(self.device) for _ in range(nstreams)]
        self.poll_s = poll_s
        self.engines = {}
        self._tri = {}

    # ------------------------------------------------------------ host side
    def engine(self, nmesh, box, kLU):
        key = (int(nmesh), float(box), _bins_key(kLU))
        eng = self.engines.get(key)
        if eng is None:
            kLU = np.asarray(kLU, dtype=np.float64)
            eng = GridEngine(nmesh, box, kLU[:, 0], kLU[:, 1], device=self.device)
            self.engines[key] = eng
        return eng

    def footprint(self, eng, nfield, tri=None):
        """HBM bytes a job holds while in flight: its shells, spectra and
        output, plus every workspace its kernels use (forward, shells, and
        the MFMA partials at the workgroup count it runs with, or the plain
        path's block sums).  Workspaces live in the stream's grow-only
        buffers, which run() releases whenever the stream goes idle, so
        counting them per job bounds what the jobs in flight hold."""
        nb = eng.nb
        sz = L.size_out()
        L.call('bsk_forward_workspace', eng._plan, L.DT_F32, ctypes.byref(sz))
        fwd = sz.value
        L.call('bsk_shells_workspace', eng._plan, L.DT_F32, nb, eng.S, ctypes.byref(sz))
        shw = sz.value
        mfma = GridEngine.mfma_supported(nb, nfield, None)
        fmt = L.FMT_BF16X2 if mfma else L.FMT_F32
        shells = nfield * eng.shell_bytes(fmt)
        spectra = nfield * eng.MX * eng.MY * eng.MZ * 8
        trw = 0
        ntri = 0 if tri is None else len(tri)
        if ntri and mfma:
            mode = {1: L.MODE_AUTO, 2: L.MODE_AAB}[nfield]
            plan = eng.pair_plan(tri, nb, mode, levels=eng.uses_levels())
            L.call('bsk_triples_mfma_workspace', plan.handle, eng.mfma_nwg(), ctypes.byref(sz))
            trw = sz.value
        elif ntri:
            L.call('bsk_triples_plain_workspace', min(ntri, 65535), eng.plain_nchunk(),
                   ctypes.byref(sz))
            trw = sz.value
        out = 8 * ntri
        return int(fwd + shw + shells + spectra + trw + out)

    def prepare(self, job):
        """Engine, triangles, path and footprint of a job (host work, untimed)."""
        eng = self.engine(job.nmesh, job.box, job.kLU)
        key = (_bins_key(job.kLU), job.nfield)
        tri = self._tri.get(key)
        if tri is None:
            tri = B.closed_triangle_indices(np.asarray(job.kLU), job.nfield).astype(np.int32)
            self._tri[key] = tri
        job.tri = tri
        job.path = PATH_MFMA if GridEngine.mfma_supported(eng.nb, job.nfield, tri) else PATH_PLAIN
        if self.autotune and job.path == PATH_MFMA and len(tri):
            from . import tune
            cfg = tune.tune_class(eng, tri=tri, field=job.fields[0], nfield=job.nfield)
            self.tuned[(eng.N, eng.S)] = {'variants': cfg.get('variants'),
                                          'mfma_nwg': cfg.get('mfma_nwg')}
        if job.path == PATH_MFMA and len(tri):
            mode = {1: L.MODE_AUTO, 2: L.MODE_AAB}[job.nfield]
            eng.pair_plan(tri, eng.nb, mode, levels=eng.uses_levels())
            if eng.uses_levels():
                eng.shell_parts(tri, job.nfield)
        eng.triangles_device(tri)
        job.footprint = self.footprint(eng, job.nfield, tri)
        if job.footprint > self.budget:
            raise MemoryError('job %d needs %.1f GB, above the %.1f GB budget'
                              % (job.jid, job.footprint / 1e9, self.budget / 1e9))
        return job

    # ---------------------------------------------------------- device side
    def _issue(self, job, si):
        """Enqueue the whole job on stream si; returns its completion event."""
        eng = self.engine(job.nmesh, job.box, job.kLU)
        s = self.streams[si]
        w = self.works[si]
        fmt = L.FMT_BF16X2 if job.path == PATH_MFMA else L.FMT_F32
        # order the job after whatever wrote its fields on the caller's
        # stream, and keep their memory alive until this stream is done
        s.wait_stream(torch.cuda.current_stream(self.device))
        for f in job.fields[:job.nfield]:
            f.record_stream(s)
        with torch.cuda.stream(s):
            sets = []
            for f in job.fields[:job.nfield]:
                dk = eng.forward(f, work=w)
                sh = eng.alloc_shells(fmt, eng.nb)
                parts = (eng.shell_parts(job.tri, job.nfield)
                         if job.path == PATH_MFMA and sh.layout is not None else None)
                sets.append(eng.shells(dk, fmt=fmt, out=sh, work=w, parts=parts))
            out = eng.triples(sets, job.tri, path=job.path, scale=job.box ** 6 / float(eng.ncell),
                              work=w)
            host = torch.empty(len(job.tri), dtype=torch.float64, pin_memory=True)
            host.copy_(out, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(s)
        job.result = host
        job.stream = si
        return ev

    def run_one(self, job):
        """Run a single job alone and wait for it (reference result)."""
        if job.tri is None:
            self.prepare(job)
        ev = self._issue(job, 0)
        ev.synchronize()
        return job.result.clone()

    def run(self, jobs, log=None):
        """Serve jobs by arrival time; fills issued/done (seconds from start)."""
        for j in jobs:
            if j.tri is None:
                self.prepare(j)
        pending = collections.deque(sorted(jobs, key=lambda j: (j.arrival, j.jid)))
        inflight = []
        per_stream = [0] * len(self.streams)
        work = [0] * len(self.streams)  # footprints in flight per stream
        used = 0
        peak = 0
        t0 = time.perf_counter()
        nextlog = 1.0
        while pending or inflight:
            now = time.perf_counter() - t0
            keep = []
            for j, ev in inflight:
                if ev.query():
                    j.done = now
                    used -= j.footprint
                    per_stream[j.stream] -= 1
                    work[j.stream] -= j.footprint
                    if per_stream[j.stream] == 0:
                        # idle stream: its workspaces are no longer counted
                        # in any footprint, so give them back
                        self.works[j.stream].release()
                else:
                    keep.append((j, ev))
            inflight = keep
            while pending and pending[0].arrival <= now:
                j = pending[0]
                if inflight and used + j.footprint > self.budget:
                    break
                load = work if self.assign == 'work' else per_stream
                room = [i for i in range(len(self.streams))
                        if self.depth is None or per_stream[i] < self.depth]
                if not room:
                    break
                pending.popleft()
                si = min(room, key=lambda i: (load[i], per_stream[i], i))
                j.issued = time.perf_counter() - t0
                ev = self._issue(j, si)
                per_stream[si] += 1
                work[si] += j.footprint
                used += j.footprint
                peak = max(peak, used)
                inflight.append((j, ev))
            if log is not None and now >= nextlog:
                ndone = sum(1 for j in jobs if j.done is not None)
                log('[serve] t=%.1fs done %d/%d in flight %d' % (now, ndone, len(jobs),
                                                               len(inflight)))
                nextlog += 1.0
            if pending and not inflight:
                wait = pending[0].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
            else:
                time.sleep(self.poll_s)
        makespan = time.perf_counter() - t0
        return {'makespan_s': makespan, 'peak_footprint_bytes': peak}


# ---------------------------------------------------------------- workloads
def serving_bins(nmesh, nbin, kf):
    """nbin bins from k_f: width 2 k_f while the last edge stays below 0.45 N k_f,
    else width k_f (edges by generate_bin_edge_list, quirk-aware kmax)."""
    w = 2 if (2 * nbin + 1) <= 0.45 * nmesh else 1
    kLU = B.generate_bin_edge_list(kf, kf + nbin * w * kf + w * kf / 2.0, w * kf)
    return kLU[:nbin]


def make_jobs(njobs, seed=0, box=1000.0, sizes=MESH_SIZES, nbin_range=(8, 40), pool=None,
              device=None):
    """Seeded job mix: N uniform over sizes, nb uniform over nbin_range, auto or
    cross with probability 1/2.  pool[N] = list of (A, B) device field pairs;
    job j uses pair j mod len(pool[N])."""
    rng = np.random.default_rng(seed)
    kf = 2 * math.pi / box
    jobs = []
    for jid in range(njobs):
        n = int(rng.choice(sizes))
        nb = int(rng.integers(nbin_range[0], nbin_range[1] + 1))
        nf = 1 if rng.random() < 0.5 else 2
        pair = pool[n][jid % len(pool[n])]
        jobs.append(Job(jid=jid, nmesh=n, box=box, kLU=serving_bins(n, nb, kf), nfield=nf,
                        fields=list(pair)))
    return jobs


def make_pool(sizes=MESH_SIZES, npair=2, box=1000.0, seed=100, device='cuda'):
    """Correlated local-quadratic field pairs per mesh size, resident in HBM."""
    from . import synth
    pool = {}
    for i, n in enumerate(sizes):
        pool[n] = [synth.correlated_pair_device(n, box, seed + 1000 * i + 2 * p,
                                                seed + 1000 * i + 2 * p + 1, device)
                   for p in range(npair)]
    return pool


def poisson_arrivals(jobs, rate, seed=0):
    """Exponential inter-arrival times at `rate` jobs/s."""
    rng = np.random.default_rng(seed + 7)
    t = 0.0
    for j in jobs:
        t += rng.exponential(1.0 / rate)
        j.arrival = t
    return jobs


def reset(jobs):
    for j in jobs:
        j.result = None
        j.issued = None
        j.done = None
        j.stream = None


def latency_stats(jobs):
    lat = np.array([j.latency for j in jobs], dtype=np.float64) * 1e3
    return {'p50_ms': float(np.percentile(lat, 50)), 'p99_ms': float(np.percentile(lat, 99)),
            'mean_ms': float(lat.mean()), 'max_ms': float(lat.max())}
