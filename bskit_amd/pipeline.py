"""Many grids through one engine with several grids in flight.

A run over many realisations or snapshots of one configuration (the fast
path of `measure_bispectrum_faster`, bskit/main.py:1784-1882, applied grid
after grid) is throughput-bound.  GridPipeline issues consecutive grids to
`inflight` HIP streams in turn; each stream owns its spectrum, shells,
output and workspace, so one grid's FFT passes (latency-bound) run beside
another's pair kernel (LDS/MFMA-bound).  A grid's sums do not depend on the
stream it ran on or on what ran beside it (the reductions have a fixed
order).

    pipe = GridPipeline(eng, tri, inflight=3, scale=L**6 / eng.ncell)
    outs = [pipe.submit(f) for f in fields]   # device tensors, in order
    pipe.synchronize()
"""
import time

import numpy as np
import torch

from . import _lib as L
from .engine import PATH_MFMA, Workspace


class GridPipeline:
    """Split-BF16 MFMA triple sums of grids with up to `inflight` in flight.

    eng: a GridEngine; tri: the (T, 3) triangle shell indices; nfield: 1
    for auto spectra <AAA>, 2 for cross spectra <AAB> (sides A, A, B; each
    grid is then two fields); scale: the factor of every sum (L^6 / S^3 for
    B_unnorm on the shell mesh); norm: optional float64 device tensor of T
    values each grid's sums are divided by (the cached N_tri); nwg: the pair
    kernel's workgroup count (default: the engine's; tune() picks one for
    this mode).
    """

    def __init__(self, eng, tri, inflight=3, scale=1.0, norm=None, nwg=None, nfield=1):
        if inflight < 1:
            raise ValueError('inflight must be >= 1')
        if nfield not in (1, 2):
            raise ValueError('nfield must be 1 (auto) or 2 (cross <AAB>)')
        self.eng = eng
        self.tri = np.ascontiguousarray(np.asarray(tri, dtype=np.int32).reshape(-1, 3))
        self.nfield = nfield
        self.scale = float(scale)
        self.norm = norm
        self.nwg = nwg
        self.fmt = L.FMT_BF16X2
        self.parts = (eng.shell_parts(self.tri, nfield) if eng.uses_levels(self.fmt)
                      else None)
        self.streams = [torch.cuda.Stream(device=eng.device) for _ in range(inflight)]
        self._bufs = []
        for s in self.streams:
            with torch.cuda.stream(s):
                self._bufs.append(dict(
                    work=Workspace(eng.device),
                    dk=torch.empty((eng.MY, eng.MX, eng.MZ), dtype=torch.complex64,
                                   device=eng.device),
                    shells=[eng.alloc_shells(self.fmt, eng.nb) for _ in range(nfield)]))
        self._next = 0

    @property
    def inflight(self):
        return len(self.streams)

    def submit(self, *fields):
        """Enqueue one grid (nfield FP32 N^3 fields: A, or A and B) on the
        next stream; returns its float64 device tensor of T sums, ready once
        that stream has run it (wait on the stream or call synchronize())."""
        if len(fields) != self.nfield:
            raise ValueError('%d field(s) per grid' % self.nfield)
        i = self._next
        self._next = (i + 1) % len(self.streams)
        s, b = self.streams[i], self._bufs[i]
        eng = self.eng
        # the fields may have just been written on the caller's stream: order
        # this grid after that work, and keep the fields' memory from being
        # reused while this stream still reads them
        s.wait_stream(torch.cuda.current_stream(eng.device))
        for f in fields:
            f.record_stream(s)
        with torch.cuda.stream(s):
            for f, sh in zip(fields, b['shells']):
                dk = eng.forward(f, stream=s, work=b['work'], out=b['dk'])
                eng.shells(dk, fmt=self.fmt, out=sh, parts=self.parts, stream=s,
                           work=b['work'])
            out = torch.empty(len(self.tri), dtype=torch.float64, device=eng.device)
            eng.triples(b['shells'], self.tri, path=PATH_MFMA, scale=self.scale, out=out,
                        stream=s, nwg=self.nwg, work=b['work'])
            if self.norm is not None:
                out = out / self.norm
        return out

    def synchronize(self):
        for s in self.streams:
            s.synchronize()

    def tune(self, *fields, candidates=(256, 512, 768, 1024), grids=100, warmup=10):
        """Time `grids` grids in flight per workgroup count of the pair kernel
        and keep the fastest: fewer workgroups leave CUs to the other streams'
        FFT passes, so the best count here differs from the one tuned for a
        grid alone.  Returns {count: grids/s}."""
        rates = {}
        for nwg in candidates:
            self.nwg = int(nwg)
            for _ in range(warmup):
                self.submit(*fields)
            self.synchronize()
            t0 = time.perf_counter()
            for _ in range(grids):
                self.submit(*fields)
            self.synchronize()
            rates[int(nwg)] = grids / (time.perf_counter() - t0)
        self.nwg = max(rates, key=rates.get)
        return rates
