"""Batched cross bispectra captured as one HIP graph.

The reference runs one measure_bispectrum_faster-style job per grid pair
(scripts/measure/measure_bs_fast.py:231-260, roles <AAB> of
bskit/main.py:624-640).  ``CrossBatch`` takes a list of pairs (A, B fields
with their own mesh size and bins), preallocates every buffer, and records
the forward transforms, shells and <AAB> triple sums of all pairs into
a hipGraph (torch.cuda.CUDAGraph).  Pairs are spread over ``nbranch``
streams by estimated cost (longest first onto the least loaded stream),
forked from and joined to the capture stream, so the graph holds one
independent chain per branch; ``per_branch=True`` records one graph per
branch stream instead, launched each on its stream.  One ``replay()`` runs
the whole batch.

Graph rules kept here: every buffer (spectra, shells, outputs, workspaces,
triangle lists, MFMA layouts) exists before capture, and nothing in the
captured region allocates, copies from host memory or clears memory with a
memset node; each replay rewrites every output in full.
"""
import numpy as np
import torch

from . import _lib as L
from . import bins as B
from .engine import PATH_MFMA, PATH_PLAIN, GridEngine, Workspace, require_device


class _Pair:
    def __init__(self, eng, a, b, tri, path, box):
        self.eng = eng
        self.fields = (a, b)
        self.tri = tri
        self.path = path
        self.scale = box ** 6 / float(eng.ncell)
        fmt = L.FMT_BF16X2 if path == PATH_MFMA else L.FMT_F32
        self.fmt = fmt
        ctype = torch.complex64
        self.dk = [torch.empty((eng.MY, eng.MX, eng.MZ), dtype=ctype, device=eng.device)
                   for _ in range(2)]
        self.sets = [eng.alloc_shells(fmt, eng.nb) for _ in range(2)]
        self.out = torch.empty(len(tri), dtype=torch.float64, device=eng.device)
        # level layout: write only the parts of each shell the sums read
        self.parts = (eng.shell_parts(tri, 2) if path == PATH_MFMA and
                      self.sets[0].layout is not None else None)
        # relative cost: two forward transforms, two shell sets, the sums
        self.cost = 2.0 * eng.N ** 3 + 2.0 * eng.nb * eng.ncell + len(tri) * eng.ncell / 64.0

    def enqueue(self, work):
        eng = self.eng
        for f, dk, sh in zip(self.fields, self.dk, self.sets):
            eng.forward(f, out=dk, work=work)
            eng.shells(dk, fmt=self.fmt, out=sh, work=work, parts=self.parts)
        eng.triples(self.sets, self.tri, path=self.path, scale=self.scale, out=self.out,
                    work=work)


class CrossBatch:
    """pairs: list of (A, B, box, kLU) with A, B float32 device N^3 fields."""

    def __init__(self, pairs, device=None, nbranch=4, autotune=True, per_branch=False):
        """autotune: tune the radix plans and the pair kernel's workgroup count
        per mesh class (Nmesh, shell mesh) on the first pair of the class
        (bskit_amd.tune.tune_class) before anything is recorded.  per_branch:
        one graph per branch stream instead of one graph of every branch."""
        self.device = require_device(device)
        self.per_branch = bool(per_branch)
        self.tuned = {}
        self.nbranch = max(1, min(int(nbranch), len(pairs)))
        self.engines = {}
        self.pairs = []
        for a, b, box, kLU in pairs:
            n = int(a.shape[0])
            kLU = np.asarray(kLU, dtype=np.float64)
            key = (n, float(box), tuple(map(tuple, kLU.tolist())))
            eng = self.engines.get(key)
            if eng is None:
                eng = GridEngine(n, box, kLU[:, 0], kLU[:, 1], device=self.device)
                self.engines[key] = eng
            tri = B.closed_triangle_indices(kLU, 2).astype(np.int32)
            path = PATH_MFMA if GridEngine.mfma_supported(eng.nb, 2, tri) else PATH_PLAIN
            if autotune and path == PATH_MFMA and len(tri):
                from . import tune
                cfg = tune.tune_class(eng, tri=tri, field=a, nfield=2)
                self.tuned[(eng.N, eng.S)] = {'variants': cfg.get('variants'),
                                              'mfma_nwg': cfg.get('mfma_nwg')}
            self.pairs.append(_Pair(eng, a, b, tri, path, float(box)))
        self.branches = [torch.cuda.Stream(device=self.device) for _ in range(self.nbranch)]
        self.works = [Workspace(self.device) for _ in range(self.nbranch)]
        # longest first onto the least loaded branch; cost ~ the cells the
        # pair's transforms, shells and triangle sums touch
        load = [0.0] * self.nbranch
        self.assign = [[] for _ in range(self.nbranch)]
        for i in sorted(range(len(self.pairs)), key=lambda i: -self.pairs[i].cost):
            b = min(range(self.nbranch), key=lambda k: (load[k], k))
            self.assign[b].append(i)
            load[b] += self.pairs[i].cost
        self.graphs = None

    def _enqueue_branch(self, b):
        for i in self.assign[b]:
            self.pairs[i].enqueue(self.works[b])

    def run_eager(self):
        """Launch the batch without graphs on the branch streams (also sizes
        workspaces and caches)."""
        main = torch.cuda.current_stream(self.device)
        for b, s in enumerate(self.branches):
            s.wait_stream(main)
            with torch.cuda.stream(s):
                self._enqueue_branch(b)
        for s in self.branches:
            main.wait_stream(s)
        return [p.out for p in self.pairs]

    def capture(self):
        """Record the batch (after one eager run): one graph whose branches
        fork from and join the capture stream (``per_branch`` False), or one
        graph per branch stream."""
        self.run_eager()
        torch.cuda.synchronize(self.device)
        graphs = []
        if self.per_branch:
            for b, s in enumerate(self.branches):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=s):
                    self._enqueue_branch(b)
                graphs.append(g)
        else:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.run_eager()
            graphs.append(g)
        torch.cuda.synchronize(self.device)
        self.graphs = graphs
        return graphs

    def replay(self):
        """Replay the batch; the caller's stream waits for all of it."""
        if self.graphs is None:
            self.capture()
        if not self.per_branch:
            self.graphs[0].replay()
            return [p.out for p in self.pairs]
        main = torch.cuda.current_stream(self.device)
        for g, s in zip(self.graphs, self.branches):
            s.wait_stream(main)
            with torch.cuda.stream(s):
                g.replay()
        for s in self.branches:
            main.wait_stream(s)
        return [p.out for p in self.pairs]

    @property
    def graph(self):
        return self.graphs

    @property
    def ntri(self):
        return sum(len(p.tri) for p in self.pairs)


def config_pairs(seed=0, box=1000.0, layout=((128, 32), (256, 24), (384, 8)), nbin_range=(8, 24),
                 device='cuda', npool=2):
    """The batched cross workload: 32 pairs at 128^3, 24 at 256^3, 8 at 384^3,
    nb uniform over nbin_range from a seeded RNG, bins of 2 k_f from k_f.  Each
    mesh size has npool seeded (A, B) field pairs, reused round robin."""
    from . import synth
    from .serve import serving_bins
    rng = np.random.default_rng(seed)
    kf = 2 * np.pi / box
    pairs = []
    for si, (n, count) in enumerate(layout):
        pool = [synth.correlated_pair_device(n, box, 500 + 100 * si + 2 * p,
                                             501 + 100 * si + 2 * p, device)
                for p in range(npool)]
        for c in range(count):
            nb = int(rng.integers(nbin_range[0], nbin_range[1] + 1))
            a, b = pool[c % npool]
            pairs.append((a, b, box, serving_bins(n, nb, kf)))
    return pairs
