"""Autotuner for the FFT bispectrum path on the device it runs on.

Tunables of one GridEngine key (Nmesh, shell mesh, retained cube, bins):

* the radix-plan variant (0: radix-8 heavy, 1: radix-4 heavy) of each of
  the six FFT passes; the three forward passes and the three inverse passes
  are timed as groups (all 2^3 combinations each);
* the workgroup count of the MFMA triple kernel, and of the fused z c2r +
  triple kernel where it applies;
* whether the fused kernel (z c2r inside the triple sums, no shell fields in
  HBM) or the separate z pass + triple kernel is faster for the workload
  (`zfused`: the engine's choice for bispectra of all its shells).

Results live in a process-wide cache (a dict guarded by a lock) persisted
as JSON at $BSK_TUNE_CACHE (default ~/.cache/bskit_amd/tune.json), keyed by
device name and workload shape (meshes, bins, field roles and a hash of the
triangle list), so a tuned shape is applied to every later engine without
re-timing.  ``tune_class`` shares one tuning among the engines of a mesh
class (Nmesh, shell mesh): the serving loop and the graph batches tune one
representative job per class and apply it to the rest.  Every candidate computes the same quantity (the
FFT variants differ only in rounding), so tuning never changes what is
measured, only how fast.
"""
import itertools
import json
import os
import threading

import numpy as np
import torch

from . import _lib as L

_LOCK = threading.Lock()
_CACHE = None


def cache_path():
    return os.environ.get('BSK_TUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'bskit_amd', 'tune.json'))


def _load():
    global _CACHE
    if _CACHE is None:
        try:
            with open(cache_path()) as f:
                _CACHE = json.load(f)
        except (OSError, ValueError):
            _CACHE = {}
    return _CACHE


def _save():
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + '.tmp%d' % os.getpid()
        with open(tmp, 'w') as f:
            json.dump(_CACHE, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home keeps the in-process cache only


def key_of(eng, ntri=0, nfield=1, tri_hash=''):
    dev = torch.cuda.get_device_name(eng.device)
    lv = int(bool(eng.uses_levels()))  # sublattice-level MFMA sums
    return '%s|N%d|S%d|m%d|nb%d|T%d|L%d|F%d|%s' % (dev, eng.N, eng.S, eng.m, eng.nb, ntri, lv,
                                                   nfield, tri_hash)


def _tri_hash(tri):
    import hashlib
    if tri is None:
        return ''
    t = np.ascontiguousarray(np.asarray(tri, dtype=np.int32))
    return hashlib.sha1(t.tobytes()).hexdigest()[:12]


def lookup(eng, ntri=0, nfield=1, tri=None):
    with _LOCK:
        return _load().get(key_of(eng, ntri, nfield, _tri_hash(tri)))


def apply(eng, cfg):
    """Set the radix variants of a tuned configuration on the engine."""
    for p, v in enumerate(cfg.get('variants', [0] * 6)):
        eng.set_variant(p, v)
    if cfg.get('mfma_nwg'):
        eng.tuned_nwg = int(cfg['mfma_nwg'])
    if cfg.get('zfused_nwg'):
        eng.tuned_zfused_nwg = int(cfg['zfused_nwg'])
    eng.prefer_zfused = bool(cfg.get('zfused', False))


def _time(fn, reps=3):
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / reps


def tune(eng, tri=None, field=None, reps=3, force=False, log=None, nfield=1):
    """Time the candidates on this engine's shapes and store the best.

    field: a float32 device field of the engine's mesh (random if None);
    tri: triangle list for the MFMA workgroup count (skipped if None);
    nfield: 1 (auto <AAA>) or 2 (cross <AAB>: the pair plan of that mode,
    timed with the field's shells on both roles).
    Returns the configuration dict (also applied to the engine).
    """
    ntri = 0 if tri is None else len(tri)
    k = key_of(eng, ntri, nfield, _tri_hash(tri))
    if not force:
        cfg = lookup(eng, ntri, nfield, tri)
        if cfg is not None:
            apply(eng, cfg)
            return cfg
    if field is None:
        g = torch.Generator(device=eng.device)
        g.manual_seed(0)
        field = torch.randn((eng.N,) * 3, generator=g, device=eng.device)
    variants = [0] * 6
    timings = {}
    # forward passes 0..2
    best = None
    for combo in itertools.product((0, 1), repeat=3):
        for p, v in enumerate(combo):
            eng.set_variant(p, v)
        t = _time(lambda: eng.forward(field), reps)
        timings['forward%s' % ''.join(map(str, combo))] = t
        if best is None or t < best[0]:
            best = (t, combo)
    variants[0:3] = best[1]
    for p, v in enumerate(best[1]):
        eng.set_variant(p, v)
    dk = eng.forward(field)
    shells = eng.alloc_shells(L.FMT_BF16X2, eng.nb)
    best = None
    for combo in itertools.product((0, 1), repeat=3):
        for p, v in enumerate(combo):
            eng.set_variant(3 + p, v)
        t = _time(lambda: eng.shells(dk, fmt=L.FMT_BF16X2, out=shells), reps)
        timings['shells%s' % ''.join(map(str, combo))] = t
        if best is None or t < best[0]:
            best = (t, combo)
    variants[3:6] = best[1]
    for p, v in enumerate(best[1]):
        eng.set_variant(3 + p, v)
    cfg = {'variants': [int(v) for v in variants], 'timings_ms': timings}
    if tri is not None and ntri and eng.mfma_supported(eng.nb, nfield, tri):
        eng.shells(dk, fmt=L.FMT_BF16X2, out=shells)
        out = torch.empty(ntri, dtype=torch.float64, device=eng.device)
        nchunk = eng.ncell // 128
        cands = sorted({n for n in (256, 512, 768, 1024, 1536, 2048) if n <= max(1, nchunk // 4)})
        cands = cands or [eng.mfma_nwg()]
        best = None
        sets = [shells] * nfield
        for n in cands:
            t = _time(lambda: eng.triples(sets, tri, path='mfma', out=out, nwg=n), reps)
            timings['mfma_nwg%d' % n] = t
            if best is None or t < best[0]:
                best = (t, n)
        cfg['mfma_nwg'] = int(best[1])
        if nfield == 1 and eng.zfused_supported(tri):
            best = None
            for n in (256, 512, 768, 1024):
                t = _time(lambda: eng.bispectrum_zfused(dk, tri, out=out, nwg=n), reps)
                timings['zfused_nwg%d' % n] = t
                if best is None or t < best[0]:
                    best = (t, n)
            cfg['zfused_nwg'] = int(best[1])
            # whole shells + sums stage, both ways, each at its best count
            sep = _time(lambda: (eng.shells(dk, fmt=L.FMT_BF16X2, out=shells),
                                 eng.triples([shells], tri, path='mfma', out=out,
                                             nwg=cfg['mfma_nwg'])), reps)
            fus = _time(lambda: eng.bispectrum_zfused(dk, tri, out=out, nwg=cfg['zfused_nwg']),
                        reps)
            timings['shells_triples_separate'] = sep
            timings['shells_triples_zfused'] = fus
            cfg['zfused'] = bool(fus < sep)
    if log:
        log('[tune] %s -> %s' % (k, {x: cfg[x] for x in cfg if x != 'timings_ms'}))
    with _LOCK:
        _load()[k] = cfg
        _save()
    apply(eng, cfg)
    return cfg


_CLASS = {}


def tune_class(eng, tri=None, field=None, nfield=1, log=None):
    """Tune the first engine of its (device, Nmesh, shell mesh) class in full
    and apply that configuration to every later engine of the class (the
    radix plans depend on the FFT lengths, the pair kernel's workgroup count
    mostly on the cell count); returns the configuration applied."""
    ck = (torch.cuda.get_device_name(eng.device), eng.N, eng.S)
    with _LOCK:
        cfg = _CLASS.get(ck)
    if cfg is None:
        cfg = tune(eng, tri=tri, field=field, nfield=nfield, log=log)
        with _LOCK:
            _CLASS[ck] = cfg
        return cfg
    apply(eng, cfg)
    return cfg
