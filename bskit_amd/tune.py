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


def _best_of(cands, reps, rounds=3):
    """Per candidate (name -> callable) the least mean time over `rounds`
    interleaved rounds of `reps` calls: a spike or drift during one round
    does not decide between candidates a few percent apart."""
    out = {}
    for _ in range(rounds):
        for name, fn in cands.items():
            t = _time(fn, reps)
            out[name] = min(out.get(name, t), t)
    return out


def tune(eng, tri=None, field=None, reps=4, force=False, log=None, nfield=1):
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

    def with_variants(first, combo, fn):
        def run():
            for p, v in enumerate(combo):
                eng.set_variant(first + p, v)
            return fn()
        return run

    # forward passes 0..2
    combos = list(itertools.product((0, 1), repeat=3))
    t = _best_of({c: with_variants(0, c, lambda: eng.forward(field)) for c in combos}, reps)
    best = min(combos, key=lambda c: t[c])
    timings.update({'forward%s' % ''.join(map(str, c)): t[c] for c in combos})
    variants[0:3] = best
    for p, v in enumerate(best):
        eng.set_variant(p, v)
    dk = eng.forward(field)
    shells = eng.alloc_shells(L.FMT_BF16X2, eng.nb)
    # the shells as the pipeline writes them: with a level plan only the
    # layout parts some launch reads
    parts = (eng.shell_parts(tri, nfield)
             if tri is not None and ntri and eng.uses_levels(L.FMT_BF16X2) else None)
    t = _best_of({c: with_variants(3, c, lambda: eng.shells(dk, fmt=L.FMT_BF16X2, out=shells,
                                                             parts=parts))
                  for c in combos}, reps)
    best = min(combos, key=lambda c: t[c])
    timings.update({'shells%s' % ''.join(map(str, c)): t[c] for c in combos})
    variants[3:6] = best
    for p, v in enumerate(best):
        eng.set_variant(3 + p, v)
    cfg = {'variants': [int(v) for v in variants], 'timings_ms': timings}
    if tri is not None and ntri and eng.mfma_supported(eng.nb, nfield, tri):
        eng.shells(dk, fmt=L.FMT_BF16X2, out=shells)
        out = torch.empty(ntri, dtype=torch.float64, device=eng.device)
        nchunk = eng.ncell // 128
        cands = sorted({n for n in (256, 512, 768, 1024, 1536, 2048) if n <= max(1, nchunk // 4)})
        cands = cands or [eng.mfma_nwg()]
        sets = [shells] * nfield
        t = _best_of({n: (lambda n=n: eng.triples(sets, tri, path='mfma', out=out, nwg=n))
                      for n in cands}, reps)
        timings.update({'mfma_nwg%d' % n: t[n] for n in cands})
        cfg['mfma_nwg'] = int(min(cands, key=lambda n: t[n]))
        if nfield == 1 and eng.zfused_supported(tri):
            zc = (256, 512, 768, 1024)
            t = _best_of({n: (lambda n=n: eng.bispectrum_zfused(dk, tri, out=out, nwg=n))
                          for n in zc}, reps)
            timings.update({'zfused_nwg%d' % n: t[n] for n in zc})
            cfg['zfused_nwg'] = int(min(zc, key=lambda n: t[n]))
            # whole shells + sums stage, both ways, each at its best count
            t = _best_of({
                'sep': lambda: (eng.shells(dk, fmt=L.FMT_BF16X2, out=shells, parts=parts),
                                eng.triples([shells], tri, path='mfma', out=out,
                                            nwg=cfg['mfma_nwg'])),
                'fus': lambda: eng.bispectrum_zfused(dk, tri, out=out, nwg=cfg['zfused_nwg'])},
                reps)
            sep, fus = t['sep'], t['fus']
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
