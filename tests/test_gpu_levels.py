"""Sublattice level layout and the level-split MFMA triple sums (GPU).

The sum over a sublattice H of the S^3 shell mesh equals |H|/S^3 of the
full sum when no nonzero vector of H's dual lattice is a frequency of the
triple product.  Levels (include/bskhip.h): 0 the whole mesh, 1 the cells
with x+y+z even, 2 the cells with x, y, z of equal parity, 3 the all-even
cells; each is a contiguous range of a shell row in the level layout.
"""
import itertools
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def _engine(N, box, k_low, k_high):
    from bskit_amd.engine import GridEngine
    return GridEngine(N, box, k_low, k_high)


def _bins(nb, width, box):
    kf = 2 * np.pi / box
    lo = kf * (1 + width * np.arange(nb))
    return lo, lo + width * kf


def _field(N, seed, dtype):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((N, N, N))
    return (f + 0.3 * (f * f - 1)).astype(dtype)


def _layout_positions(S, nc, ne):
    """Row position of every cell (C order) in the level layout, from its
    definition: parts P0 = C0 minus E at 0, P1 = E at nc, P2 = O at nc + ne,
    P3 = C1 minus O at nc + 2 ne; within a part ((c S/2 + x/2) S/2 + y/2) S/2
    + z/2 with the row class c = 0 in E and O, 2 (x%2) + y%2 - 1 in P0 and
    2 (x%2) + y%2 in P3."""
    h = S // 2
    pos = np.empty((S, S, S), np.int64)
    part = np.empty((S, S, S), np.int64)
    for x in range(S):
        for y in range(S):
            for z in range(S):
                xy = 2 * (x % 2) + y % 2
                if (x + y + z) % 2 == 0:
                    p, base, c = (1, nc, 0) if x % 2 == y % 2 == z % 2 == 0 else (0, 0, xy - 1)
                else:
                    p, base, c = (2, nc + ne, 0) if x % 2 == y % 2 == z % 2 == 1 else (
                        3, nc + 2 * ne, xy)
                pos[x, y, z] = base + ((c * h + x // 2) * h + y // 2) * h + z // 2
                part[x, y, z] = p
    return pos.reshape(-1), part.reshape(-1)


@pytest.mark.parametrize('N,nb,width,fmt', [(64, 8, 2, 'bf16x2'), (128, 12, 2, 'f32'),
                                            (64, 6, 3, 'f64'), (256, 4, 16, 'bf16x2')])
def test_level_layout_holds_each_part(N, nb, width, fmt):
    """The z pass's level layout: each part holds exactly its cells, bit for
    bit the values of the C-order shells, in the documented order, and
    zeros in its padding whatever the buffer held before."""
    from bskit_amd import _lib as L
    box = 800.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    S, H = eng.S, eng.S // 2
    lay = eng.layout
    code = {'bf16x2': L.FMT_BF16X2, 'f32': L.FMT_F32, 'f64': L.FMT_F64}[fmt]
    dt = np.float64 if fmt == 'f64' else np.float32
    tdt = torch.float64 if fmt == 'f64' else torch.float32
    dk = eng.forward(torch.from_numpy(_field(N, 5, dt)).cuda())
    ref = eng.shells(dk, fmt=code, dtype=tdt, out=eng.alloc_shells(code, nb, levels=False))
    out = eng.alloc_shells(code, nb, levels=True)
    assert out.layout == lay and out.stride == lay.stride
    nc, ne = lay.off[2], lay.ncell[3]
    assert lay.off == (0, 0, nc, nc) and lay.ncell == (2 * nc + 2 * ne, nc + ne, 2 * ne, ne)
    assert nc == -(-3 * H ** 3 // 128) * 128 and ne == -(-H ** 3 // 128) * 128
    planes = [(out.data, ref.data)] + ([(out.lo, ref.lo)] if out.lo is not None else [])
    for p, _ in planes:
        p.fill_(7)
    eng.shells(dk, fmt=code, dtype=tdt, out=out)
    pos, part = _layout_positions(S, nc, ne)
    # a permutation of the cells: every position taken once, the rest padding
    assert len(np.unique(pos)) == S ** 3
    assert np.array_equal(np.bincount(part), [3 * H ** 3, H ** 3, H ** 3, 3 * H ** 3])
    pad = np.ones(lay.stride, bool)
    pad[pos] = False
    pos_d = torch.from_numpy(pos).cuda()
    pad_d = torch.from_numpy(pad).cuda()
    for p, r in planes:
        assert torch.equal(p[:, pos_d], r[:, :S ** 3])
        assert bool((p[:, pad_d] == 0).all())
    assert torch.equal(out.as_float(), ref.as_float())
    print('level layout N=%d S=%d nb=%d %s: stride %d (S^3 = %d)' % (N, S, nb, fmt, lay.stride,
                                                                    S ** 3))


def _exact_py(S, L, rinf, r2):
    """The level rule restated: no nonzero dual vector (S/2)(a, b, c) of the
    level's parity class with every component <= rinf and norm <= r2."""
    h = S // 2
    A = rinf // h
    for a, b, c in itertools.product(range(-A, A + 1), repeat=3):
        if (a, b, c) == (0, 0, 0):
            continue
        cls = [a % 2 == 0 and b % 2 == 0 and c % 2 == 0, a % 2 == b % 2 == c % 2,
               (a + b + c) % 2 == 0, True][L]
        if cls and h * h * (a * a + b * b + c * c) <= r2 * r2 * (1 + 1e-12):
            return False
    return True


@pytest.mark.parametrize('N,nb,width', [(64, 10, 1), (128, 12, 2), (512, 32, 2)])
def test_triangle_levels_follow_the_dual_lattice_rule(N, nb, width):
    from bskit_amd import bins
    box = 1000.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1)
    lv = eng.triangle_levels(tri)
    m = np.array([math.isqrt(int(v)) if v >= 0 else 0 for v in eng.n2hi])
    r = np.sqrt(np.maximum(eng.n2hi, 0).astype(float))
    ref = []
    for t in tri:
        rinf, r2 = int(m[t].sum()), float(r[t].sum())
        ref.append(next(L for L in (3, 2, 1, 0) if L == 0 or _exact_py(eng.S, L, rinf, r2)))
    assert np.array_equal(lv, np.array(ref))
    counts = np.bincount(lv, minlength=4)
    print('levels N=%d S=%d: %s of %d triangles at levels 0..3' % (N, eng.S, counts.tolist(),
                                                                  len(tri)))
    if N == 512:
        assert (counts > 0).all()


def test_level_sums_are_exact_in_fp64():
    """FP64 shells in the level layout: for every triangle 2^L x (sum over the
    cells of its level L) equals the whole-mesh sum to rounding, and for each
    level above 0 some triangle not exact there shows aliased terms (the
    rule is not vacuous)."""
    from bskit_amd import _lib as L
    from bskit_amd import bins
    N, nb, box = 128, 12, 1000.0
    lo, hi = _bins(nb, 2, box)
    eng = _engine(N, box, lo, hi)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1)
    lv = eng.triangle_levels(tri)
    counts = np.bincount(lv, minlength=4)
    assert (counts > 0).all(), counts
    dk = eng.forward(torch.from_numpy(_field(N, 9, np.float64)).cuda())
    sh = eng.shells(dk, fmt=L.FMT_F64, dtype=torch.float64,
                    out=eng.alloc_shells(L.FMT_F64, nb, levels=True))
    a = sh.level_values(0)
    full = torch.stack([(a[i] * a[j] * a[k]).sum() for i, j, k in tri]).cpu().numpy()
    absum = torch.stack([(a[i].abs() * a[j].abs() * a[k].abs()).sum()
                         for i, j, k in tri]).cpu().numpy()
    for L_ in range(1, 4):
        d = sh.level_values(L_)
        part = 2.0 ** L_ * torch.stack([(d[i] * d[j] * d[k]).sum()
                                        for i, j, k in tri]).cpu().numpy()
        err = np.abs(part - full) / absum
        # the dual of level 3 contains the duals of levels 1 and 2, so a
        # triangle exact at 3 is exact at every level
        exact = (lv == L_) | (lv == 3)
        below = lv == 0 if L_ == 1 else lv < L_
        print('fp64 level %d: exact max %.2e, below-level max %.2e'
              % (L_, err[exact].max(), err[below].max()))
        assert err[exact].max() < 1e-12
        assert err[below].max() > 1e-6


@pytest.mark.parametrize('N,nb,width,mode', [(128, 12, 2, 0), (512, 32, 2, 0), (128, 12, 2, 1),
                                             (96, 10, 2, 2)])
def test_level_mfma_matches_one_level_and_fp64(N, nb, width, mode):
    """Split-BF16 sums with the level plan against the one-level plan on the
    same fields and against the FP64 plain path."""
    from bskit_amd import _lib as L
    from bskit_amd import bins
    box = 1000.0
    lo, hi = _bins(nb, width, box)
    eng = _engine(N, box, lo, hi)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1)
    if mode:
        rng = np.random.default_rng(3)
        tri = np.concatenate([tri, rng.integers(0, nb, size=(200, 3))]).astype(np.int32)
    sc = 1.0 / eng.ncell
    nf = mode + 1
    fs = [_field(N, 40 + i, np.float64) for i in range(nf)]
    d32 = [eng.forward(torch.from_numpy(f.astype(np.float32)).cuda()) for f in fs]
    lvl = [eng.shells(d, fmt=L.FMT_BF16X2, out=eng.alloc_shells(L.FMT_BF16X2, nb, True))
           for d in d32]
    one = [eng.shells(d, fmt=L.FMT_BF16X2, out=eng.alloc_shells(L.FMT_BF16X2, nb, False))
           for d in d32]
    plan = eng.pair_plan(tri, nb, mode, levels=True)
    assert len(plan.segments) > 1
    got2 = eng.triples(lvl, tri, path='mfma', scale=sc).cpu().numpy()
    got1 = eng.triples(one, tri, path='mfma', scale=sc).cpu().numpy()
    s64 = [eng.shells(eng.forward(torch.from_numpy(f).cuda()), fmt=L.FMT_F64,
                      dtype=torch.float64) for f in fs]
    ref = eng.triples(s64, tri, path='plain', scale=sc).cpu().numpy()
    a = [x.data for x in s64]
    ra, rb, rc = {0: (0, 0, 0), 1: (0, 0, 1), 2: (0, 1, 2)}[mode]
    absum = torch.stack([(a[ra][i].abs() * a[rb][j].abs() * a[rc][k].abs()).sum()
                         for i, j, k in tri]).cpu().numpy() * sc
    e21 = (np.abs(got2 - got1) / absum).max()
    e2r = (np.abs(got2 - ref) / absum).max()
    print('levels N=%d nb=%d mode %d: segments (parts, first, count, launches) %s, '
          'vs one-level %.2e, vs fp64 %.2e' % (N, nb, mode, plan.segments, e21, e2r))
    assert e21 < 5e-5
    assert e2r < 5e-5


@pytest.mark.parametrize('nfield,zvar', [(1, 0), (2, 0), (1, 1), (2, 1)])
def test_part_masked_shells_give_the_same_sums(nfield, zvar):
    """Shells written only in the layout parts the plan reads: the MFMA sums
    are bit-identical to those over fully written shells, the parts left out
    keep what the buffer held, and a plan that reads more is refused.  zvar
    1: the z pass on its 10 x 10 plan (also with the rows gathered for
    E/O-only shells), whose sums match the 4 x 5 x 5 plan's to FP32
    rounding."""
    from bskit_amd import _lib as L
    from bskit_amd import bins
    N, nb, box = 512, 32, 1000.0
    lo, hi = _bins(nb, 2, box)
    eng = _engine(N, box, lo, hi)
    assert eng.S == 200
    eng.set_variant(5, zvar)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), nfield).astype(np.int32)
    parts = eng.shell_parts(tri, nfield)
    assert parts.host.shape == (nb,) and (parts.host & 2).all()  # E: every shell
    assert (parts.host != 0xf).any()  # something is left out
    fs = [_field(N, 60 + i, np.float32) for i in range(nfield)]
    d32 = [eng.forward(torch.from_numpy(f).cuda()) for f in fs]
    full = [eng.shells(d, fmt=L.FMT_BF16X2, out=eng.alloc_shells(L.FMT_BF16X2, nb, True))
            for d in d32]
    masked = []
    for d in d32:
        sh = eng.alloc_shells(L.FMT_BF16X2, nb, True)
        sh.data.fill_(0x1234)
        sh.lo.fill_(0x1234)
        masked.append(eng.shells(d, fmt=L.FMT_BF16X2, out=sh, parts=parts))
    sc = 1.0 / eng.ncell
    a = eng.triples(full, tri, path='mfma', scale=sc).cpu().numpy()
    b = eng.triples(masked, tri, path='mfma', scale=sc).cpu().numpy()
    assert np.array_equal(a, b)
    if zvar:
        eng.set_variant(5, 0)
        ref = [eng.shells(d, fmt=L.FMT_BF16X2, out=eng.alloc_shells(L.FMT_BF16X2, nb, True),
                          parts=parts) for d in d32]
        c = eng.triples(ref, tri, path='mfma', scale=sc).cpu().numpy()
        eng.set_variant(5, zvar)
        dv = np.abs(b - c).max() / np.abs(c).max()
        print('z plans 10x10 vs 4x5x5: max |dB| / max |B| = %.2e' % dv)
        assert dv < 1e-5
    # parts not written keep the fill; written parts equal the full shells
    lay = eng.layout
    pc = lay.part_cells()
    starts = np.cumsum((0,) + pc[:3])
    for s_full, s_mask in zip(full, masked):
        for sh_i in range(nb):
            for p in range(4):
                seg = slice(starts[p], starts[p] + pc[p])
                if parts.host[sh_i] >> p & 1:
                    assert torch.equal(s_mask.data[sh_i, seg], s_full.data[sh_i, seg])
                else:
                    assert bool((s_mask.data[sh_i, seg] == 0x1234).all())
    # shells written for a smaller list are refused for this one
    small = eng.shell_parts(np.array([[0, 0, 1]], np.int32), nfield)
    assert small.host[0] == 2 and small.host[2:].max() == 0  # E of shells 0, 1 only
    thin = [eng.shells(d, fmt=L.FMT_BF16X2, out=eng.alloc_shells(L.FMT_BF16X2, nb, True),
                       parts=small) for d in d32]
    with pytest.raises(ValueError, match='another triangle list'):
        eng.triples(thin, tri, path='mfma', scale=sc)
    print('part masks nfield=%d: %s' % (nfield, np.unique(parts.host).tolist()))


def test_grids_in_flight_on_streams_equal_one_at_a_time():
    """Grids issued on three HIP streams in turn, each stream with its own
    shells, output and workspace (the throughput run of bench.py), give the
    sums of the same grids run one at a time, bit for bit, over two rounds
    that reuse every stream's buffers."""
    from bskit_amd import _lib as L
    from bskit_amd import bins
    from bskit_amd.engine import Workspace
    N, nb, box = 128, 12, 1000.0
    lo, hi = _bins(nb, 2, box)
    eng = _engine(N, box, lo, hi)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1).astype(np.int32)
    parts = eng.shell_parts(tri)
    fields = [torch.from_numpy(_field(N, 80 + i, np.float32)).cuda() for i in range(5)]
    sc = 1.0 / eng.ncell

    def grid(f, stream=None, work=None, sh=None, out=None):
        dk = eng.forward(f, stream=stream, work=work)
        sh = eng.shells(dk, fmt=L.FMT_BF16X2, out=sh, parts=parts, stream=stream, work=work)
        return eng.triples([sh], tri, scale=sc, out=out, stream=stream, work=work)

    ref = [grid(f).cpu().numpy() for f in fields]
    ns = 3
    streams = [torch.cuda.Stream() for _ in range(ns)]
    ctx = []
    for s in streams:
        with torch.cuda.stream(s):
            ctx.append(dict(work=Workspace(eng.device),
                            sh=eng.alloc_shells(L.FMT_BF16X2, eng.nb),
                            out=torch.empty(len(tri), dtype=torch.float64, device='cuda')))
    torch.cuda.synchronize()
    for _ in range(2):
        got = []
        for i, f in enumerate(fields):
            s, c = streams[i % ns], ctx[i % ns]
            with torch.cuda.stream(s):
                o = grid(f, stream=s, work=c['work'], sh=c['sh'], out=c['out'])
                got.append(o.clone())  # on s, before the stream's next grid reuses out
        torch.cuda.synchronize()
        for i, g in enumerate(got):
            assert np.array_equal(g.cpu().numpy(), ref[i]), i


def test_grid_pipeline_equals_one_at_a_time():
    """bskit_amd.pipeline.GridPipeline (three grids in flight, normalised by
    a device tensor) returns every grid's sums bit for bit as the engine's
    one-at-a-time path divided by the same tensor; tune() keeps one of its
    candidate workgroup counts, with which the sums again equal the
    one-at-a-time path run with that count."""
    from bskit_amd import _lib as L
    from bskit_amd import bins
    from bskit_amd.pipeline import GridPipeline
    N, nb, box = 128, 12, 1000.0
    lo, hi = _bins(nb, 2, box)
    eng = _engine(N, box, lo, hi)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 1).astype(np.int32)
    fields = [torch.from_numpy(_field(N, 90 + i, np.float32)).cuda() for i in range(4)]
    sc = 1.0 / eng.ncell
    norm = torch.arange(1, len(tri) + 1, dtype=torch.float64, device='cuda')

    def one(f, nwg=None):
        sh = eng.shells(eng.forward(f), fmt=L.FMT_BF16X2,
                        out=eng.alloc_shells(L.FMT_BF16X2, eng.nb))
        return (eng.triples([sh], tri, scale=sc, nwg=nwg) / norm).cpu().numpy()

    ref = [one(f) for f in fields]
    pipe = GridPipeline(eng, tri, inflight=3, scale=sc, norm=norm)
    for _ in range(2):
        outs = [pipe.submit(f) for f in fields]
        pipe.synchronize()
        for i, o in enumerate(outs):
            assert np.array_equal(o.cpu().numpy(), ref[i]), i
    rates = pipe.tune(fields[0], candidates=(32, 64), grids=3, warmup=1)
    assert set(rates) == {32, 64} and pipe.nwg in rates
    out = pipe.submit(fields[1])
    pipe.synchronize()
    assert np.array_equal(out.cpu().numpy(), one(fields[1], nwg=pipe.nwg))
    with pytest.raises(ValueError):
        GridPipeline(eng, tri, inflight=0)


def test_grid_pipeline_cross_equals_one_at_a_time():
    """GridPipeline with nfield=2 (<AAB>: each grid is two fields) returns the
    sums of the engine's one-at-a-time cross path bit for bit, and refuses a
    grid given with the wrong number of fields."""
    from bskit_amd import _lib as L
    from bskit_amd import bins
    from bskit_amd.pipeline import GridPipeline
    N, nb, box = 128, 10, 1000.0
    lo, hi = _bins(nb, 2, box)
    eng = _engine(N, box, lo, hi)
    tri = bins.closed_triangle_indices(np.stack([lo, hi], 1), 2).astype(np.int32)
    pairs = [(torch.from_numpy(_field(N, 100 + 2 * i, np.float32)).cuda(),
              torch.from_numpy(_field(N, 101 + 2 * i, np.float32)).cuda()) for i in range(4)]
    sc = 1.0 / eng.ncell

    def one(a, b):
        sets = [eng.shells(eng.forward(f), fmt=L.FMT_BF16X2,
                           out=eng.alloc_shells(L.FMT_BF16X2, eng.nb)) for f in (a, b)]
        return eng.triples(sets, tri, scale=sc).cpu().numpy()

    ref = [one(a, b) for a, b in pairs]
    pipe = GridPipeline(eng, tri, inflight=3, scale=sc, nfield=2)
    for _ in range(2):
        outs = [pipe.submit(a, b) for a, b in pairs]
        pipe.synchronize()
        for i, o in enumerate(outs):
            assert np.array_equal(o.cpu().numpy(), ref[i]), i
    with pytest.raises(ValueError):
        pipe.submit(pairs[0][0])
    with pytest.raises(ValueError):
        GridPipeline(eng, tri, nfield=3)


def _deep_lists(case):
    """Triangle lists whose level-0 plan (forced) is one segment of few
    items with 3-item units (third sides in three 16-shell blocks) over a
    32-shell window (one prefetch segment per thread): replicated wave lists
    of 3 or 4 items."""
    if case == 3:
        return [(7, 9, k) for k in range(5, 37)]
    fam = [(i, j, k) for i in range(20, 28) for j in range(30, 34) for k in range(5, 37, 2)]
    return fam + [(6, 8, k) for k in range(8, 16)]


@pytest.mark.parametrize('ni', [3, 4])
def test_replicated_lists_with_two_staged_chunks(monkeypatch, ni):
    """Segments with replicated 3- and 4-item wave lists run the pair kernel
    with two staged chunks in flight per workgroup; over the whole mesh
    (every triangle at level 0) their sums match the FP64 plain path."""
    from bskit_amd import _lib as L
    N, nb, box = 128, 40, 1000.0
    lo, hi = _bins(nb, 1, box)
    eng = _engine(N, box, lo, hi)
    assert eng.layout is not None
    tri = np.array(_deep_lists(ni), np.int32)
    monkeypatch.setattr(eng, 'triangle_levels', lambda t: np.zeros(len(t), np.int32))
    plan = eng.pair_plan(tri, nb, 0, levels=True)
    parts, first, count, npass, items, rep = plan.segments[0]
    assert len(plan.segments) == 1 and (parts, npass, items, rep) == (15, 1, ni, 4)
    assert count <= 32  # one prefetch segment per thread: the two-chunk path
    sc = 1.0 / eng.ncell
    f = _field(N, 61, np.float64)
    d32 = eng.forward(torch.from_numpy(f.astype(np.float32)).cuda())
    lvl = eng.shells(d32, fmt=L.FMT_BF16X2, out=eng.alloc_shells(L.FMT_BF16X2, nb, True))
    got = eng.triples([lvl], tri, path='mfma', scale=sc).cpu().numpy()
    s64 = eng.shells(eng.forward(torch.from_numpy(f).cuda()), fmt=L.FMT_F64, dtype=torch.float64)
    ref = eng.triples([s64], tri, path='plain', scale=sc).cpu().numpy()
    a = s64.data
    absum = torch.stack([(a[i].abs() * a[j].abs() * a[k].abs()).sum()
                         for i, j, k in tri]).cpu().numpy() * sc
    err = (np.abs(got - ref) / absum).max()
    print('replicated %d-item lists: segments %s, vs fp64 %.2e' % (ni, plan.segments, err))
    assert err < 5e-5
