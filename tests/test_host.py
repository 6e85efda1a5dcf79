"""Host-side logic (CPU): bins, output formats, combiner, library ABI surface."""
import ctypes
import itertools
import os
import re
import subprocess

import numpy as np
import pytest

import bskit_amd
from bskit_amd import bins as pb
from bskit_amd import io as bio
from oracle import bins as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_product_bins_match_oracle_random():
    rng = np.random.default_rng(11)
    n = 0
    for _ in range(150):
        kmin = rng.uniform(0.001, 0.05)
        dk = rng.uniform(0.002, 0.03)
        nb = int(rng.integers(1, 12))
        kmax = kmin + dk * (nb + rng.uniform(0.2, 1.8))
        nl = int(rng.integers(0, 4))
        dkh = dk * rng.uniform(1, 3) if nl else -1.
        e1 = ob.bin_edges(kmin, kmax, dk, nl, dkh)
        assert np.array_equal(e1, pb.generate_bin_edge_list(kmin, kmax, dk, nl, dkh))
        if not len(e1):
            continue
        n += 1
        for nf in (1, 2, 3):
            a = ob.all_triangles(kmin, kmax, dk, nf, nl, dkh, return_indices=True)
            b = pb.generate_triangle_bin_list(kmin, kmax, dk, num_fields=nf, num_lowk_bins=nl,
                                              dk_high=dkh, return_indices=True)
            assert np.array_equal(a, b)
            assert np.array_equal(ob.all_triangles(kmin, kmax, dk, nf, nl, dkh),
                                  pb.generate_triangle_bin_list(kmin, kmax, dk, num_fields=nf,
                                                                num_lowk_bins=nl, dk_high=dkh))
        for m in (1.0, 2.0, 3.0):
            assert np.array_equal(ob.isosceles(kmin, kmax, dk, m, 0.1, nl, dkh, True),
                                  pb.generate_isosceles_triangle_bin_list(kmin, kmax, dk, m, 0.1,
                                                                          nl, dkh, True))
    assert n > 100


def test_fixture_triangle_list(golden_dir):
    ref = np.loadtxt(os.path.join(golden_dir, 'Lbox1000_512_kf_3kf_3lowkbins.dat'))
    e = pb.generate_triangle_bin_list(0.00314, 0.1, 0.00628, num_lowk_bins=3, dk_high=0.01884)
    assert np.abs(e - ref[:, 4:10]).max() < 5e-10


def test_config_triangle_counts():
    kf = 2 * np.pi / 1000.
    for nb, w, T, T2 in [(8, 2, 98, 231), (16, 3, 564, 1469), (32, 2, 3624, 10041)]:
        kmax = kf + nb * w * kf + w * kf / 2
        assert len(pb.generate_bin_edge_list(kf, kmax, w * kf)) == nb
        assert len(pb.generate_triangle_bin_list(kf, kmax, w * kf)) == T
        assert len(pb.generate_triangle_bin_list(kf, kmax, w * kf, num_fields=2)) == T2


def test_output_formats_match_fixture_lines(golden_dir):
    ref_lines = open(os.path.join(golden_dir, 'Lbox1000_512_kf_3kf_3lowkbins.dat')).read().splitlines()
    ref = np.loadtxt(os.path.join(golden_dir, 'Lbox1000_512_kf_3kf_3lowkbins.dat'))
    for r, line in zip(ref[:5], ref_lines[:5]):
        s = bio.format_row(bio.FAST_GRIDINFO, int(r[0]), r[4:10], r[1:4], ntri=r[10])
        assert s.rstrip('\n') == line
    ub = open(os.path.join(golden_dir, 'test_grid_512_1_unnormbs_kf_3kf_3lowbins.dat')).read().splitlines()
    arr = np.loadtxt(os.path.join(golden_dir, 'test_grid_512_1_unnormbs_kf_3kf_3lowbins.dat'))
    for r, line in zip(arr[:5], ub[:5]):
        assert bio.format_row(bio.FAST_B, int(r[0]), r[1:7], B=r[7]).rstrip('\n') == line
    sl = open(os.path.join(golden_dir, 'test_grid_512_1_bs_kf_3kf_3lowbins_slow.dat')).read().splitlines()
    arr = np.loadtxt(os.path.join(golden_dir, 'test_grid_512_1_bs_kf_3kf_3lowbins_slow.dat'))
    for r, line in zip(arr[:5], sl[:5]):
        s = bio.format_row(bio.SLOW_FULL, int(r[0]), r[4:10], r[1:4], B=r[10], ntri=r[11])
        assert s.rstrip('\n') == line


def test_combiner_reproduces_reference_relation(tmp_path, golden_dir):
    """combine(grid info, unnormalised B) matches the reference's combined file."""
    gi = os.path.join(golden_dir, 'Lbox1000_512_kf_3kf_3lowkbins.dat')
    bu = os.path.join(golden_dir, 'test_grid_512_1_unnormbs_kf_3kf_3lowbins.dat')
    out = tmp_path / 'comb.dat'
    comb = bio.combine(gi, bu, 0.1, str(out))
    ref = np.loadtxt(os.path.join(golden_dir, 'test_grid_512_1_bs_comb_kf_3kf_3lowbins.dat'))
    assert comb.shape == ref.shape
    assert np.array_equal(comb[:, 0], ref[:, 0])
    assert np.allclose(comb[:, 1:], ref[:, 1:], rtol=1e-6)
    header = open(out).readline()
    ref_header = open(os.path.join(golden_dir, 'test_grid_512_1_bs_comb_kf_3kf_3lowbins.dat')).readline()
    assert header == ref_header


def test_combiner_checks(tmp_path):
    gi = tmp_path / 'gi.dat'
    bu = tmp_path / 'b.dat'
    gi.write_text('0 1 1 1 0.1 0.2 0.1 0.2 0.1 0.2 5\n1 1 1 1 0.2 0.3 0.1 0.2 0.1 0.2 5\n')
    bu.write_text('0 0.1 0.2 0.1 0.2 0.1 0.2 10\n2 0.2 0.3 0.1 0.2 0.1 0.2 10\n')
    with pytest.raises(ValueError, match='index mismatch'):
        bio.combine(str(gi), str(bu), 1.0, str(tmp_path / 'o.dat'))
    bu.write_text('0 0.1 0.2 0.1 0.2 0.1 0.2 10\n1 0.25 0.3 0.1 0.2 0.1 0.2 10\n')
    with pytest.raises(ValueError, match='bound mismatch'):
        bio.combine(str(gi), str(bu), 1.0, str(tmp_path / 'o.dat'))


def test_append_rows_chunks_and_resume(tmp_path):
    p = tmp_path / 'b.dat'
    idx = np.arange(250)
    edges = np.tile(np.array([1, 2, 1, 2, 1, 2.]), (250, 1))
    bio.append_rows(str(p), bio.FAST_B, idx[:130], edges[:130], B=np.ones(130))
    assert bio.resume_index(str(p)) == 130
    bio.append_rows(str(p), bio.FAST_B, idx[130:], edges[130:], B=np.ones(120))
    a = np.loadtxt(p)
    assert np.array_equal(a[:, 0], idx)
    # a partial trailing line (interrupted write) is ignored by resume_index
    with open(p, 'a') as f:
        f.write('250 1.0')
    assert bio.resume_index(str(p)) == 250


def test_save_bispectrum(tmp_path):
    b = {'index': np.arange(2), 'k_edge': np.ones((2, 6)), 'k_mean': np.ones((2, 3)),
         'B': np.array([1., 2.]), 'N_tri': np.array([3., 4.])}
    p = tmp_path / 's.dat'
    bio.save_bispectrum(str(p), b)
    a = np.loadtxt(p)
    assert a.shape == (2, 11)
    assert open(p).readline().startswith('# k1_mean')


def _declared_symbols():
    hdr = open(os.path.join(ROOT, 'include', 'bskhip.h')).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char\*)\s+(bsk_\w+)\(', hdr, re.M)))


def test_library_exports_every_declared_symbol():
    lib_path = os.path.join(ROOT, 'bskit_amd', 'libbskhip.so')
    if not os.path.exists(lib_path):
        pytest.skip('libbskhip.so not built')
    out = subprocess.check_output(['nm', '-D', '--defined-only', lib_path]).decode()
    exported = set(re.findall(r' T (bsk_\w+)', out))
    declared = _declared_symbols()
    assert declared, 'no declarations parsed'
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    # the ctypes table covers the same set
    from bskit_amd import _lib
    assert sorted(_lib.exported_symbols()) == declared


def test_library_loads_and_host_calls_work():
    """Host-only ABI calls (no GPU work): thresholds, slots, workspace sizes."""
    from bskit_amd import _lib as L
    try:
        lib = L.lib()
    except (ImportError, OSError) as e:
        pytest.skip('library not loadable here: %s' % e)
    assert lib.bsk_version() == L.ABI_VERSION
    lo = (ctypes.c_int * 2)()
    hi = (ctypes.c_int * 2)()
    kf = 2 * np.pi / 100.
    L.call('bsk_shell_thresholds', 64, 100.0, 2, L.dbl_array([kf, 3 * kf]),
           L.dbl_array([3 * kf, 5 * kf]), lo, hi)
    assert list(lo) == [1, 9] and list(hi) == [9, 25]
    n = ctypes.c_size_t()
    from bskit_amd.engine import PairPlan
    plan = PairPlan(np.array([[2, 1, 0], [0, 1, 2], [1, 1, 1]]), 3, L.MODE_AUTO, upload=False)
    sl = plan.slots()
    assert sl.shape == (3, 4) and (sl[:, 1:] == -1).all()  # one segment
    assert np.array_equal(sl[0], sl[1])  # symmetric spectra: any side order maps to one slot
    L.call('bsk_triples_mfma_workspace', plan.handle, 10, ctypes.byref(n))
    # FP32 partial tiles per workgroup, then the reduction's FP64 sums over 16
    # workgroup ranges per slot
    assert n.value == 10 * max(plan.nitem, 1) * 256 * 4 + 16 * max(plan.nitem, 1) * 256 * 8
    for mode, nsh in ((L.MODE_AUTO, 129), (L.MODE_ABC, 97)):
        with pytest.raises(L.BskError):
            PairPlan(np.array([[0, 0, 0]]), nsh, mode, upload=False)
    PairPlan(np.array([[127, 64, 63]]), 128, L.MODE_AUTO, upload=False)
    with pytest.raises(L.BskError):
        PairPlan(np.array([[3, 0, 0]]), 3, L.MODE_AUTO, upload=False)


# parts of the level layout (P0 = C0 minus E, P1 = E, P2 = O, P3 = C1 minus O)
# in each sublattice level
_LEVEL_PARTS = {0: (0, 1, 2, 3), 1: (0, 1), 2: (1, 2), 3: (1,)}


def _emulate_pair_plan(plan, sets, tri, mode, nshell, part_sets=None):
    """FP64 numpy emulation of the MFMA kernel's tables: per launch item the
    16x16 tile sum_x (d_a d_b)(x) d_k(x) over the cells of its segment, stored
    at its partial position (each position belongs to one segment), read
    back through the triangle slots (one per segment, summed, scaled by
    2^level).  part_sets[p]:
    the role fields on the cells of layout part p; without it the plan must be
    one segment of all parts, summed over `sets`."""
    items, outpos, prs = plan.tables()
    # the kernel re-forms pair products only at flagged items: an unflagged
    # item continues the previous item's tile in the same wave's list
    # (flag 29: padding)
    e = 0
    for _, _, _, npass, ni, rep in plan.segments:
        for w in range(8 * npass):
            lst = items[e:e + ni]
            e += ni
            assert lst[0] < 0
            for s in range(1, ni):
                if (lst[s] >> 29) & 1:
                    assert all((x >> 29) & 1 for x in lst[s:])  # padding closes the list
                    break
                if lst[s] >= 0:
                    assert (lst[s] & 0xffff) == (lst[s - 1] & 0xffff)
    assert e == len(items)
    assert plan.segments[0][0] & 2  # E first
    part = np.zeros((max(plan.nitem, 1), 16, 16))
    written = np.zeros(len(part), bool)
    launch, e0 = 0, 0
    for g, (parts, w0, wn, npass, ni, rep) in enumerate(plan.segments):
        assert rep in (1, 2, 4) and (rep == 1 or npass == 1)
        assert 0 <= w0 and w0 + wn <= nshell
        if part_sets is None:
            assert parts == 0xf and len(plan.segments) == 1
            f0, f1, f2 = sets
        else:
            f0, f1, f2 = (np.concatenate([part_sets[q][r] for q in range(4) if parts >> q & 1],
                                         axis=1) for r in range(3))
        z = np.zeros_like(f0[0])
        for _ in range(npass):
            # replicas: waves w and w + 8/rep hold the same list; one counts
            lists = items[e0:e0 + 8 * ni].reshape(8, ni)
            assert all(np.array_equal(lists[w], lists[w % (8 // rep)]) for w in range(8))
            for e in range(e0, e0 + (8 // rep) * ni):
                d, pos = items[e], outpos[e]
                if (d >> 29) & 1:
                    assert pos == -1
                    continue  # padding
                tile = d & 0xffff
                kb = ((d >> 16) & 0xff) + w0 // 16
                tsum = np.zeros((16, 16))
                for row in range(16):
                    pr = prs[tile * 16 + row]
                    a, b = pr & 0xffff, (pr >> 16) & 0xffff
                    for v in (a, b):
                        assert v == 0xffff or w0 <= v < w0 + wn  # staged in this segment
                    pa = (f0[a] if a != 0xffff else z) * (f1[b] if b != 0xffff else z)
                    for col in range(16):
                        k = 16 * kb + col
                        if w0 <= k < w0 + wn:  # other third-side rows are zero
                            tsum[row, col] = np.sum(pa * f2[k])
                assert not written[pos]
                part[pos], written[pos] = tsum, True
            launch += 1
            e0 += 8 * ni
    assert launch == plan.npass
    assert written.all() or plan.nitem == 0
    flat = part.reshape(-1)
    slots = plan.slots()
    assert slots.shape == (len(tri), 4) and (slots[:, 0] >= 0).all()
    assert (slots[:, len(plan.segments):] == -1).all()
    level = (slots[:, 0] >> 28) & 3
    assert all(((s >> 28) & 3) == level[t] for t in range(len(tri)) for s in slots[t] if s >= 0)
    tot = np.zeros(len(tri))
    for g in range(4):
        ok = slots[:, g] >= 0
        tot[ok] += flat[slots[ok, g] & 0x0fffffff]
    return tot * 2.0 ** level


def _level_reference(part_sets, tri, levels):
    return np.array([2.0 ** lv * sum(np.sum(part_sets[p][0][i] * part_sets[p][1][j] *
                                            part_sets[p][2][k]) for p in _LEVEL_PARTS[lv])
                     for (i, j, k), lv in zip(tri, levels)])


@pytest.mark.parametrize('mode,nshell,ntri,seed', [(0, 7, 40, 1), (0, 33, 300, 2), (1, 20, 150, 3),
                                                   (2, 18, 200, 4), (0, 40, 900, 5)])
def test_pair_plan_tables_compute_every_triangle(mode, nshell, ntri, seed):
    """The host plan's tiles, items and slots give every triangle its sum."""
    from bskit_amd import _lib as L
    from bskit_amd.engine import PairPlan
    try:
        L.lib()
    except (ImportError, OSError) as e:
        pytest.skip('library not loadable here: %s' % e)
    rng = np.random.default_rng(seed)
    ncell = 24
    nset = {0: 1, 1: 2, 2: 3}[mode]
    fields = [rng.standard_normal((nshell, ncell)) for _ in range(nset)]
    roles = {0: (0, 0, 0), 1: (0, 0, 1), 2: (0, 1, 2)}[mode]
    tri = rng.integers(0, nshell, size=(ntri, 3))
    plan = PairPlan(tri, nshell, mode, upload=False)
    assert plan.npair <= ntri and plan.items_per_wave <= 8
    sets = [fields[r] for r in roles]
    got = _emulate_pair_plan(plan, sets, tri, mode, nshell)
    ref = np.array([np.sum(sets[0][i] * sets[1][j] * sets[2][k]) for i, j, k in tri])
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize('mode,nshell,ntri,seed,probs', [
    (0, 33, 400, 11, (0.1, 0.3, 0.4, 0.2)), (0, 20, 150, 12, (0, 0, 0, 1)),
    (1, 20, 150, 13, (0.25, 0.25, 0.25, 0.25)), (2, 18, 200, 14, (1, 0, 0, 0)),
    (0, 40, 600, 15, (0, 0.5, 0, 0.5)), (0, 24, 300, 16, (0.5, 0, 0.5, 0))])
def test_level_pair_plan(mode, nshell, ntri, seed, probs):
    """A plan over sublattice levels: each triangle's item is summed over
    the layout parts of its level (their own fields here) through the
    segment launches and scaled by 2^level."""
    from bskit_amd import _lib as L
    from bskit_amd.engine import PairPlan
    try:
        L.lib()
    except (ImportError, OSError) as e:
        pytest.skip('library not loadable here: %s' % e)
    rng = np.random.default_rng(seed)
    nset = {0: 1, 1: 2, 2: 3}[mode]
    roles = {0: (0, 0, 0), 1: (0, 0, 1), 2: (0, 1, 2)}[mode]
    part_sets = []
    for p in range(4):
        fields = [rng.standard_normal((nshell, (12, 4, 4, 12)[p])) for _ in range(nset)]
        part_sets.append([fields[r] for r in roles])
    tri = rng.integers(0, nshell, size=(ntri, 3))
    levels = rng.choice(4, size=ntri, p=probs).astype(np.int32)
    plan = PairPlan(tri, nshell, mode, upload=False, levels=levels)
    # segments: maximal runs of parts with the same levels present
    present = set(levels.tolist())
    covered = 0
    for parts, w0, wn, npass, ni, rep in plan.segments:
        assert ni <= 4
        lv = [{L_ for L_ in present if p in _LEVEL_PARTS[L_]} for p in range(4) if parts >> p & 1]
        assert all(x == lv[0] for x in lv) and npass > 0
        assert not covered & parts
        covered |= parts
    got = _emulate_pair_plan(plan, None, tri, mode, nshell, part_sets)
    ref = _level_reference(part_sets, tri, levels)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)
    # all triangles at level 0: one segment of all parts, the items of
    # bsk_pairs_create (which never replicates its wave lists: the C-order and
    # fused entry points run it)
    p0 = PairPlan(tri, nshell, mode, upload=False, levels=np.zeros(ntri, np.int32))
    p1 = PairPlan(tri, nshell, mode, upload=False)
    assert len(p1.segments) == 1 and p1.segments[0][0] == 0xf and p1.segments[0][5] == 1
    assert p0.segments[0][:4] == p1.segments[0][:4] and np.array_equal(p0.slots(), p1.slots())


@pytest.mark.parametrize('mode', [0, 1, 2])
def test_level_pair_plan_stages_shell_windows(mode):
    """Each segment stages only the shell range its levels' triangles use
    (windows not aligned to 16: third-side rows offset in LDS) and the
    emulated sums still match."""
    from bskit_amd import _lib as L
    from bskit_amd.engine import PairPlan
    try:
        L.lib()
    except (ImportError, OSError) as e:
        pytest.skip('library not loadable here: %s' % e)
    rng = np.random.default_rng(20 + mode)
    nshell = 40
    spans = [(25, 40), (10, 30), (0, 18), (3, 9)]  # shells used per level
    tri, levels = [], []
    for lv, (a, b) in enumerate(spans):
        t = rng.integers(a, b, size=(60, 3))
        t[0] = (a, b - 1, a)  # both ends of the span
        tri.append(t)
        levels += [lv] * len(t)
    tri = np.concatenate(tri)
    levels = np.array(levels, np.int32)
    nset = {0: 1, 1: 2, 2: 3}[mode]
    roles = {0: (0, 0, 0), 1: (0, 0, 1), 2: (0, 1, 2)}[mode]
    part_sets = []
    for p in range(4):
        fields = [rng.standard_normal((nshell, 8)) for _ in range(nset)]
        part_sets.append([fields[r] for r in roles])
    plan = PairPlan(tri, nshell, mode, upload=False, levels=levels)
    for parts, w0, wn, npass, ni, rep in plan.segments:
        lv = [L_ for L_ in range(4) if any(parts >> p & 1 for p in _LEVEL_PARTS[L_])]
        assert (w0, w0 + wn) == (min(spans[L_][0] for L_ in lv), max(spans[L_][1] for L_ in lv))
    got = _emulate_pair_plan(plan, None, tri, mode, nshell, part_sets)
    ref = _level_reference(part_sets, tri, levels)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_two_level_pair_plan_rejects_bad_levels():
    from bskit_amd import _lib as L
    from bskit_amd.engine import PairPlan
    try:
        L.lib()
    except (ImportError, OSError) as e:
        pytest.skip('library not loadable here: %s' % e)
    tri = np.array([[0, 1, 2], [1, 1, 1]])
    with pytest.raises(L.BskError, match='level'):
        PairPlan(tri, 4, 0, upload=False, levels=np.array([0, 4], np.int32))
    with pytest.raises(L.BskError, match='level'):
        PairPlan(tri, 4, 0, upload=False, levels=np.array([-1, 0], np.int32))


@pytest.mark.parametrize('mode,nshell', [(2, 70), (1, 100)])
def test_pair_plan_splits_wide_tiles_over_waves(mode, nshell):
    """Where the item count per wave is capped below a tile's kb blocks
    (four prefetch segments per role set), the tile's items are split over
    waves, each part forming the pair products again."""
    from bskit_amd import _lib as L
    from bskit_amd.engine import PairPlan
    try:
        L.lib()
    except (ImportError, OSError) as e:
        pytest.skip('library not loadable here: %s' % e)
    tri = np.array([(i, j, k) for i in range(5) for j in range(4) for k in range(0, nshell, 3)])
    plan = PairPlan(tri, nshell, mode, upload=False)
    items, _, _ = plan.tables()
    ni = plan.items_per_wave
    waves = {}
    for g, d in enumerate(items):
        if (d & 0xffff) != plan.ntile:
            waves.setdefault(d & 0xffff, set()).add(g // ni)
    assert any(len(w) > 1 for w in waves.values())
    rng = np.random.default_rng(8)
    nset = {1: 2, 2: 3}[mode]
    fields = [rng.standard_normal((nshell, 16)) for _ in range(nset)]
    roles = {1: (0, 0, 1), 2: (0, 1, 2)}[mode]
    sets = [fields[r] for r in roles]
    got = _emulate_pair_plan(plan, sets, tri, mode, nshell)
    ref = np.array([np.sum(sets[0][i] * sets[1][j] * sets[2][k]) for i, j, k in tri])
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_shell_threshold_abi_matches_oracle():
    from bskit_amd import _lib as L
    from oracle import estimator as est
    try:
        L.lib()
    except (ImportError, OSError) as e:
        pytest.skip('library not loadable here: %s' % e)
    from bskit_amd.engine import shell_limits
    rng = np.random.default_rng(3)
    kf = 2 * np.pi / 1000.
    lows = rng.uniform(0, 40, 200) * kf
    highs = lows + rng.uniform(0, 10, 200) * kf
    # exact ties at odd multiples of k_f
    lows[:20] = kf * (1 + 2 * np.arange(20))
    highs[:20] = lows[:20] + 2 * kf
    a = shell_limits(512, 1000., lows, highs)
    b = est.shell_limits(lows, highs, kf)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_api_surface_matches_reference_names():
    names = ['FFTBispectrum', 'k_field', 'number_field', 'compute_Nbin', 'compute_k_means_on_grid',
             'compute_bk_FFT_value', 'bk_FFT_full', 'bk_FFT_grid_info', 'bk_FFT_unnormalized_value',
             'bk_binned', 'generate_bin_edge_list', 'generate_equilateral_triangle_bin_list',
             'generate_squeezed_triangle_bin_list', 'generate_isosceles_triangle_bin_list',
             'generate_triangle_bin_list', 'geometric_k_mean']
    for n in names:
        assert hasattr(bskit_amd, n), n
    import inspect
    sig = inspect.signature(bskit_amd.FFTBispectrum.__init__)
    ref_params = ['source', 'Nmesh', 'BoxSize', 'dk', 'kmin', 'kmax', 'num_lowk_bins', 'dk_high',
                  'dmu', 'mu_min', 'mu_max', 'pos_units_mpcoverh', 'k_edges', 'second', 'third',
                  'triangle_type', 'isos_mult', 'isos_tol', 'squeezed_bin_index',
                  'for_grid_info_only']
    got = [p for p in sig.parameters if p != 'self'][:len(ref_params)]
    assert got == ref_params
    for meth in ('measure_bispectrum', 'measure_bispectrum_faster', 'measure_gridinfo_faster'):
        params = list(inspect.signature(getattr(bskit_amd.FFTBispectrum, meth)).parameters)
        assert params[1:7] == ['imin', 'imax', 'kmeas_min', 'kmeas_max', 'out_file', 'verbose']


def test_constructor_errors_without_gpu():
    field = np.zeros((16, 16, 16), dtype=np.float32)
    with pytest.raises(ValueError, match='Must specify either'):
        bskit_amd.FFTBispectrum(field, BoxSize=100., for_grid_info_only=True)
    with pytest.raises(ValueError, match='Need second source'):
        bskit_amd.FFTBispectrum(field, BoxSize=100., kmin=0.1, kmax=1., third=field,
                                for_grid_info_only=True)
    with pytest.raises(ValueError, match='BoxSize is required'):
        bskit_amd.FFTBispectrum(field, kmin=0.1, kmax=1., for_grid_info_only=True)
    with pytest.raises(NotImplementedError):
        bskit_amd.FFTBispectrum(np.zeros((16, 16, 8), np.float32), BoxSize=100., kmin=0.1,
                                kmax=1., for_grid_info_only=True)


def test_truncate_partial_and_cli_errors(tmp_path):
    from bskit_amd import measure
    p = tmp_path / 'rows.dat'
    p.write_text('0 1.0\n1 2.0\n2 3.0e+0')           # interrupted mid-line
    assert bio.truncate_partial(str(p)) == len('0 1.0\n1 2.0\n')
    assert p.read_text() == '0 1.0\n1 2.0\n'
    assert bio.resume_index(str(p)) == 2
    assert bio.truncate_partial(str(tmp_path / 'missing.dat')) == 0
    g = tmp_path / 'g.npy'
    np.save(g, np.zeros((16, 16, 16), dtype=np.float32))
    base = ['npy_grid', str(g), '16', '16', '100.', '0', '10', str(tmp_path / 'o.dat')]
    with pytest.raises(RuntimeError, match='k-bin file or kmin/kmax/dk'):
        measure.main(base)
    with pytest.raises(NotImplementedError):
        measure.main(['gadget'] + base[1:] + ['--k_bin_info', '0.1', '0.5', '0.1'])
    with pytest.raises(ValueError, match='isos_mult'):
        measure.main(base + ['--k_bin_info', '0.1', '0.5', '0.1', '--triangle_type', 'isosceles'])


def test_subbox_index_roundtrip_and_cuts():
    from bskit_amd import subbox
    for n in (1, 2, 3, 4):
        for i in range(n ** 3):
            m = subbox.subbox_index_to_multiindex(i, n)
            assert m.dtype == np.float64
            assert subbox.subbox_multiindex_to_index(m, n) == i
    f = np.arange(8 ** 3, dtype=np.float64).reshape(8, 8, 8)
    sm = subbox.field_subbox_pm(np.array([1., 0., 1.]), 2, f, BoxSize=100.)
    assert np.array_equal(sm.array, f[4:8, 0:4, 4:8])
    assert np.allclose(sm.attrs['BoxSize'], 50.) and list(sm.attrs['Nmesh']) == [4, 4, 4]
    sl = subbox.field_subbox_slice(f, [(50., 100.), (0., 50.), (50., 100.)], BoxSize=100.)
    assert np.array_equal(sl.array, sm.array)


@pytest.mark.parametrize('S', [8, 12])
def test_sublattice_duals_by_brute_force(S):
    """The dual lattices behind the sublattice levels, by brute force on a
    small mesh: the frequencies q (mod S) at which sum_{x in H} e^{2 pi i q.x/S}
    does not vanish are exactly (S/2)(a, b, c) with a, b, c all even (level 0,
    the whole mesh), all of one parity (1: x+y+z even), of even sum (2: x, y,
    z of equal parity) or any (3: all even)."""
    i = np.arange(S)
    x, y, z = np.meshgrid(i, i, i, indexing='ij')
    cells = {0: np.ones_like(x, bool), 1: (x + y + z) % 2 == 0,
             2: (x % 2 == y % 2) & (y % 2 == z % 2),
             3: (x % 2 == 0) & (y % 2 == 0) & (z % 2 == 0)}
    h = S // 2
    for L, mask in cells.items():
        assert mask.sum() * 2 ** L == S ** 3
        # the sum over H of e^{2 pi i q.x / S} for every q: an FFT of the mask
        f = np.fft.fftn(mask.astype(float))
        dual = {tuple(q) for q in np.argwhere(np.abs(f) > 1e-6 * mask.sum())}
        rule = set()
        for a, b, c in itertools.product(range(2), repeat=3):
            cls = [a % 2 == 0 and b % 2 == 0 and c % 2 == 0, a % 2 == b % 2 == c % 2,
                   (a + b + c) % 2 == 0, True][L]
            if cls:
                rule.add(((h * a) % S, (h * b) % S, (h * c) % S))
        assert dual == rule, (L, sorted(dual), sorted(rule))


def test_server_rejects_unknown_stream_assignment_and_depth():
    """The serving loop's stream choice is 'work' (least outstanding
    footprint) or 'count' (fewest jobs in flight); anything else is refused
    before the device is touched."""
    from bskit_amd import serve
    with pytest.raises(ValueError):
        serve.Server(assign='random')
    with pytest.raises(ValueError):
        serve.Server(depth=0)
