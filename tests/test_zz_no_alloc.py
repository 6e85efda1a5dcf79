"""bsk_triples_mfma_levels_noalloc (library scratch, no allocation) against
bsk_triples_mfma_levels (caller workspace), through the C ABI: the same
bits on two calls in a row, the second with a smaller input, and a clear
refusal above the scratch size."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

from bskit_amd import _lib as L  # noqa: E402
from bskit_amd import bins as B  # noqa: E402
from bskit_amd import synth  # noqa: E402
from bskit_amd.engine import GridEngine  # noqa: E402

SCRATCH = 64 << 20  # BSK_NOALLOC_SCRATCH_BYTES


def _case(N, nb, seed):
    box = 1000.0
    kf = 2 * np.pi / box
    kLU = B.generate_bin_edge_list(kf, kf + nb * 2 * kf + kf, 2 * kf)
    eng = GridEngine(N, box, kLU[:, 0], kLU[:, 1])
    tri = B.closed_triangle_indices(kLU, 1).astype(np.int32)
    field = torch.from_numpy(synth.quadratic_field(N, box, seed=seed)).cuda()
    sh = eng.shells(eng.forward(field), fmt=L.FMT_BF16X2)
    plan = eng.pair_plan(tri, eng.nb, L.MODE_AUTO, levels=sh.layout is not None)
    lay = sh.layout
    off = (ctypes.c_longlong * L.LEVELS)(*lay.off)
    cells = (ctypes.c_longlong * L.LEVELS)(*lay.ncell)
    args = (plan.handle, L.ptr(sh.data), L.ptr(sh.lo), L.ptr(sh.data), L.ptr(sh.lo),
            L.ptr(sh.data), L.ptr(sh.lo), sh.stride, off, cells)
    # the raw pointers in args are only valid while the shell tensors live
    return eng, plan, tri, args, sh


def _both(eng, plan, tri, args, nwg):
    scale = 1.0 / eng.ncell
    n = L.size_out()
    L.call('bsk_triples_mfma_workspace', plan.handle, nwg, ctypes.byref(n))
    assert n.value <= SCRATCH
    work = torch.empty(n.value, dtype=torch.uint8, device='cuda')
    ref = torch.empty(len(tri), dtype=torch.float64, device='cuda')
    got = torch.full((len(tri),), float('nan'), dtype=torch.float64, device='cuda')
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.call('bsk_triples_mfma_levels', *args, nwg, scale, L.ptr(ref), L.ptr(work), s)
    L.call('bsk_triples_mfma_levels_noalloc', *args, nwg, scale, L.ptr(got), s)
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_equals_workspace_entry_point():
    big = _case(128, 24, seed=3)
    small = _case(64, 10, seed=4)
    for (eng, plan, tri, args, _), nwg in ((big, 512), (small, 128), (big, 256)):
        ref, got = _both(eng, plan, tri, args, nwg)
        print('N=%d T=%d nwg=%d: max |noalloc - workspace| = %g'
              % (eng.N, len(tri), nwg, np.abs(got - ref).max()))
        assert np.array_equal(got, ref)
    # a plan and workgroup count needing more than the scratch is refused
    eng, plan, tri, args, _ = big
    n = L.size_out()
    nwg = 1
    while True:
        L.call('bsk_triples_mfma_workspace', plan.handle, nwg, ctypes.byref(n))
        if n.value > SCRATCH:
            break
        nwg *= 2
    out = torch.zeros(len(tri), dtype=torch.float64, device='cuda')
    with pytest.raises(L.BskError, match='scratch'):
        L.call('bsk_triples_mfma_levels_noalloc', *args, nwg, 1.0, L.ptr(out),
               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
