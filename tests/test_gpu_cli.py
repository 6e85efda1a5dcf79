"""Command-line measurement: interruption and resume (GPU, subprocesses)."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, timeout=240):
    env = dict(os.environ, PYTHONPATH=ROOT + os.pathsep + os.environ.get('PYTHONPATH', ''))
    print('[cli]', ' '.join(args), flush=True)
    return subprocess.run([sys.executable, '-m', 'bskit_amd.measure'] + args, cwd=ROOT, env=env,
                          timeout=timeout, capture_output=True, text=True)


def test_cli_interrupt_and_resume_matches_full_run(tmp_path):
    rng = np.random.default_rng(3)
    f = rng.standard_normal((32, 32, 32)).astype(np.float32)
    grid = tmp_path / 'grid.npy'
    np.save(grid, f)
    kf = 2 * np.pi / 500.0
    common = ['npy_grid', str(grid), '32', '32', '500.', '0', '100000']
    kb = ['--k_bin_info', repr(kf), repr(kf + 8 * kf + kf / 2), repr(kf), '--block', '20']
    full = tmp_path / 'full.dat'
    r = _run(common + [str(full)] + kb)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = full.read_text().splitlines()
    assert len(rows) > 60
    part = tmp_path / 'part.dat'
    r = _run(common + [str(part)] + kb + ['--stop-after', '2'])
    assert r.returncode == 3, r.stdout + r.stderr
    assert len(part.read_text().splitlines()) == 40
    with open(part, 'a') as fh:             # a torn final write
        fh.write('40 1.23')
    r = _run(common + [str(part)] + kb + ['--resume'])
    assert r.returncode == 0, r.stdout + r.stderr
    assert 'Resuming after triangle 39' in r.stdout
    assert part.read_text() == full.read_text()
    # grid info through the same entry point
    gi = tmp_path / 'gi.dat'
    r = _run(common + [str(gi)] + kb + ['--meas_type', 'grid_info'])
    assert r.returncode == 0, r.stdout + r.stderr
    a = np.loadtxt(gi)
    assert a.shape[1] == 11 and np.all(a[:, 10] == np.rint(a[:, 10]))
