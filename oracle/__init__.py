"""CPU oracle for the FFT bispectrum path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product (``bskit_amd``) imports, links or runs anything from
this package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Contents
--------
``bins``       restatement of the reference's bin-edge and triangle-list
               generators (``bskit/main.py:408-428``, ``:1008-1389``).
``estimator``  FP64 numpy restatement of the FFT estimator: forward r2c with
               1/N^3 normalisation (``bskit/main.py:1608-1621``), closed-shell
               masked unnormalised c2r (``:227-329``, ``:1846-1861``), triple
               sums, N_tri and k-means (``:1871-1882``, ``:2024-2064``), and the
               output-record conventions (``:1885-1929``, ``:2098-2106``).
``direct``     ctypes wrapper of ``direct.c``: exact periodic enumeration of
               closed triangles (integer N_tri, FP64 k sums and FP64 sums of
               delta(k1) delta(k2) delta(k3)); the periodic form of
               ``bk_binned`` (``bskit/main.py:834-1004``).
``baseline``   reference-structured multi-process CPU path that ``bench.py``
               times as its CPU baseline.

Pinning (what checks each function, see ``tests/test_oracle_*.py``):

* bins: the reference fixtures' edge and index columns
  (``examples/tests/output_ref/Lbox1000_512_kf_3kf_3lowkbins.dat`` and the
  equilateral cross files) and hand-checkable closed forms.
* shells / N_tri / k-means: the grid-info fixtures at 512^3 (L=1000) and
  256^3 (L=500) through ``direct`` (exact counts) and through ``estimator``
  (FFT counts), plus agreement of the two on small grids.
* B: brute-force periodic sums from ``direct`` on small grids, a
  single-cosine closed form, and the invariants B(a*delta) = a^3 B and
  L^6 scaling.
"""
