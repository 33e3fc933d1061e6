"""CPU oracle for the coarse-graining hot path -- TEST INFRASTRUCTURE ONLY.

This package is a plain NumPy (float64 / complex128) restatement of the
reference's HOTRG 2D/3D, ATRG 3D, truncated-SVD, randomized-SVD and tensor
container modules.  Each function cites the reference file:line it follows.

Who may use it: ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` -- and only as the checker or as the
timed CPU baseline.  The product package ``trg_backup_amd`` never imports it;
the product runs on the GPU through ``libtrg.so`` or raises.

Pinning (what checks each function against something other than itself):
  * HOTRG_2d  -- Onsager's closed form (2D Ising), brute-force Z of the 4x4
                 torus in the untruncated regime, beta=0 limit ln 2.
  * HOTRG_3d  -- brute-force Z of the 2x2x2 Ising torus (untruncated),
                 beta=0 limit; projector = LAPACK eigh of the environment.
  * ATRG_3d_new -- brute-force Z of the 2x2x2 Ising and q=3 Potts tori
                 (untruncated regime), beta=0 limits ln 2 / ln 3,
                 rSVD == dense SVD when the sketch covers the rank.
  * truncated_svd -- LAPACK identities (A = u s vh), hand-worked rank cases.
  * randomized_svd -- equals the dense SVD when l >= rank.
  * philox      -- Random123 Philox4x32-10 known-answer vectors.
  * exact       -- Onsager by two independent quadratures; brute force by
                   transfer-matrix cross-check.
Functions with none of the above are marked "parity unpinned" in their
docstring.
"""
