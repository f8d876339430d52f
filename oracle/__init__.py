"""CPU oracle for the SPECTER per-RK-substep hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``specter_amd``) imports
this package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline, never as the thing measured or shipped.

``oracle.specter_ref`` is a float64 numpy/scipy restatement of the
reference's algorithm (SPECTER, Fortran).  Every function cites the reference
file:line it follows.  Arrays use C-order numpy shapes that reproduce the
reference's Fortran memory layout byte for byte:

* spectral  Fortran ``(nz, ny, nkx)``  ->  numpy ``(nkx, ny, nz)`` complex128
* real      Fortran ``(nx, ny, nz)``   ->  numpy ``(nz, ny, nx)``  float64

Pinning (what fixes each function independently of this restatement) is
documented at the top of ``specter_ref.py`` and exercised in
``tests/test_oracle_*.py``.
"""
