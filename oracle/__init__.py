"""CPU oracle for the spectral-element time-stepping path (TEST INFRASTRUCTURE).

This package is a from-scratch NumPy/SciPy restatement, in FP64, of the
reference's MATLAB engines for the flat-shell and 3D-solid spectral element
solvers (pawelkudela/ma-shm).  Every function cites the reference file:line it
follows.  It exists only to CHECK the HIP implementation:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it, and only as the checker / CPU baseline;
* the product path (``ma-shm_amd/``) never imports, links or runs it.

Pinning status of each part (what independent check guards it) is listed in the
module docstrings of ``common``, ``mesh``, ``shell``, ``solid``, ``pzt``, ``adjoint``
and ``grid``.  Parts that no independent check covers are marked "parity unpinned"
there.
"""
