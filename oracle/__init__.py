"""CPU oracle for the FTLE hot path (test infrastructure, not product code).

This package restates pyFTLE's advection -> Jacobian -> Cauchy-Green -> FTLE
path on the CPU in FP64 so that the HIP kernels in ``pyftle_amd`` can be
checked against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
the timed CPU baseline; the product path never routes through it.

Contents
--------
``numpy_oracle``  NumPy FP64 restatement (interpolation, integrators, particle
                  layout, Jacobian, Cauchy-Green, eigenvalues, FTLE, window
                  driver in reference and time-linear semantics).
``flows``         closed-form velocity fields (double gyre, ABC, saddle).
``cpu_path.c``    C restatement of the reference CPU path (scalar grid
                  interpolation + full-array RK4), the ``port`` CPU baseline.
``Makefile``      builds ``cpu_path.so`` and, when the reference sources are
                  present, ``_ref/ginterp*.so`` from the reference's own
                  ``GridInterp.cpp``.
``native``        ctypes loaders for the two shared objects above.

Pinning status of every oracle function is listed in the module docstrings;
no function here is "parity unpinned" (see ``tests/test_oracle.py``).
"""
