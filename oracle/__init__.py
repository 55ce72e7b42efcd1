"""FP64 CPU restatement of HyMD's particle-field hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker, never the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to compare the HIP path against it or to time it as the
reported CPU baseline.  ``hymd_amd`` never imports it and has no CPU fallback.

Modules
-------
field      CIC paint/readout, r2c/c2r, Gaussian filter, W'(phi), spectral
           gradient, PME (field.py:152-403,428-703; hamiltonian.py:34-486)
bonded     ctypes wrapper of ``bonded.c`` (compute_*_forces*.f90 and
           dipole_reconstruction*.f90 restated in C) and of the real
           reference kernels compiled into ``_ref/`` by ``Makefile``
md         rRESPA outer step (main.py:802-1169, integrator.py:9-73)

Pinning status (what checks each function against something other than a
retyping of its own formula) is listed in each module's header.
"""
