"""FP64 CPU restatement of HyMD's particle-field hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker, never the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to compare the HIP path against it or to time it as the
reported CPU baseline.  ``hymd_amd`` never imports it and has no CPU fallback.

Modules
-------
field      CIC paint/readout, r2c/c2r, Gaussian filter, W'(phi), spectral
           gradient, PME (field.py:152-403,428-703; hamiltonian.py:34-486)
gpe        variable-dielectric electrostatics (field.py:706-761, 764-1112)
bonded     ctypes wrapper of ``bonded.c`` (compute_*_forces*.f90 and
           dipole_reconstruction*.f90 restated in C, built by ``Makefile``;
           the reference's Fortran itself is not built: it is reached only
           through f2py, which needs meson, absent here)
md         rRESPA outer step (main.py:802-1169, integrator.py:9-73)
pressure   comp_pressure / comp_laplacian (pressure.py:9-200)
barostat   Berendsen and SCR barostats (barostat.py, barostat_scr.py)
thermostat CSVR thermostat and COM removal (thermostat.py)
topology   prepare_bonds and its networkx traversal orders (force.py:266-728)

Pinning status (what checks each function against something other than a
retyping of its own formula) is listed in each module's header.
"""
