"""The f2py module surface ``force_kernels`` (setup.py:26-38, force.py:8-23) on HIP."""
from .force import (  # noqa: F401
    compute_angle_forces__fortran as caf,
    compute_angle_forces__fortran__double as caf_d,
    compute_bond_forces__fortran as cbf,
    compute_bond_forces__fortran__double as cbf_d,
    compute_dihedral_forces__fortran as cdf,
    compute_dihedral_forces__fortran__double as cdf_d,
)
