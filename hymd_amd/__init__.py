"""MI355X-native HyMD particle-field hot path.

Drop-in modules over ``libhymd_amd.so`` (hand-written gfx950 HIP kernels
behind the C-ABI in include/hymd_amd.h):

* ``hymd_amd.field``          field.py: initialize_pm, update_field,
                               compute_field_force, update_field_force_q,
                               compute_self_energy_q, compute_field_and_kinetic_energy
* ``hymd_amd.hamiltonian``    hamiltonian.py: DefaultNoChi, DefaultWithChi, SquaredPhi
* ``hymd_amd.force_kernels``  the f2py module force.py binds: cbf/caf/cdf(+_d) on the HIP kernels
* ``hymd_amd.force``          force.py: compute_*_forces__fortran(+__double) (bound from
                               force_kernels, as force.py:8-23), dipole_forces_redistribution
* ``hymd_amd.md``             the rRESPA loop of main.py on the device (+ thermostat / COM coupling)
* ``hymd_amd.topology``       force.py: prepare_bonds (connectivity table or TOML topology)
* ``hymd_amd.thermostat``     thermostat.py: csvr_thermostat, cancel_com_momentum
* ``hymd_amd.pressure``       pressure.py: comp_pressure
* ``hymd_amd.barostat``       barostat.py / barostat_scr.py: isotropic, semiisotropic
* ``hymd_amd.synth``          seeded synthetic systems
"""
__version__ = "0.1.0"
