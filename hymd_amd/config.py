"""Configuration fields used on the particle-field hot path.

Mirrors the subset of ``hymd.input_parser.Config`` (input_parser.py:17-312)
that field.py, hamiltonian.py, force.py and the rRESPA loop read, with the
same names, units and defaults (kJ/mol, nm, ps, amu).
"""
from dataclasses import dataclass, field
from typing import ClassVar, List, Optional

import numpy as np


@dataclass
class Chi:
    """chi mixing energy between two named species (force.py:224-264)."""

    atom_1: str
    atom_2: str
    interaction_energy: float


@dataclass
class Target_pressure:
    """Target pressures in bar, lateral and normal (barostat.py:37-40)."""

    P_L: object = None
    P_N: object = None


@dataclass
class Config:
    gas_constant: ClassVar[float] = 0.0083144621  # kJ mol-1 K-1 (input_parser.py:159)
    coulomb_constant: ClassVar[float] = 138.935458  # kJ nm mol-1 e-2 (input_parser.py:160)

    n_steps: int
    time_step: float
    mesh_size: object
    sigma: float
    kappa: float

    dtype: object = np.float32
    box_size: Optional[np.ndarray] = None
    n_print: Optional[int] = None
    mass: float = 72.0                      # input_parser.py:949-967 default
    hamiltonian: str = "DefaultWithChi"
    respa_inner: int = 1
    chi: List[Chi] = field(default_factory=list)
    n_particles: Optional[int] = None
    n_types: Optional[int] = None
    unique_names: Optional[List[str]] = None
    type_to_name_map: Optional[dict] = None
    coulombtype: Optional[str] = None
    dielectric_const: Optional[float] = None
    self_energy: Optional[float] = None
    type_charges: Optional[list] = None
    # PIC_Spectral_GPE (input_parser.py:195-200, 1256-1262): one dielectric per
    # type (after sort_dielectric_by_type_id), mixing 0.6 and tolerance 1e-6
    # by default, convergence "max_diff" by default (main.py:137-165)
    dielectric_type: Optional[list] = None
    pol_mixing: Optional[float] = None
    conv_crit: Optional[float] = None
    convergence_type: Optional[str] = None
    barostat: Optional[str] = None           # "isotropic" | "semiisotropic"
    rho0: Optional[float] = None
    a: Optional[float] = None
    m: Optional[list] = None                 # per-type paint mass (input_parser.py:1138-1141)
    target_temperature: Optional[float] = None
    tau: Optional[float] = None
    start_temperature: Optional[float] = None
    thermostat_work: float = 0.0             # input_parser.py:189
    thermostat_coupling_groups: List[List[str]] = field(default_factory=list)
    cancel_com_momentum: object = False      # int period or False (input_parser.py:192)
    n_b: Optional[int] = None                # thermostat / barostat period; check_n_b defaults it to 1
    pressure: bool = False                   # input_parser.py:204-211
    barostat_type: Optional[str] = None      # "berendsen" | "scr"
    tau_p: Optional[float] = None
    target_pressure: Optional[Target_pressure] = None

    def __post_init__(self):
        if self.box_size is not None:
            self.box_size = np.asarray(self.box_size, dtype=np.float64)
        if self.n_types is not None and self.m is None:
            self.m = [1.0] * self.n_types

    @property
    def mesh(self):
        """Mesh edges as a 3-tuple (``mesh_size`` may be a scalar or 3 values)."""
        return tuple(int(v) for v in np.broadcast_to(np.asarray(self.mesh_size), (3,)))

    @property
    def simulation_volume(self):
        return float(np.prod(self.box_size))


def chi_matrix(config):
    """Symmetric (T,T) chi matrix from ``config.chi`` (names mapped through
    ``type_to_name_map``, hamiltonian.py:384-387); zero diagonal."""
    T = config.n_types
    chi = np.zeros((T, T))
    if not config.chi:
        return chi
    if isinstance(config.chi, np.ndarray):
        chi = np.array(config.chi, dtype=np.float64)
        np.fill_diagonal(chi, 0.0)
        return chi
    name_of = config.type_to_name_map or {t: n for t, n in enumerate(config.unique_names)}
    table = {tuple(sorted([c.atom_1, c.atom_2])): float(c.interaction_energy) for c in config.chi}
    for i in range(T):
        for j in range(T):
            if i != j:
                chi[i, j] = table.get(tuple(sorted([name_of[i], name_of[j]])), 0.0)
    return chi
