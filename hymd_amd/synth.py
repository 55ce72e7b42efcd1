"""Seeded synthetic systems for the benchmark and parity workloads.

Inputs are generated, not read (no datasets ship): linear chains are random
walks with 0.47 nm steps in isotropic random directions from uniform origins,
wrapped into the box; velocities are Maxwell-Boltzmann at 300 K for the bead
mass (sigma_v = sqrt(R T / m), R = 0.0083144621 kJ/mol/K) with the centre of
mass motion removed.  numpy PCG64 streams, seed = 1000 + workload index.

* ``diblock_melt``   10,000 beads, 10 nm box, 32^3 mesh, 1,000 A5B5 chains
                     (chi_AB = 15, kappa = 0.05, sigma = 0.5, dt = 0.03, inner 5)
* ``lipid_water``    1,048,576 beads, 40 nm box, 128^3 mesh, 4 types H/L/T/W
                     at fractions 0.1/0.1/0.4/0.4: 52,428 twelve-bead chains
                     [H H L L T*8] plus 419,440 free water beads; bonds
                     r0 = 0.47, k = 1250; angles k = 25 with theta0 = 120 deg at
                     an L vertex and 180 deg otherwise; chi 4x4 symmetric
                     U[0,30]; kappa = 0.03, sigma = 0.5, dt = 0.03, inner 10.
Both accept smaller chain / solvent counts for fast parity cases.
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

GAS_CONSTANT = 0.0083144621


@dataclass
class System:
    name: str
    box: np.ndarray
    mesh: tuple
    n_types: int
    type_names: list
    types: np.ndarray          # (N,) int32
    positions: np.ndarray      # (N,3) float64
    velocities: np.ndarray     # (N,3) float64
    bonds: tuple               # (a, b, r0, k)
    angles: tuple              # (a, b, c, t0, k)
    chi: np.ndarray            # (T,T)
    kappa: float
    sigma: float
    time_step: float
    respa_inner: int
    n_steps: int
    mass: float = 72.0
    temperature: float = 300.0
    charges: Optional[np.ndarray] = None
    dielectric: Optional[float] = None
    extra: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.types)

    @property
    def rho0(self):
        return self.n / float(np.prod(self.box))


def random_walk_chains(rng, n_chains, length, box, step=0.47):
    """(n_chains*length, 3) chain positions, wrapped into [0, L)."""
    box = np.asarray(box, dtype=np.float64)
    origin = rng.uniform(0.0, 1.0, (n_chains, 1, 3)) * box
    d = rng.normal(size=(n_chains, length - 1, 3))
    d *= step / np.linalg.norm(d, axis=2, keepdims=True)
    steps = np.concatenate([np.zeros((n_chains, 1, 3)), np.cumsum(d, axis=1)], axis=1)
    return np.mod(origin + steps, box).reshape(-1, 3)


def maxwell_velocities(rng, n, mass, temperature=300.0):
    v = rng.normal(scale=np.sqrt(GAS_CONSTANT * temperature / mass), size=(n, 3))
    return v - v.mean(axis=0)


def chain_terms(n_chains, length, offset=0):
    """Bond and angle index arrays of consecutive linear chains."""
    base = offset + length * np.arange(n_chains)[:, None]
    i = np.arange(length)
    b_a = (base + i[None, :-1]).ravel()
    a_a = (base + i[None, :-2]).ravel()
    return (b_a, b_a + 1), (a_a, a_a + 1, a_a + 2)


def diblock_melt(n_chains=1000, box=10.0, mesh=32, seed=1000):
    rng = np.random.default_rng(seed)
    L = 10
    box = np.full(3, float(box))
    x = random_walk_chains(rng, n_chains, L, box)
    types = np.tile(np.array([0] * 5 + [1] * 5, dtype=np.int32), n_chains)
    (ba, bb), (aa, ab, ac) = chain_terms(n_chains, L)
    bonds = (ba, bb, np.full(len(ba), 0.47), np.full(len(ba), 1250.0))
    angles = (aa, ab, ac, np.full(len(aa), np.pi), np.full(len(aa), 25.0))
    chi = np.array([[0.0, 15.0], [15.0, 0.0]])
    v = maxwell_velocities(rng, len(types), 72.0)
    return System("diblock_melt", box, (mesh,) * 3, 2, ["A", "B"], types, x, v, bonds, angles, chi, 0.05, 0.5,
                  0.03, 5, 100)


def lipid_water(n_chains=52428, n_water=419440, box=40.0, mesh=128, seed=1001):
    rng = np.random.default_rng(seed)
    L = 12
    box = np.full(3, float(box))
    chain_types = np.array([0, 0, 1, 1] + [2] * 8, dtype=np.int32)
    xc = random_walk_chains(rng, n_chains, L, box)
    xw = rng.uniform(0.0, 1.0, (n_water, 3)) * box
    x = np.concatenate([xc, xw])
    types = np.concatenate([np.tile(chain_types, n_chains), np.full(n_water, 3, dtype=np.int32)])
    (ba, bb), (aa, ab, ac) = chain_terms(n_chains, L)
    bonds = (ba, bb, np.full(len(ba), 0.47), np.full(len(ba), 1250.0))
    t0 = np.where(types[ab] == 1, np.radians(120.0), np.pi)
    angles = (aa, ab, ac, t0, np.full(len(aa), 25.0))
    c = rng.uniform(0.0, 30.0, (4, 4))
    chi = np.triu(c, 1)
    chi = chi + chi.T
    v = maxwell_velocities(rng, len(types), 72.0)
    return System("lipid_water", box, (mesh,) * 3, 4, ["H", "L", "T", "W"], types, x, v, bonds, angles, chi, 0.03,
                  0.5, 0.03, 10, 1000)


def hamiltonian_form(system):
    """(A, b, c0) of DefaultWithChi for ``system`` (rho0 = a = N/V)."""
    T = system.n_types
    rho0 = system.rho0
    inv = 1.0 / (system.kappa * rho0)
    chi = np.array(system.chi, dtype=np.float64)
    np.fill_diagonal(chi, 0.0)
    A = np.full((T, T), inv) + chi / rho0
    return A, np.full(T, -rho0 * inv), 0.5 * rho0 * rho0 * inv
