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
* ``peptide_solvent`` 2,097,152 beads, 50 nm box, 160^3 mesh (shared by the
                     density and PME), 4 types BB/SC1/SC2/W: 8,738 chains of
                     40 residues x 3 beads (backbone + 2 side beads) = 50 %
                     peptide, the rest free solvent.  Choices the workload
                     leaves open, fixed here: backbone bond angles drawn in
                     [100, 130] deg; bonds BB-BB 0.38 nm, BB-SC1 and SC1-SC2
                     0.30 nm (k = 1250); angles BB-BB-SC1 100 deg and
                     BB-SC1-SC2 180 deg (k = 25); every backbone dihedral is a
                     combined bending-torsion term (type 1) with cosine
                     coefficients n = 1..4 ~ U[-5,5] kJ/mol, K(phi) = 60 kJ/mol
                     and the chain's last dihedral flagged; 10 % of the beads
                     (side beads and solvent) carry +-1 with zero net charge;
                     dielectric 45; chi 4x4 symmetric U[0,30]; kappa 0.03;
                     sigma 0.5; dt 0.03; respa_inner 1.
* ``markov_melt``    16,777,216 beads, 100 nm box, 256^3 mesh, 6 types in
                     1,048,576 sixteen-bead linear chains, each chain's types a
                     Markov sequence (first bead uniform; seeded 6x6 transition
                     matrix U[0,1) + 3 I, rows normalised, so runs of one type
                     are favoured); bonds r0 = 0.47, k = 1250 and angles
                     theta0 = 180 deg, k = 25 (the values of the other chain
                     workloads; the workload leaves them open); chi 6x6
                     symmetric U[0,30]; kappa 0.02; sigma 0.6; 5 % of the beads
                     carry +-1 with zero net charge, PME on the same mesh with
                     dielectric 45 (left open, fixed as in peptide_solvent);
                     dt 0.03, respa_inner 1, 200 steps.
* ``replica_set``    64 independent A5B5 diblock melts for the batched
                     chi-gradient workload: N_r ~ U[50k, 400k] rounded down to
                     whole 10-bead chains, L_r = (N_r / 8.4)^(1/3) nm, mesh
                     edge the one of {48, 64, 96} nearest L_r / 0.6 nm (only
                     48 and 64 occur), chi_AB ~ U[5, 25], sigma ~ U[0.4, 0.7];
                     kappa = 0.05 (left open by the workload, fixed here).
                     Each replica's chains are stored in the mesh order of
                     their first bead (x-major cells, as an MD code keeps its
                     particles; ``spatial=False`` keeps generation order).
All accept smaller chain / solvent counts for fast parity cases.
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


def _bent_backbone(rng, n_chains, n_res, box, step=0.38, angle_range=(100.0, 130.0)):
    """Backbone random walks whose bond angles lie in ``angle_range`` (deg)."""
    x = np.zeros((n_chains, n_res, 3))
    x[:, 0] = rng.uniform(0.0, 1.0, (n_chains, 3)) * box
    d = rng.normal(size=(n_chains, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x[:, 1] = x[:, 0] + step * d
    for i in range(2, n_res):
        gamma = np.radians(rng.uniform(*angle_range, n_chains))
        alpha = np.pi - gamma                       # turn away from the previous direction
        u = rng.normal(size=(n_chains, 3))
        u -= np.sum(u * d, axis=1, keepdims=True) * d
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        d = np.cos(alpha)[:, None] * d + np.sin(alpha)[:, None] * u
        x[:, i] = x[:, i - 1] + step * d
    return x


def peptide_solvent(n_chains=8738, n_res=40, n_solvent=1048592, box=50.0, mesh=160, seed=1002):
    rng = np.random.default_rng(seed)
    box = np.full(3, float(box))
    bb = _bent_backbone(rng, n_chains, n_res, box)
    side_dir = rng.normal(size=(n_chains, n_res, 3))
    side_dir /= np.linalg.norm(side_dir, axis=2, keepdims=True)
    sc1 = bb + 0.30 * side_dir
    sc2 = sc1 + 0.30 * side_dir
    chain = np.stack([bb, sc1, sc2], axis=2).reshape(n_chains, 3 * n_res, 3)
    x = np.concatenate([np.mod(chain.reshape(-1, 3), box), rng.uniform(0.0, 1.0, (n_solvent, 3)) * box])
    per_chain = 3 * n_res
    types = np.concatenate([np.tile(np.array([0, 1, 2], dtype=np.int32), n_chains * n_res),
                            np.full(n_solvent, 3, dtype=np.int32)])
    base = per_chain * np.arange(n_chains)[:, None]
    r = np.arange(n_res)[None, :]
    bbi = base + 3 * r                                  # backbone bead of residue r
    ba = np.concatenate([bbi[:, :-1].ravel(), bbi.ravel(), (bbi + 1).ravel()])
    bb_ = np.concatenate([bbi[:, 1:].ravel(), (bbi + 1).ravel(), (bbi + 2).ravel()])
    r0 = np.concatenate([np.full(n_chains * (n_res - 1), 0.38), np.full(2 * n_chains * n_res, 0.30)])
    bonds = (ba, bb_, r0, np.full(len(ba), 1250.0))
    # angles BB(i-1)-BB(i)-SC1(i) and BB(i)-SC1(i)-SC2(i)
    aa = np.concatenate([bbi[:, :-1].ravel(), bbi.ravel()])
    ab = np.concatenate([bbi[:, 1:].ravel(), (bbi + 1).ravel()])
    ac = np.concatenate([(bbi[:, 1:] + 1).ravel(), (bbi + 2).ravel()])
    t0 = np.concatenate([np.full(n_chains * (n_res - 1), np.radians(100.0)), np.full(n_chains * n_res, np.pi)])
    angles = (aa, ab, ac, t0, np.full(len(aa), 25.0))
    # backbone dihedrals, combined bending-torsion (type 1)
    da = bbi[:, :-3].ravel()
    m4 = len(da)
    coeff = np.zeros((m4, 6, 5), dtype=np.float32)
    coeff[:, 0, 1:] = rng.uniform(-5.0, 5.0, (m4, 4))
    coeff[:, 4, 0] = 30.0
    last = np.zeros((n_chains, n_res - 3), dtype=np.int32)
    last[:, -1] = 1
    dihedrals = (da, da + 3, da + 6, da + 9, coeff, np.ones(m4, dtype=np.int32), last.ravel())
    # +-1 charges on 10 % of the beads, chosen among side beads and solvent
    n = len(types)
    n_q = (n // 10) // 2 * 2
    candidates = np.flatnonzero(types != 0)
    pick = rng.choice(candidates, n_q, replace=False)
    charges = np.zeros(n, dtype=np.float32)
    charges[pick[: n_q // 2]] = 1.0
    charges[pick[n_q // 2:]] = -1.0
    c = rng.uniform(0.0, 30.0, (4, 4))
    chi = np.triu(c, 1)
    chi = chi + chi.T
    v = maxwell_velocities(rng, n, 72.0)
    s = System("peptide_solvent", box, (mesh,) * 3, 4, ["BB", "SC1", "SC2", "W"], types, x, v, bonds, angles, chi,
               0.03, 0.5, 0.03, 1, 500, charges=charges, dielectric=45.0)
    s.extra["dihedrals"] = dihedrals
    return s


def markov_types(rng, n_chains, length, n_types, stay=3.0):
    """(n_chains * length,) int32 bead types: per chain a Markov sequence with
    a uniform first bead and the seeded transition matrix U[0,1) + stay I
    (rows normalised)."""
    P = rng.uniform(0.0, 1.0, (n_types, n_types)) + stay * np.eye(n_types)
    cum = np.cumsum(P / P.sum(axis=1, keepdims=True), axis=1)
    cum[:, -1] = 1.0
    t = np.empty((n_chains, length), dtype=np.int32)
    t[:, 0] = rng.integers(0, n_types, n_chains)
    for j in range(1, length):
        u = rng.uniform(0.0, 1.0, n_chains)
        t[:, j] = np.minimum((u[:, None] >= cum[t[:, j - 1]]).sum(axis=1), n_types - 1)
    return t.ravel()


def markov_melt(n_chains=1048576, length=16, box=100.0, mesh=256, n_types=6, charged=0.05, seed=1004):
    rng = np.random.default_rng(seed)
    box = np.full(3, float(box))
    x = random_walk_chains(rng, n_chains, length, box)
    types = markov_types(rng, n_chains, length, n_types)
    (ba, bb), (aa, ab, ac) = chain_terms(n_chains, length)
    bonds = (ba, bb, np.full(len(ba), 0.47), np.full(len(ba), 1250.0))
    angles = (aa, ab, ac, np.full(len(aa), np.pi), np.full(len(aa), 25.0))
    c = rng.uniform(0.0, 30.0, (n_types, n_types))
    chi = np.triu(c, 1)
    chi = chi + chi.T
    n = len(types)
    n_q = int(n * charged) // 2 * 2
    pick = rng.choice(n, n_q, replace=False)
    charges = np.zeros(n, dtype=np.float32)
    charges[pick[: n_q // 2]] = 1.0
    charges[pick[n_q // 2:]] = -1.0
    v = maxwell_velocities(rng, n, 72.0)
    names = [chr(ord("A") + t) for t in range(n_types)]
    return System("markov_melt", box, (mesh,) * 3, n_types, names, types, x, v, bonds, angles, chi, 0.02, 0.6, 0.03,
                  1, 200, charges=charges if n_q else None, dielectric=45.0)


@dataclass
class ReplicaSet:
    counts: np.ndarray      # (R,) particles per replica
    meshes: np.ndarray      # (R,3) int
    boxes: np.ndarray       # (R,3) nm
    positions: np.ndarray   # (sum N_r, 3) float64, replica after replica
    types: np.ndarray       # (sum N_r,) int32 (0 = A, 1 = B)
    chi: np.ndarray         # (R,2,2) symmetric, zero diagonal
    sigma: np.ndarray       # (R,)
    kappa: np.ndarray       # (R,)

    @property
    def n_replicas(self):
        return len(self.counts)

    def replica(self, r):
        """(positions, types) of replica r."""
        o = np.concatenate([[0], np.cumsum(self.counts)])
        return self.positions[o[r]:o[r + 1]], self.types[o[r]:o[r + 1]]


def mesh_ordered_chains(x, length, box, mesh):
    """Chains of ``length`` beads (consecutive rows of x) reordered by the
    x-major mesh cell of their first bead (stable)."""
    c = x.reshape(-1, length, 3)
    cell = np.floor(c[:, 0] / np.asarray(box) * mesh).astype(np.int64) % mesh
    key = (cell[:, 0] * mesh + cell[:, 1]) * mesh + cell[:, 2]
    return c[np.argsort(key, kind="stable")].reshape(-1, 3)


def replica_set(n_replicas=64, n_min=50_000, n_max=400_000, meshes=(48, 64, 96), spacing=0.6, density=8.4,
                seed=1003, spatial=True):
    rng = np.random.default_rng(seed)
    counts = (rng.integers(n_min, n_max + 1, n_replicas) // 10) * 10
    L = (counts / density) ** (1.0 / 3.0)
    choice = np.asarray(meshes)
    mesh = choice[np.argmin(np.abs(choice[None, :] - (L / spacing)[:, None]), axis=1)]
    chi_ab = rng.uniform(5.0, 25.0, n_replicas)
    sigma = rng.uniform(0.4, 0.7, n_replicas)
    xs, ts = [], []
    for r in range(n_replicas):
        box = np.full(3, L[r])
        xr = random_walk_chains(rng, int(counts[r]) // 10, 10, box)
        xs.append(mesh_ordered_chains(xr, 10, box, int(mesh[r])) if spatial else xr)
        ts.append(np.tile(np.array([0] * 5 + [1] * 5, dtype=np.int32), int(counts[r]) // 10))
    chi = np.zeros((n_replicas, 2, 2))
    chi[:, 0, 1] = chi[:, 1, 0] = chi_ab
    return ReplicaSet(counts.astype(np.int64), np.repeat(mesh[:, None], 3, axis=1).astype(np.int32),
                      np.repeat(L[:, None], 3, axis=1), np.concatenate(xs) if xs else np.zeros((0, 3)),
                      np.concatenate(ts) if ts else np.zeros(0, np.int32), chi, sigma, np.full(n_replicas, 0.05))


def hamiltonian_form(system):
    """(A, b, c0) of DefaultWithChi for ``system`` (rho0 = a = N/V)."""
    T = system.n_types
    rho0 = system.rho0
    inv = 1.0 / (system.kappa * rho0)
    chi = np.array(system.chi, dtype=np.float64)
    np.fill_diagonal(chi, 0.0)
    A = np.full((T, T), inv) + chi / rho0
    return A, np.full(T, -rho0 * inv), 0.5 * rho0 * rho0 * inv
