"""Internal pressure (oracle; test infrastructure only).

comp_pressure: pressure.py:9-179 with comp_laplacian (pressure.py:182-200,
field.py:406-425), restated in FP64 real space on the oracle's grids:

    p_kin   = 2 K / (3 V),  K = 1/2 m sum v^2
    p0      = -(dV / V) sum_cells [w(phi~) + w_elec]          w_elec = 1/2 phi_q psi - E_self / V
    p1      =  (dV / V) sum_t sum_cells V_bar_t phi~_t          V_bar_t = (A phi~ + b)_t [+ q_t psi]
    p2_d    =  (dV / V) sigma^2 sum_t sum_cells V_bar_t lap_d phi~_t,   lap_d phi~_t = c2r(-k_d^2 phi~^_t)
    p_bond  = bond_pr / V,  p_angle = angle_pr / V,  p_dihedral = 0
    p_tot_d = p_kin + p0 + p1 + p2_d + p_bond_d + p_angle_d

returned as the reference's 18-vector.  phi~ are the filtered densities and
phi~^ their spectra (pmesh normalisation), psi the electrostatic potential
and phi_q the charge density of the PME path; the electrostatic terms enter
only with ``electrostatics`` (coulombtype PIC_Spectral), and V_bar then
carries q_t psi with the per-type charges (hamiltonian.py:167-183).  k_d on
the Nyquist plane of an even edge is -pi/h, so lap_d uses (pi/h)^2 there.

Pinning (tests/test_oracle_pressure.py): the field part p0 + p1 + p2_d equals
-(1/V) dE/d(eps_d) of the oracle's own field energy under the affine
stretch L_d -> L_d (1 + eps_d) with fixed particle fractional coordinates
(central differences): scaling dV and the filter's k_d reproduces exactly
these three terms, the sigma^2 Laplacian from dH/d(eps_d) = sigma^2 k_d^2 H,
which also fixes the Nyquist value of k_d^2; a uniform density gives the
textbook p = rho w'(rho) - w(rho) with p2 = 0; the electrostatic terms by
a uniform potential (p1_elec = psi0 sum_t q_t N_t / V, p2_elec = 0,
p0_elec = -psi0 Q / (2V) + E_self / V).
"""
import numpy as np

from . import field as OF


def comp_laplacian(phi_fourier, mesh, box):
    """[T][3] real grids c2r(-k_d^2 phi~^_t) (pressure.py:182-200)."""
    M = OF._mesh3(mesh)
    ks = OF.kvectors(M, box)
    return [[OF.c2r(-(ks[d] ** 2) * pf, M) for d in range(3)] for pf in phi_fourier]


def comp_pressure(phi, phi_fourier, ham, velocities, mass, mesh, box, sigma, bond_pr, angle_pr,
                  phi_q=None, psi=None, type_charges=None, self_energy=0.0):
    """The reference's 18 pressure components (pressure.py:9-179)."""
    M = OF._mesh3(mesh)
    L = OF._box3(box)
    V = float(np.prod(L))
    dV = V / float(M[0] * M[1] * M[2])
    T = len(phi)
    elec = psi is not None
    w = ham.w(phi) * dV
    if elec:
        w = w + (0.5 * phi_q * psi - self_energy / V) * dV
    K = 0.5 * mass * float(np.sum(np.asarray(velocities, dtype=np.float64) ** 2))
    p_kin = 2.0 / (3.0 * V) * K
    p0 = -1.0 / V * float(np.sum(w))
    v_bar = [ham.v_ext(t, phi) + (type_charges[t] * psi if elec else 0.0) for t in range(T)]
    p1 = float(sum(np.sum(dV / V * v_bar[t] * phi[t]) for t in range(T)))
    lap = comp_laplacian(phi_fourier, M, L)
    p2 = np.array([sum(float(np.sum(dV / V * sigma ** 2 * v_bar[t] * lap[t][d])) for t in range(T))
                   for d in range(3)])
    p_bond = np.asarray(bond_pr, dtype=np.float64) / V
    p_angle = np.asarray(angle_pr, dtype=np.float64) / V
    p_dih = np.zeros(3)
    p_tot = p_kin + p0 + p1 + p2 + p_bond + p_angle + p_dih
    return np.array([p_kin, p0, p1, p2[0], p2[1], p2[2], *p_bond, *p_angle, *p_dih, *p_tot])
