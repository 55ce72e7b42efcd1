"""FP64 restatement of HyMD's variable-dielectric electrostatics (oracle; test
infrastructure only): update_field_force_q_GPE (field.py:764-1112) and
compute_field_energy_q_GPE (field.py:706-761).

Step by step, in the reference's order and with its masked divisions
(``np.divide(..., where=...)`` leaves the output's previous value where the
mask is false, so ``phi_eps``, ``phi_q`` and ``elec_field_contrib`` keep the
caller's values there):

1. phi_q = c2r(H r2c(paint(q)/dV)) (field.py:971-980);
2. phi_eps = sum_t eps_t phi_t / sum_t phi_t where the denominator > 1e-6
   (phi = the filtered densities of update_field; field.py:982-989);
3. phi_q /= phi_eps where phi_eps > 1e-6 (:992);
4. phi_eta_d = c2r(i k_d r2c(phi_eps)) / phi_eps where phi_eps > 1e-6 (:999-1007);
5. polarisation charges starting from the caller's phi_pol_prev grid (the
   loop rebinds the name, so the caller's grid is not changed: HyMD's main
   loop passes its zero grid every call), at most 100 iterations while
   delta > conv_crit
   (:1012-1042): E_d = c2r(-i k_d r2c(phi_q + phi_pol_prev) / |k|^2)
   (|k|^2 := 1 at k = 0), phi_pol = w (-sum_d phi_eta_d E_d) + (1 - w) phi_pol_prev,
   delta = conv(|phi_pol - phi_pol_prev|), phi_pol_prev = phi_pol;
6. psi = c2r(r2c(4 pi k_e (phi_q + phi_pol)) / |k|^2), E_d = c2r(-i k_d r2c(psi))
   (:1048-1063);
7. elec_dot = sum_d E_d^2, contrib = elec_dot / sum_t phi_t where > 1e-6
   (:1068-1079); V_t = q_t psi - (0.5 / (4 pi k_e)) (eps_t - phi_eps) contrib
   (:1081-1086);
8. force grids c2r(-i k_d H r2c(V_t)), forces = per-type readout (:1089-1110);
9. energy dV (0.5 / (4 pi k_e)) sum phi_eps elec_dot (:755-759).

conv: "max_diff" (the reference's default, np.max), "csum" (pmesh csum: the
sum) or "euclidean_norm" (pmesh cnorm: the sum of squares) (main.py:137-165).
The spectral derivative uses k_d = 0 on Nyquist planes and |k|^2 the full k,
as the rest of the field oracle (field.py restatement).

Pinning (tests/test_oracle_gpe.py): a uniform dielectric reduces the solver
to the plain Poisson solve of update_field_force_q with dielectric eps (no
polarisation, potential equal to the PME potential, forces equal to the
per-type charge times the filtered PME field, field energy equal to the
charge-potential energy 1/2 dV sum rho_smeared psi); for a varying dielectric the converged potential
satisfies the generalised Poisson equation div(eps grad psi) =
-4 pi k_e rho_smeared (residual against the fixed-point tolerance), and the
converged polarisation satisfies its own fixed point.
"""
import numpy as np

from . import field as F


def _conv(diff, kind):
    if kind == "max_diff":
        return float(np.max(diff))
    if kind == "csum":
        return float(np.sum(diff))
    if kind == "euclidean_norm":
        return float(np.sum(diff * diff))
    raise ValueError(kind)


def update_field_force_q_gpe(x, types, charges, phi, mesh, box, sigma, dielectric_type, type_charges,
                             coulomb_constant, conv_crit=1e-6, pol_mixing=0.6, conv="max_diff", max_iter=100,
                             phi_eps=None, phi_pol_prev=None, elec_field_contrib=None):
    M = F._mesh3(mesh)
    L = F._box3(box)
    T = len(dielectric_type)
    dV = float(np.prod(L)) / float(M[0] * M[1] * M[2])
    H = F.filter_H(M, L, sigma)
    k2 = F.k_squared(M, L)
    k2[0, 0, 0] = 1.0
    kd = [F.derivative_k(M, L, d) for d in range(3)]
    q = np.asarray(charges, dtype=np.float64)
    phi_eps = np.zeros(M) if phi_eps is None else np.array(phi_eps, dtype=np.float64)
    phi_pol_prev = np.zeros(M) if phi_pol_prev is None else np.array(phi_pol_prev, dtype=np.float64)
    contrib = np.zeros(M) if elec_field_contrib is None else np.array(elec_field_contrib, dtype=np.float64)
    # 1. smeared charge density
    phi_q = F.paint(x, M, L, mass=q) / dV
    phi_q = F.c2r(F.r2c(phi_q) * H, M)
    # 2. dielectric field from the filtered densities
    den = np.zeros(M)
    num = np.zeros(M)
    for t in range(T):
        num = num + dielectric_type[t] * phi[t]
        den = den + phi[t]
    np.divide(num, den, where=np.abs(den > 1e-6), out=phi_eps)
    eps_f = F.r2c(phi_eps)
    # 3.
    np.divide(phi_q, phi_eps, where=np.abs(phi_eps > 1e-6), out=phi_q)
    # 4.
    phi_eta = []
    for d in range(3):
        g = F.c2r(1j * kd[d] * eps_f, M)
        np.divide(g, phi_eps, where=np.abs(phi_eps > 1e-6), out=g)
        phi_eta.append(g)
    # 5. polarisation charges
    it, delta = 0, 1.0
    w = pol_mixing
    phi_pol = phi_pol_prev.copy()
    while it < max_iter and delta > conv_crit:
        rf = F.r2c(phi_q + phi_pol_prev)
        E = [F.c2r(rf * (-1j * kd[d]) / k2, M) for d in range(3)]
        phi_pol = -(phi_eta[0] * E[0] + phi_eta[1] * E[1] + phi_eta[2] * E[2])
        phi_pol = w * phi_pol + (1.0 - w) * phi_pol_prev
        delta = _conv(np.abs(phi_pol - phi_pol_prev), conv)
        phi_pol_prev = phi_pol.copy()
        it += 1
    # 6. potential and field
    eps0_inv = coulomb_constant * 4.0 * np.pi
    psi_f = F.r2c(eps0_inv * (phi_q + phi_pol)) / k2
    psi = F.c2r(psi_f, M)
    E = [F.c2r(-1j * kd[d] * psi_f, M) for d in range(3)]
    # 7.
    elec_dot = E[0] * E[0] + E[1] * E[1] + E[2] * E[2]
    np.divide(elec_dot, den, where=np.abs(den > 1e-6), out=contrib)
    Vbar = [type_charges[t] * psi - (0.5 / eps0_inv) * (dielectric_type[t] - phi_eps) * contrib for t in range(T)]
    # 8. forces
    types = np.asarray(types)
    forces = np.zeros((len(types), 3))
    force_mesh = []
    for t in range(T):
        vf = F.r2c(Vbar[t]) * H
        fm = [F.c2r(-1j * kd[d] * vf, M) for d in range(3)]
        force_mesh.append(fm)
        sel = types == t
        for d in range(3):
            forces[sel, d] = F.readout(fm[d], x[sel], M, L)
    energy = field_q_energy_gpe(phi_eps, elec_dot, coulomb_constant, M, L)
    return {"phi_q": phi_q, "phi_eps": phi_eps, "phi_eta": phi_eta, "phi_pol": phi_pol, "psi": psi,
            "elec_field": E, "elec_dot": elec_dot, "elec_field_contrib": contrib, "Vbar": Vbar,
            "force_mesh": force_mesh, "forces": forces, "iterations": it, "delta": delta, "energy": energy}


def field_q_energy_gpe(phi_eps, elec_dot, coulomb_constant, mesh, box):
    """compute_field_energy_q_GPE (field.py:755-759)."""
    M = F._mesh3(mesh)
    L = F._box3(box)
    dV = float(np.prod(L)) / float(M[0] * M[1] * M[2])
    eps_0 = 1.0 / (coulomb_constant * 4.0 * np.pi)
    return float(dV * (0.5 * eps_0) * np.sum(phi_eps * elec_dot))
