"""The GPE oracle (oracle/gpe.py, field.py:706-761, 764-1112) pinned by what
the electrostatics fixes, not by its own formulas:

* a uniform dielectric eps makes the solver the plain Poisson solve of
  update_field_force_q with dielectric eps: no polarisation charge, the PME
  potential, forces q_t readout(c2r(-i k H psi_pme)) (the GPE filters the
  potential once more), and a field energy eps/(8 pi k_e) dV sum |E|^2 equal
  to the charge-potential energy 1/2 dV sum rho_smeared psi (Gauss / Parseval);
* a varying dielectric: the converged potential solves the generalised Poisson
  equation div(eps grad psi) = -4 pi k_e rho_smeared (spectral derivatives on
  the grid; the same potential without polarisation charges misses it by
  O(1)), and the polarisation charge is the fixed point of its iteration;
* the reference's masked divisions keep the caller's previous values where the
  density or the dielectric vanishes, and a start from a converged
  polarisation stops at once.
"""
import numpy as np
import pytest

from oracle import field as F
from oracle import gpe as G

KE = 138.935458


def _system(seed=3, n=8000, M=(24, 24, 24), L=(6.0, 6.0, 6.0), sigma=0.6, slab=True):
    rng = np.random.default_rng(seed)
    L = np.asarray(L, dtype=np.float64)
    x = rng.uniform(0, 1, (n, 3)) * L
    types = (x[:, 0] > L[0] / 2).astype(np.int32) if slab else rng.integers(0, 2, n).astype(np.int32)
    q = np.zeros(n)
    q[: n // 20] = 1.0
    q[n // 20: n // 10] = -1.0
    out = F.update_field(x, types, 2, M, L, sigma, F.Quadratic(np.eye(2), np.zeros(2), 0.0))
    return x, types, q, out["phi"], M, L, sigma


def _div_eps_grad(eps, psi, M, L):
    kd = [F.derivative_k(M, L, d) for d in range(3)]
    return sum(F.c2r(1j * kd[d] * F.r2c(eps * F.c2r(1j * kd[d] * F.r2c(psi), M)), M) for d in range(3))


def _smeared(x, q, M, L, sigma):
    dV = np.prod(L) / np.prod(M)
    return F.c2r(F.r2c(F.paint(x, M, L, mass=q) / dV) * F.filter_H(M, L, sigma), M)


def test_uniform_dielectric_is_the_poisson_solve():
    x, types, q, phi, M, L, sigma = _system(slab=False)
    eps = 45.0
    tq = [0.7, -1.3]
    r = G.update_field_force_q_gpe(x, types, q, phi, M, L, sigma, [eps, eps], tq, KE)
    pme = F.update_field_force_q(x, q, M, L, sigma, KE, eps)
    assert np.abs(r["phi_pol"]).max() < 1e-9 * np.abs(r["phi_q"]).max()
    np.testing.assert_allclose(r["psi"], pme["psi"], rtol=0, atol=1e-10 * np.abs(pme["psi"]).max())
    # forces: q_t times the readout of the once more filtered PME field
    H = F.filter_H(M, L, sigma)
    for t in range(2):
        sel = types == t
        for d in range(3):
            e = F.c2r(-1j * F.derivative_k(M, L, d) * pme["psi_fourier"] * H, M)
            want = tq[t] * F.readout(e, x[sel], M, L)
            np.testing.assert_allclose(r["forces"][sel, d], want, rtol=0, atol=1e-9 * np.abs(want).max())
    # field energy = 1/2 dV sum rho_smeared psi
    dV = np.prod(L) / np.prod(M)
    want = 0.5 * dV * np.sum(_smeared(x, q, M, L, sigma) * pme["psi"])
    assert r["energy"] == pytest.approx(want, rel=1e-6)


def test_varying_dielectric_solves_the_generalised_poisson_equation():
    x, types, q, phi, M, L, sigma = _system()
    r = G.update_field_force_q_gpe(x, types, q, phi, M, L, sigma, [5.0, 80.0], [0.0, 0.0], KE, conv_crit=1e-8,
                                   max_iter=400)
    assert r["delta"] <= 1e-8 and r["iterations"] < 400
    eps = r["phi_eps"]
    assert eps.min() > 5.0 and eps.max() < 80.0 and eps.max() - eps.min() > 50.0
    rhs = -4.0 * np.pi * KE * _smeared(x, q, M, L, sigma)
    res = _div_eps_grad(eps, r["psi"], M, L) - rhs
    rel = np.sqrt(np.mean(res ** 2)) / np.sqrt(np.mean(rhs ** 2))
    assert rel < 2e-3
    # the same potential without polarisation charges misses the equation
    k2 = F.k_squared(M, L)
    k2[0, 0, 0] = 1.0
    psi0 = F.c2r(F.r2c(4.0 * np.pi * KE * r["phi_q"]) / k2, M)
    res0 = _div_eps_grad(eps, psi0, M, L) - rhs
    assert np.sqrt(np.mean(res0 ** 2)) / np.sqrt(np.mean(rhs ** 2)) > 50 * rel
    # the polarisation charge is its iteration's fixed point
    kd = [F.derivative_k(M, L, d) for d in range(3)]
    rf = F.r2c(r["phi_q"] + r["phi_pol"])
    E = [F.c2r(rf * (-1j * kd[d]) / k2, M) for d in range(3)]
    fixed = -(r["phi_eta"][0] * E[0] + r["phi_eta"][1] * E[1] + r["phi_eta"][2] * E[2])
    assert np.abs(fixed - r["phi_pol"]).max() < 1e-6 * np.abs(r["phi_pol"]).max()


def test_masked_divisions_and_warm_start():
    # sparse particles, no filter: empty cells have zero density, so phi_eps
    # (and elec_field_contrib) keep the caller's values there
    x, types, q, phi, M, L, _ = _system(n=300, sigma=0.0)
    den = phi[0] + phi[1]
    empty = den <= 1e-6
    assert empty.any() and (~empty).any()
    prev_eps = np.full(M, 7.0)
    prev_contrib = np.full(M, -3.0)
    r = G.update_field_force_q_gpe(x, types, q, phi, M, L, 0.0, [5.0, 80.0], [1.0, -1.0], KE,
                                   phi_eps=prev_eps, elec_field_contrib=prev_contrib, max_iter=3)
    assert np.all(r["phi_eps"][empty] == 7.0)
    assert np.all(r["elec_field_contrib"][empty] == -3.0)
    assert np.all((r["phi_eps"][~empty] >= 5.0 - 1e-6) & (r["phi_eps"][~empty] <= 80.0 + 1e-6))
    # started from the converged polarisation, the iteration stops at once
    x, types, q, phi, M, L, sigma = _system()
    a = G.update_field_force_q_gpe(x, types, q, phi, M, L, sigma, [5.0, 80.0], [0.0, 0.0], KE)
    b = G.update_field_force_q_gpe(x, types, q, phi, M, L, sigma, [5.0, 80.0], [0.0, 0.0], KE,
                                   phi_eps=a["phi_eps"], phi_pol_prev=a["phi_pol"])
    assert a["iterations"] > 5 and b["iterations"] <= 2


@pytest.mark.parametrize("conv", ["max_diff", "csum", "euclidean_norm"])
def test_convergence_measures(conv):
    x, types, q, phi, M, L, sigma = _system(n=4000, M=(16, 16, 16), L=(4.0, 4.0, 4.0))
    r = G.update_field_force_q_gpe(x, types, q, phi, M, L, sigma, [5.0, 80.0], [0.0, 0.0], KE, conv=conv,
                                   conv_crit=1e-6)
    assert r["delta"] <= 1e-6 and 1 < r["iterations"] < 100
