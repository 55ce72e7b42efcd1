"""The pressure oracle (oracle/pressure.py) against what the mathematics
fixes: the field terms are minus the strain derivative of the field energy,
a uniform density gives rho w'(rho) - w(rho), a uniform potential fixes the
electrostatic terms."""
import numpy as np
import pytest

from oracle import field as OF
from oracle import pressure as OP


def _system(seed, n=3000, T=3, box=(4.0, 5.0, 6.0)):
    rng = np.random.default_rng(seed)
    box = np.array(box)
    x = rng.uniform(0.0, 1.0, (n, 3)) * box
    types = rng.integers(0, T, n)
    return x, types, box


def _energy(x, types, T, mesh, box, sigma, ham):
    out = OF.update_field(x, types, T, mesh, box, sigma, ham)
    return OF.field_energy(out["phi"], ham, mesh, box), out


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("which,sigma", [("with_chi", 0.55), ("squared_phi", 0.55), ("with_chi", 0.12)])
def test_field_pressure_is_minus_the_strain_derivative(d, which, sigma):
    """Stretching L_d by (1 + eps) at fixed fractional coordinates scales dV
    and the filter's k_d; -(1/V) dE/d(eps) must equal p0 + p1 + p2_d.  With
    sigma = 0.12 (below the cell size) the Nyquist planes carry weight, so
    this also fixes k_d^2 = (pi/h)^2 there."""
    T, mesh = 3, (16, 12, 20)
    x, types, box = _system(1 + d)
    rho0 = len(x) / np.prod(box)
    if which == "with_chi":
        chi = np.array([[0, 12.0, -4.0], [12.0, 0, 7.5], [-4.0, 7.5, 0]])
        ham = OF.default_with_chi(T, 0.05, rho0, rho0, chi)
    else:
        ham = OF.squared_phi(T, 0.1, rho0)
    E0, out = _energy(x, types, T, mesh, box, sigma, ham)
    p = OP.comp_pressure(out["phi"], out["phi_fourier"], ham, np.zeros((1, 3)), 72.0, mesh, box, sigma,
                         np.zeros(3), np.zeros(3))
    field_d = p[1] + p[2] + p[3 + d]
    eps = 1e-4
    E = []
    for s in (+1, -1):
        f = np.ones(3)
        f[d] += s * eps
        E.append(_energy(x * f, types, T, mesh, box * f, sigma, ham)[0])
    fd = -(E[0] - E[1]) / (2 * eps) / np.prod(box)
    assert field_d == pytest.approx(fd, rel=2e-7, abs=1e-9 * abs(p[1]))
    # the sigma^2 Laplacian term is what makes the directions differ
    assert abs(p[3 + d]) > 1e-6 * abs(p[1])


def test_uniform_density_gives_the_thermodynamic_pressure():
    T, mesh, box = 2, (8, 8, 8), np.array([3.0, 3.0, 3.0])
    rho = np.array([4.0, 2.5])
    ham = OF.default_with_chi(T, 0.05, 6.5, 6.5, np.array([[0, 10.0], [10.0, 0]]))
    phi = [np.full(mesh, r) for r in rho]
    pf = [OF.r2c(p) for p in phi]
    p = OP.comp_pressure(phi, pf, ham, np.zeros((1, 3)), 72.0, mesh, box, 0.5, np.zeros(3), np.zeros(3))
    w = 0.5 * rho @ ham.A @ rho + ham.b @ rho + ham.c0
    dw = ham.A @ rho + ham.b
    assert p[1] + p[2] == pytest.approx(rho @ dw - w, rel=1e-12)
    assert np.allclose(p[3:6], 0.0, atol=1e-12)


def test_uniform_potential_fixes_the_electrostatic_terms():
    T, mesh, box = 2, (8, 10, 12), np.array([3.0, 4.0, 5.0])
    x, types, _ = _system(7, n=500, T=T, box=box)
    ham = OF.Quadratic(np.zeros((T, T)), np.zeros(T), 0.0)   # no field energy: only the electrostatic terms
    out = OF.update_field(x, types, T, mesh, box, 0.5, ham)
    rng = np.random.default_rng(3)
    phi_q = rng.normal(size=mesh)
    psi0, qt, e_self = 0.7, np.array([1.5, -0.5]), 3.25
    psi = np.full(mesh, psi0)
    V = np.prod(box)
    dV = V / np.prod(mesh)
    p = OP.comp_pressure(out["phi"], out["phi_fourier"], ham, np.zeros((1, 3)), 72.0, mesh, box, 0.5,
                         np.zeros(3), np.zeros(3), phi_q=phi_q, psi=psi, type_charges=qt, self_energy=e_self)
    N_t = np.array([np.sum(types == t) for t in range(T)])
    assert p[2] == pytest.approx(psi0 * (qt @ N_t) / V, rel=1e-10)
    assert np.allclose(p[3:6], 0.0, atol=1e-10)
    assert p[1] == pytest.approx(-0.5 * psi0 * np.sum(phi_q) * dV / V + e_self / V, rel=1e-12)


def test_kinetic_and_bonded_terms():
    T, mesh, box = 1, (8, 8, 8), np.array([2.0, 3.0, 4.0])
    ham = OF.Quadratic(np.zeros((T, T)), np.zeros(T), 0.0)
    phi = [np.zeros(mesh)]
    v = np.random.default_rng(0).normal(size=(100, 3))
    bp, ap = np.array([1.0, -2.0, 3.0]), np.array([0.5, 0.25, -1.0])
    p = OP.comp_pressure(phi, [OF.r2c(phi[0])], ham, v, 72.0, mesh, box, 0.5, bp, ap)
    V = 24.0
    K = 0.5 * 72.0 * np.sum(v ** 2)
    assert p[0] == pytest.approx(2 * K / (3 * V), rel=1e-14)
    np.testing.assert_allclose(p[6:9], bp / V, rtol=1e-15)
    np.testing.assert_allclose(p[9:12], ap / V, rtol=1e-15)
    np.testing.assert_array_equal(p[12:15], 0.0)
    np.testing.assert_allclose(p[15:18], p[0] + p[6:9] + p[9:12], rtol=1e-14)
