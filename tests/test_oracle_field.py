"""Pin the field oracle (oracle/field.py) before trusting it.

* paint mass conservation (test/test_hamiltonian.py:106-109) and brute force;
* fixed-point paint: exact integer mass, integer-cell translation = roll;
* Gaussian-core closed-form energies at 160^3 / 15 nm
  (test/test_hamiltonian.py:172-227, 270-360) and the Mathematica values;
* field forces on grid nodes = continuum Gaussian-core pair forces
  (compute_gaussian_core.f90:60-89 with a coupling matrix);
* sum of forces = 0, 90-degree rotation permutes forces;
* PME: two opposite charges = Gaussian-smeared Coulomb; neutral sum F = 0;
* PME energy (field_q_energy with the self energy removed): the Gaussian-smeared
  Coulomb pair energies of a neutral set of charges on nodes, with the
  periodic (tinfoil) dipole term of a cubic lattice.
"""
import numpy as np
import pytest
from scipy.special import erf

from oracle import field as F


def test_paint_mass_conservation(goldens):
    g = goldens["energy_functional_small"]
    x = np.array(g["positions"])
    types = np.array(g["types"])
    for t, count in ((0, 2.0), (1, 3.0)):
        grid = F.paint(x[types == t], g["mesh"], g["box"])
        assert grid.sum() == pytest.approx(count, abs=1e-14)


def test_paint_brute_force():
    rng = np.random.default_rng(0)
    M, L = (6, 8, 10), np.array([3.0, 4.5, 5.0])
    x = rng.uniform(-1, 1, (7, 3)) * L * 1.5
    m = rng.uniform(0.5, 2.0, 7)
    grid = F.paint(x, M, L, mass=m)
    ref = np.zeros(M)
    for p in range(len(x)):
        g = [x[p, d] / (L[d] / M[d]) for d in range(3)]
        base = [int(np.floor(v)) for v in g]
        for i in range(2):
            for j in range(2):
                for k in range(2):
                    w = 1.0
                    for d, o in enumerate((i, j, k)):
                        f = g[d] - base[d]
                        w *= f if o else 1.0 - f
                    ref[(base[0] + i) % M[0], (base[1] + j) % M[1], (base[2] + k) % M[2]] += m[p] * w
    np.testing.assert_allclose(grid, ref, atol=1e-13)


def test_paint_fixed_exact_mass_and_translation():
    rng = np.random.default_rng(1)
    M = (16, 16, 16)
    L = np.full(3, 4.0)  # h = 0.25: dyadic, so shifts by whole cells are exact
    x = np.round(rng.uniform(0, 4, (300, 3)) * 2 ** 20) / 2 ** 20
    q = F.paint_fixed(x, M, L)
    assert int(q.sum()) == 300 * 2 ** 32
    np.testing.assert_allclose(q / 2.0 ** 32, F.paint(x, M, L), atol=8 * 2.0 ** -33)
    shifted = np.mod(x + np.array([3, 5, 7]) * 0.25, L)
    q2 = F.paint_fixed(shifted, M, L)
    np.testing.assert_array_equal(q2, np.roll(q, (3, 5, 7), axis=(0, 1, 2)))


def test_r2c_c2r_round_trip():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(12, 10, 8))
    np.testing.assert_allclose(F.c2r(F.r2c(x), x.shape), x, atol=1e-12)
    np.testing.assert_allclose(F.r2c(x)[0, 0, 0].real, x.mean(), atol=1e-14)


def _gc_energy_closed_form(r, types, A, b, c0, sigma, box):
    e, _ = F.gaussian_core_pairs(r, types, A, sigma, box)
    counts = np.bincount(types, minlength=len(b))
    return e + float(np.dot(b, counts)) + c0 * float(np.prod(box))


@pytest.mark.parametrize("case", [0, 1])
def test_gaussian_core_no_chi(goldens, case):
    """SquaredPhi and DefaultNoChi energies on the grid equal the continuum
    closed form of test/test_hamiltonian.py:213-227 to 1e-6."""
    g = goldens["gaussian_core_no_chi"]
    kappa, sigma = g["params"][case]
    r = np.array(g["positions"])
    box = np.array(g["box"])
    N = len(r)
    rho0 = N / np.prod(box)
    types = np.zeros(N, dtype=int)
    c = 16.0 * np.pi ** 1.5 * kappa * sigma ** 3 * rho0
    diag = N / c
    offdiag = sum(2 * np.exp(-np.sum((r[i] - r[j]) ** 2) / (4 * sigma ** 2)) / c
                  for i in range(N) for j in range(i + 1, N))
    for ham, E in ((F.squared_phi(1, kappa, rho0), diag + offdiag),
                   (F.default_no_chi(1, kappa, rho0, rho0), diag + offdiag - 0.5 * N / kappa)):
        out = F.update_field(r, types, 1, g["mesh"], box, sigma, ham)
        assert F.field_energy(out["phi"], ham, g["mesh"], box) == pytest.approx(E, abs=g["tolerance"])
        assert _gc_energy_closed_form(r, types, ham.A, ham.b, ham.c0, sigma, box) == pytest.approx(E, abs=1e-9)


def test_gaussian_core_mathematica(goldens):
    """Mathematica values of test/test_hamiltonian.py:243-258 for w = phi^2
    (sigma of case 0): the closed form to 1e-15 and the 160^3 grid to 1e-12.
    The W = phi^2/(2 phi0 kappa) values (:229-241) have an unstated kappa with
    phi0 = N/V; their ratios to the one-particle value are kappa-free."""
    g = goldens["gaussian_core_no_chi"]
    sigma = g["params"][0][1]
    box = np.array(g["box"])
    r = np.array(g["positions"])
    ham = F.Quadratic([[2.0]], [0.0], 0.0)  # w = phi^2
    for n, m_key, g_key in ((1, "mathematica_phi2_one_particle", "grid160_phi2_one_particle"),
                            (2, "mathematica_phi2_two_particles", "grid160_phi2_two_particles")):
        e, _ = F.gaussian_core_pairs(r[:n], np.zeros(n, dtype=int), ham.A, sigma, box)
        assert e == pytest.approx(g[m_key], rel=1e-15)
        out = F.update_field(r[:n], np.zeros(n, dtype=int), 1, g["mesh"], box, sigma, ham)
        assert F.field_energy(out["phi"], ham, g["mesh"], box) == pytest.approx(g[g_key], abs=1e-12)
    g0, _ = F.gaussian_core_pairs(r[:1], np.zeros(1, dtype=int), ham.A, sigma, box)
    for n, key in ((2, "mathematica_prefactor_two_particles"), (3, "mathematica_prefactor_three_particles")):
        e, _ = F.gaussian_core_pairs(r[:n], np.zeros(n, dtype=int), ham.A, sigma, box)
        # prefactor V/(2 kappa N): energy ratio to the one-particle value is e/(n g0)
        assert e / (n * g0) == pytest.approx(g[key] / g["mathematica_prefactor_one_particle"], rel=1e-14)


@pytest.mark.parametrize("case", [0, 1])
def test_gaussian_core_with_chi(goldens, case):
    """DefaultWithChi = closed form + sum 2 kappa chi_ij exp(-r^2/4 sigma^2)/c
    (test/test_hamiltonian.py:334-360)."""
    g = goldens["gaussian_core_with_chi"]
    kappa, sigma = g["params"][case]
    r = np.array(g["positions"])
    types = np.array(g["types"])
    box = np.array(g["box"])
    N = len(r)
    rho0 = N / np.prod(box)
    chi = np.zeros((3, 3))
    for key, (i, j) in (("AB", (0, 1)), ("AC", (0, 2)), ("BC", (1, 2))):
        chi[i, j] = chi[j, i] = g["chi"][key]
    ham = F.default_with_chi(3, kappa, rho0, rho0, chi)
    out = F.update_field(r, types, 3, g["mesh"], box, sigma, ham)
    c = 16.0 * np.pi ** 1.5 * kappa * sigma ** 3 * rho0
    E = N / c - 0.5 * N / kappa
    for i in range(N):
        for j in range(i + 1, N):
            e = np.exp(-np.sum((r[i] - r[j]) ** 2) / (4 * sigma ** 2))
            E += 2 * e / c + (2 * kappa * chi[types[i], types[j]] * e / c if types[i] != types[j] else 0.0)
    assert F.field_energy(out["phi"], ham, g["mesh"], box) == pytest.approx(E, abs=g["tolerance"])


def _nodes_system(seed, n, M, L, T):
    rng = np.random.default_rng(seed)
    h = L / M
    idx = rng.choice(M ** 3, n, replace=False)
    nodes = np.stack(np.unravel_index(idx, (M, M, M)), axis=1)
    return nodes * h, rng.integers(0, T, n)


def test_field_forces_on_nodes_match_pair_forces():
    """Particles on grid nodes: paint and readout are exact deltas, the filter
    is band-limited, so the grid forces equal the continuum pair forces."""
    M, L, sigma = 64, 10.0, 0.5
    x, types = _nodes_system(3, 12, M, L, 2)
    rho0 = 16.0
    chi = np.array([[0.0, 20.0], [20.0, 0.0]])
    ham = F.default_with_chi(2, 0.05, rho0, rho0, chi)
    out = F.update_field(x, types, 2, M, np.full(3, L), sigma, ham)
    f = F.compute_field_force(x, types, out["force_mesh"], M, np.full(3, L))
    _, f_ref = F.gaussian_core_pairs(x, types, ham.A, sigma, np.full(3, L))
    np.testing.assert_allclose(f, f_ref, atol=1e-9 * np.abs(f_ref).max())


def test_field_force_sum_and_rotation():
    rng = np.random.default_rng(4)
    M, L = 32, 8.0
    x = rng.uniform(0, L, (400, 3))
    types = rng.integers(0, 3, 400)
    chi = np.array([[0, 10.0, 5.0], [10.0, 0, 25.0], [5.0, 25.0, 0]])
    ham = F.default_with_chi(3, 0.05, 400 / L ** 3, 400 / L ** 3, chi)
    box = np.full(3, L)
    f = F.compute_field_force(x, types, F.update_field(x, types, 3, M, box, 0.45, ham)["force_mesh"], M, box)
    np.testing.assert_allclose(f.sum(axis=0), 0.0, atol=1e-10 * np.abs(f).max() * len(x))
    # rotation by 90 degrees about z: (x, y, z) -> (L - y, x, z)
    xr = np.stack([np.mod(L - x[:, 1], L), x[:, 0], x[:, 2]], axis=1)
    fr = F.compute_field_force(xr, types, F.update_field(xr, types, 3, M, box, 0.45, ham)["force_mesh"], M, box)
    np.testing.assert_allclose(fr, np.stack([-f[:, 1], f[:, 0], f[:, 2]], axis=1), atol=1e-9 * np.abs(f).max())


def test_pme_two_charges():
    """Opposite unit charges on nodes: the filtered Poisson kernel
    4 pi k_e H/(eps k^2) gives F = -q1 q2 k_e/eps d/dr[erf(r/(sqrt2 sigma))/r]."""
    M, L, sigma, ke, eps = 128, 24.0, 0.5, 138.935458, 45.0
    h = L / M
    x = np.array([[60, 64, 64], [68, 64, 64]], dtype=float) * h
    q = np.array([1.0, -1.0])
    out = F.update_field_force_q(x, q, M, np.full(3, L), sigma, ke, eps)
    r = 8 * h
    s = np.sqrt(2.0) * sigma
    dpot = (2 / np.sqrt(np.pi) * np.exp(-(r / s) ** 2) / s) / r - erf(r / s) / r ** 2
    f_expect = -q[0] * q[1] * ke / eps * dpot  # along +x on particle 1 (attractive: toward particle 0)
    assert out["forces"][1, 0] == pytest.approx(f_expect, rel=1e-2)
    assert out["forces"][0, 0] == pytest.approx(-f_expect, rel=1e-2)
    np.testing.assert_allclose(out["forces"].sum(axis=0), 0.0, atol=1e-10)


def test_self_energy():
    q = np.array([1.0, -1.0, 0.5])
    e = F.self_energy_q(q, 0.5, 138.935458, 45.0)
    assert e == pytest.approx(138.935458 / 45.0 / np.sqrt(2 * np.pi * 0.25) * 2.25, rel=1e-14)


@pytest.mark.parametrize("sep, tol", [(8, 1e-5), (16, 2e-4)])
def test_pme_energy_two_charges(sep, tol):
    """E_q = dV sum(1/2 phi_q psi) - E_self (field.py:697-699) for charges on
    nodes.  With the filter applied once the pair potential is
    k_e/eps erf(r/(sqrt2 sigma))/r and its r -> 0 limit is exactly E_self's
    per-charge term (field.py:231-238), so the self terms cancel and
    E_q = sum_{i<j} k_e/eps q_i q_j erf(r_ij/(sqrt2 sigma))/r_ij
          - 2 pi/(3V) k_e/eps |sum_i q_i r_i|^2
    for a neutral set in a cubic periodic box without the k = 0 term (the
    r^2 term of the periodic Coulomb potential; the next term is O(r^4/L^5)).
    A dropped 1/2, a missing self energy or a wrong dV fail by O(1)."""
    M, L, sigma, ke, eps = 128, 24.0, 0.5, 138.935458, 45.0
    h = L / M
    box = np.full(3, L)
    x = np.array([[60, 64, 64], [60 + sep, 64, 64]], dtype=float) * h
    q = np.array([1.0, -1.0])
    out = F.update_field_force_q(x, q, M, box, sigma, ke, eps)
    e_self = F.self_energy_q(q, sigma, ke, eps)
    e = F.field_q_energy(out["phi_q"], out["psi"], e_self, M, box)
    r = sep * h
    dipole = sep * h  # |sum q_i r_i|
    expect = ke / eps * (-1.0) * erf(r / (np.sqrt(2.0) * sigma)) / r - 2 * np.pi / (3 * L ** 3) * ke / eps * dipole ** 2
    assert e == pytest.approx(expect, rel=tol)
    # the self energy is not small against the pair energy: leaving it out fails
    assert abs(e + e_self - expect) > 100 * tol * abs(expect)


def test_pme_energy_neutral_cluster():
    """Four charges (+1, -0.5, -1.5, +1) on nodes: the pair-sum closed form of
    test_pme_energy_two_charges within the O(r^4/L^5) lattice term."""
    M, L, sigma, ke, eps = 128, 24.0, 0.5, 138.935458, 45.0
    h = L / M
    box = np.full(3, L)
    x = np.array([[40, 64, 64], [70, 60, 64], [50, 80, 70], [44, 71, 59]], dtype=float) * h
    q = np.array([1.0, -0.5, -1.5, 1.0])
    out = F.update_field_force_q(x, q, M, box, sigma, ke, eps)
    e = F.field_q_energy(out["phi_q"], out["psi"], F.self_energy_q(q, sigma, ke, eps), M, box)
    pair = 0.0
    for i in range(4):
        for j in range(i + 1, 4):
            r = np.linalg.norm(x[i] - x[j])
            pair += ke / eps * q[i] * q[j] * erf(r / (np.sqrt(2.0) * sigma)) / r
    p = (q[:, None] * x).sum(axis=0)
    expect = pair - 2 * np.pi / (3 * L ** 3) * ke / eps * float(p @ p)
    assert e == pytest.approx(expect, rel=2e-3)
