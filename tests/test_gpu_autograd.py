"""Differentiable field energy: exact chi / kappa / sigma derivatives of the
grid energy against central finite differences (h = 1e-3 relative) of the
FP64 oracle energy, and the autograd.Function's position gradient against
the reference forces.  FP32 device spectra, FP64 moments: 1e-5 relative."""
import numpy as np
import pytest
import torch

from oracle import field as F

pytestmark = pytest.mark.gpu


def _system(kind):
    from hymd_amd import synth
    if kind == "diblock":
        s = synth.diblock_melt(n_chains=500, box=7.5, mesh=32, seed=1000)
    else:
        s = synth.lipid_water(n_chains=800, n_water=6000, box=9.0, mesh=36, seed=17)
    return s


def _oracle_energy(s, x, chi, kappa, sigma):
    T = s.n_types
    rho0 = s.n / np.prod(s.box)
    ham = F.default_with_chi(T, kappa, rho0, rho0, chi)
    out = F.update_field(x, s.types, T, s.mesh, s.box, sigma, ham)
    return F.field_energy(out["phi"], ham, s.mesh, s.box)


@pytest.mark.parametrize("kind", ["diblock", "lipid"])
def test_parameter_gradients_vs_fd(hip, kind):
    from hymd_amd.autograd import energy_and_gradients, moments
    from hymd_amd.field import ParticleMesh
    s = _system(kind)
    T = s.n_types
    rho0 = s.n / np.prod(s.box)
    x32 = s.positions.astype(np.float32)
    x64 = x32.astype(np.float64)
    pm = ParticleMesh(s.mesh, s.box, T)
    P, Q, phi0 = moments(pm, torch.as_tensor(x32).cuda(), torch.as_tensor(s.types).cuda(), T, s.sigma)
    E, dchi, dk, ds, _ = energy_and_gradients(P, Q, phi0, s.chi, s.kappa, rho0, rho0, s.sigma, np.prod(s.box))
    E_ref = _oracle_energy(s, x64, s.chi, s.kappa, s.sigma)
    assert E == pytest.approx(E_ref, rel=1e-6)
    h = 1e-3
    # chi of one pair (both matrix entries)
    t, u = 0, 1
    c = s.chi[t, u]
    cp, cm = s.chi.copy(), s.chi.copy()
    cp[t, u] = cp[u, t] = c * (1 + h)
    cm[t, u] = cm[u, t] = c * (1 - h)
    fd = (_oracle_energy(s, x64, cp, s.kappa, s.sigma) - _oracle_energy(s, x64, cm, s.kappa, s.sigma)) / (2 * c * h)
    assert dchi[t, u] == pytest.approx(fd, rel=1e-5, abs=1e-6 * abs(E))
    k = s.kappa
    fd = (_oracle_energy(s, x64, s.chi, k * (1 + h), s.sigma) - _oracle_energy(s, x64, s.chi, k * (1 - h), s.sigma)) / (
        2 * k * h)
    assert dk == pytest.approx(fd, rel=1e-5)
    sg = s.sigma
    fd = (_oracle_energy(s, x64, s.chi, s.kappa, sg * (1 + h)) - _oracle_energy(s, x64, s.chi, s.kappa, sg * (1 - h))) / (
        2 * sg * h)
    assert ds == pytest.approx(fd, rel=1e-4)


def test_autograd_function(hip):
    from hymd_amd.autograd import field_energy
    from hymd_amd.field import ParticleMesh
    from hymd_amd.synth import hamiltonian_form
    s = _system("diblock")
    T = s.n_types
    rho0 = s.n / np.prod(s.box)
    pm = ParticleMesh(s.mesh, s.box, T)
    x = torch.as_tensor(s.positions.astype(np.float32)).cuda().requires_grad_(True)
    chi = torch.as_tensor(s.chi, dtype=torch.float64).cuda().requires_grad_(True)
    kappa = torch.tensor(s.kappa, dtype=torch.float64, requires_grad=True)
    sigma = torch.tensor(s.sigma, dtype=torch.float64, requires_grad=True)
    types = torch.as_tensor(s.types).cuda()
    E = field_energy(x, chi, kappa, sigma, pm, types, rho0, rho0)
    E.backward()
    # position gradient = -(reference forces)
    A, b, c0 = hamiltonian_form(s)
    ham = F.Quadratic(A, b, c0)
    x64 = s.positions.astype(np.float32).astype(np.float64)
    ref = F.update_field(x64, s.types, T, s.mesh, s.box, s.sigma, ham)
    f_ref = F.compute_field_force(x64, s.types, ref["force_mesh"], s.mesh, s.box)
    g = -x.grad.detach().cpu().numpy()
    assert np.sqrt(np.mean((g - f_ref) ** 2)) / np.sqrt(np.mean(f_ref ** 2)) < 1e-5
    # a symmetric chi perturbation picks up both entries
    gchi = chi.grad.cpu().numpy()
    assert gchi[0, 1] == pytest.approx(gchi[1, 0])
    assert gchi[0, 0] == 0.0
    assert float(kappa.grad) != 0.0 and float(sigma.grad) != 0.0


def _oracle_forces(s, x64, A, sigma, b=None):
    T = s.n_types
    ham = F.Quadratic(A, np.zeros(T) if b is None else b, 0.0)
    out = F.update_field(x64, s.types, T, s.mesh, s.box, sigma, ham)
    return F.compute_field_force(x64, s.types, out["force_mesh"], s.mesh, s.box)


@pytest.mark.parametrize("kind", ["diblock", "lipid"])
def test_adjoint_against_oracle_linearity(hip, kind):
    """The forces are linear in the coupling A and the shift b does not reach
    them, so dL/dA_ts for L = sum_i g_i . F_i is exactly sum_i g_i . F_i of
    the unit coupling E_ts: the FP64 oracle's forward forces pin the adjoint
    (a transposed A, a missing filter factor or a sign error fail it).
    dL/dsigma against central differences of the oracle's L (h = 1e-3
    relative).  FP32 spectra: 1e-5 relative to the largest entry."""
    from hymd_amd.autograd import field_adjoint
    from hymd_amd.field import ParticleMesh
    from hymd_amd.synth import hamiltonian_form
    s = _system(kind)
    T = s.n_types
    x32 = s.positions.astype(np.float32)
    x64 = x32.astype(np.float64)
    rng = np.random.default_rng(5)
    g = rng.normal(size=(s.n, 3))
    pm = ParticleMesh(s.mesh, s.box, T)
    A, b, _ = hamiltonian_form(s)
    xd = torch.as_tensor(x32).cuda()
    td = torch.as_tensor(s.types).cuda()
    dA, ds = field_adjoint(pm, xd, td, T, s.sigma, A, torch.as_tensor(g, dtype=torch.float32).cuda())
    ref = np.zeros((T, T))
    for t in range(T):
        for u in range(T):
            E = np.zeros((T, T))
            E[t, u] = 1.0
            ref[t, u] = np.sum(g * _oracle_forces(s, x64, E, s.sigma))
    assert not np.allclose(ref, ref.T, rtol=1e-3)   # the test tells t from u
    np.testing.assert_allclose(dA, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
    # b reaches no force
    f0 = _oracle_forces(s, x64, A, s.sigma)
    f1 = _oracle_forces(s, x64, A, s.sigma, b=b)
    np.testing.assert_allclose(f1, f0, rtol=0, atol=1e-12 * np.abs(f0).max())
    h = 1e-3 * s.sigma
    fd = (np.sum(g * _oracle_forces(s, x64, A, s.sigma + h)) - np.sum(g * _oracle_forces(s, x64, A, s.sigma - h))) / (
        2 * h)
    assert ds == pytest.approx(fd, rel=1e-4)


def test_adjoint_reuses_moments_binning(hip):
    """reuse = 1 after a moments call gives the same bits as a fresh call;
    reuse without a preceding moments call is refused."""
    from hymd_amd._lib import HymdError
    from hymd_amd.autograd import field_adjoint, moments
    from hymd_amd.field import ParticleMesh
    from hymd_amd.synth import hamiltonian_form
    s = _system("lipid")
    T = s.n_types
    pm = ParticleMesh(s.mesh, s.box, T)
    A, _, _ = hamiltonian_form(s)
    xd = torch.as_tensor(s.positions.astype(np.float32)).cuda()
    td = torch.as_tensor(s.types).cuda()
    g = torch.as_tensor(np.random.default_rng(9).normal(size=(s.n, 3)), dtype=torch.float32).cuda()
    fresh = field_adjoint(pm, xd, td, T, s.sigma, A, g)
    pm2 = ParticleMesh(s.mesh, s.box, T)
    with pytest.raises(HymdError):
        field_adjoint(pm2, xd, td, T, s.sigma, A, g, reuse=True)
    moments(pm2, xd, td, T, s.sigma)
    again = field_adjoint(pm2, xd, td, T, s.sigma, A, g, reuse=True)
    np.testing.assert_array_equal(again[0], fresh[0])
    assert again[1] == fresh[1]
    zero = field_adjoint(pm2, xd, td, T, s.sigma, A, torch.zeros_like(g), reuse=True)
    assert not np.any(zero[0]) and zero[1] == 0.0


def test_energy_forces_function_vs_fd(hip):
    """L = E + sum_i w_i . F_i through FieldEnergyForces: dL/dchi, dL/dkappa
    and dL/dsigma against central differences of the FP64 oracle's L."""
    from hymd_amd.autograd import field_energy_forces
    from hymd_amd.field import ParticleMesh
    s = _system("lipid")
    T = s.n_types
    rho0 = s.n / np.prod(s.box)
    x32 = s.positions.astype(np.float32)
    x64 = x32.astype(np.float64)
    w = np.random.default_rng(3).normal(size=(s.n, 3)) * 50.0
    pm = ParticleMesh(s.mesh, s.box, T)
    x = torch.as_tensor(x32).cuda()
    chi = torch.as_tensor(s.chi, dtype=torch.float64).cuda().requires_grad_(True)
    kappa = torch.tensor(s.kappa, dtype=torch.float64, requires_grad=True)
    sigma = torch.tensor(s.sigma, dtype=torch.float64, requires_grad=True)
    types = torch.as_tensor(s.types).cuda()
    E, Fd = field_energy_forces(x, chi, kappa, sigma, pm, types, rho0, rho0)
    L = E + torch.sum(torch.as_tensor(w, dtype=torch.float32).cuda() * Fd)
    L.backward()

    def loss(chi_, kappa_, sigma_):
        ham = F.default_with_chi(T, kappa_, rho0, rho0, chi_)
        out = F.update_field(x64, s.types, T, s.mesh, s.box, sigma_, ham)
        f = F.compute_field_force(x64, s.types, out["force_mesh"], s.mesh, s.box)
        return F.field_energy(out["phi"], ham, s.mesh, s.box) + np.sum(w * f)

    L0 = loss(s.chi, s.kappa, s.sigma)
    assert float(L.detach()) == pytest.approx(L0, rel=1e-5)
    h = 1e-3
    t, u = 1, 3
    c = s.chi[t, u]
    cp, cm = s.chi.copy(), s.chi.copy()
    cp[t, u] = cp[u, t] = c * (1 + h)
    cm[t, u] = cm[u, t] = c * (1 - h)
    fd = (loss(cp, s.kappa, s.sigma) - loss(cm, s.kappa, s.sigma)) / (2 * c * h)
    gchi = chi.grad.cpu().numpy()
    assert gchi[t, u] + gchi[u, t] == pytest.approx(fd, rel=1e-4)
    k = s.kappa
    fd = (loss(s.chi, k * (1 + h), s.sigma) - loss(s.chi, k * (1 - h), s.sigma)) / (2 * k * h)
    assert float(kappa.grad) == pytest.approx(fd, rel=1e-4)
    sg = s.sigma
    fd = (loss(s.chi, s.kappa, sg * (1 + h)) - loss(s.chi, s.kappa, sg * (1 - h))) / (2 * sg * h)
    assert float(sigma.grad) == pytest.approx(fd, rel=1e-4)
