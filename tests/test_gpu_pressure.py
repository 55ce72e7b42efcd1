"""Device pressure (hymd_amd.pressure) against the FP64 oracle
(oracle/pressure.py) on the same positions: the field terms p0, p1, p2_xyz
from the FP32 spectra of update_field, by Parseval sums, against the
oracle's real-space sums over its Laplacian grids.  Tolerance: 2e-5 of the
largest field term (FP32 paint / FFT / filter against FP64)."""
import numpy as np
import pytest
import torch

from oracle import field as OF
from oracle import pressure as OP

pytestmark = pytest.mark.gpu


def _cfg(T, mesh, box, sigma, **kw):
    from hymd_amd.config import Config
    cfg = Config(n_steps=0, time_step=0.03, mesh_size=list(mesh), sigma=sigma, kappa=0.05, box_size=box,
                 n_types=T, **kw)
    return cfg


def _device_fields(x, types, T, mesh, box, sigma, ham_q):
    """update_field through the drop-in: (pm, phi list, phi_fourier list)."""
    from hymd_amd.field import initialize_pm, update_field
    cfg = _cfg(T, mesh, box, sigma)
    pm, fl, _, _ = initialize_pm(None, cfg)
    phi, phi_fourier, force_on_grid, v_ext_fourier, v_ext, phi_transfer, phi_laplacian = fl

    class Ham:
        def quadratic(self):
            return ham_q.A, ham_q.b, ham_q.c0

    update_field(phi, phi_laplacian, phi_transfer, None, force_on_grid, Ham(), pm,
                 torch.as_tensor(x, dtype=torch.float32).cuda(), torch.as_tensor(types, dtype=torch.int32).cuda(),
                 cfg, v_ext, phi_fourier, v_ext_fourier, None)
    return pm, cfg, Ham(), fl


CASES = [
    # mesh, box, T, sigma, n
    ((32, 24, 40), (5.0, 4.0, 6.0), 3, 0.5, 20_000),
    ((32, 24, 40), (5.0, 4.0, 6.0), 3, 0.12, 20_000),   # the Nyquist planes carry weight
    ((48, 48, 48), (6.0, 6.0, 6.0), 2, 0.45, 50_000),
    ((16, 20, 24), (3.0, 3.5, 4.0), 5, 0.3, 8_000),
]


@pytest.mark.parametrize("mesh,box,T,sigma,n", CASES)
def test_field_pressure_matches_oracle(hip, mesh, box, T, sigma, n):
    from hymd_amd.pressure import comp_pressure
    print(f"pressure mesh={mesh} T={T} sigma={sigma}", flush=True)
    rng = np.random.default_rng(n + T)
    box = np.array(box)
    x = rng.uniform(0.0, 1.0, (n, 3)) * box
    types = rng.integers(0, T, n)
    rho0 = n / np.prod(box)
    chi = np.triu(rng.uniform(-5.0, 25.0, (T, T)), 1)
    ham = OF.default_with_chi(T, 0.05, rho0, rho0, chi + chi.T)
    pm, cfg, H, fl = _device_fields(x, types, T, mesh, box, sigma, ham)
    v = rng.normal(size=(n, 3))
    bp, ap = rng.normal(size=3), rng.normal(size=3)
    got = comp_pressure(fl[0], None, None, H, torch.as_tensor(v).cuda(), cfg, fl[1], fl[6], fl[5], None, bp, ap)
    ref = OF.update_field(x, types, T, mesh, box, sigma, ham)
    want = OP.comp_pressure(ref["phi"], ref["phi_fourier"], ham, v, cfg.mass, mesh, box, sigma, bp, ap)
    scale = max(abs(want[1]), abs(want[2]))
    assert np.max(np.abs(got[1:6] - want[1:6])) < 2e-5 * scale, (got[1:6], want[1:6])
    np.testing.assert_allclose(got[0], want[0], rtol=1e-12)
    np.testing.assert_allclose(got[6:15], want[6:15], rtol=1e-14, atol=0)
    np.testing.assert_allclose(got[15:], want[15:], rtol=0, atol=2e-5 * scale)


def test_electrostatic_terms_match_oracle(hip):
    """psi and phi_q of the PME path enter V_bar (type charges) and w_elec."""
    from hymd_amd.field import update_field_force_q
    from hymd_amd.pressure import comp_pressure
    T, mesh, box, sigma, n = 2, (32, 32, 32), np.array([5.0, 5.0, 5.0]), 0.5, 20_000
    rng = np.random.default_rng(5)
    x = rng.uniform(0.0, 1.0, (n, 3)) * box
    types = rng.integers(0, T, n)
    qt = np.array([1.0, -1.0])
    q = qt[types]
    rho0 = n / np.prod(box)
    ham = OF.default_with_chi(T, 0.05, rho0, rho0, np.array([[0, 8.0], [8.0, 0]]))
    pm, cfg, H, fl = _device_fields(x, types, T, mesh, box, sigma, ham)
    cfg.coulombtype = "PIC_Spectral"
    cfg.dielectric_const = 45.0
    cfg.type_charges = list(qt)
    e_self = OF.self_energy_q(q, sigma, cfg.coulomb_constant, cfg.dielectric_const)
    cfg.self_energy = e_self
    phi_q = pm.create("real")
    psi = pm.create("real")
    ef = pm.create("real", batch=3)
    forces = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    update_field_force_q(torch.as_tensor(q, dtype=torch.float32).cuda(), phi_q, None, psi, None, None,
                         [ef[d] for d in range(3)], forces, None, H, pm,
                         torch.as_tensor(x, dtype=torch.float32).cuda(), cfg)
    v = np.zeros((1, 3))
    got = comp_pressure(fl[0], phi_q, psi, H, torch.as_tensor(v).cuda(), cfg, fl[1], fl[6], fl[5], None,
                        np.zeros(3), np.zeros(3))
    ref = OF.update_field(x, types, T, mesh, box, sigma, ham)
    rq = OF.update_field_force_q(x, q, mesh, box, sigma, cfg.coulomb_constant, cfg.dielectric_const)
    want = OP.comp_pressure(ref["phi"], ref["phi_fourier"], ham, v, cfg.mass, mesh, box, sigma, np.zeros(3),
                            np.zeros(3), phi_q=rq["phi_q"], psi=rq["psi"], type_charges=qt, self_energy=e_self)
    scale = max(abs(want[1]), abs(want[2]))
    assert np.max(np.abs(got[1:6] - want[1:6])) < 2e-5 * scale, (got[1:6], want[1:6])


def test_lipid_water_pressure_at_full_size(hip):
    """1,048,576 beads / 128^3: the device field terms against the oracle on
    the same positions."""
    from hymd_amd import synth
    from hymd_amd.field import ParticleMesh
    from hymd_amd.pressure import field_pressure
    s = synth.lipid_water(seed=1001)
    A, b, c0 = synth.hamiltonian_form(s)
    T = s.n_types
    pm = ParticleMesh(s.mesh, s.box, T)
    x32 = torch.as_tensor(np.mod(s.positions, s.box), dtype=torch.float32).cuda()
    ty = torch.as_tensor(s.types, dtype=torch.int32).cuda()
    fm = pm.create("real", batch=(T, 3))
    pf = pm.create("complex", batch=T)
    pm.field_update(x32, ty, T, A, b, c0, [1.0] * T, s.sigma, fm, phi_fourier=pf)
    got = field_pressure(pm, pf, T, A, b, c0, s.sigma)
    ham = OF.Quadratic(A, b, c0)
    ref = OF.update_field(x32.cpu().numpy().astype(np.float64), s.types, T, s.mesh, s.box, s.sigma, ham)
    want = OP.comp_pressure(ref["phi"], ref["phi_fourier"], ham, np.zeros((1, 3)), s.mass, s.mesh, s.box, s.sigma,
                            np.zeros(3), np.zeros(3))[1:6]
    print("pressure field terms device", got, "oracle", want, flush=True)
    scale = max(abs(want[0]), abs(want[1]))
    assert np.max(np.abs(got - want)) < 2e-5 * scale
    pm.close()
