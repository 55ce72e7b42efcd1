"""HIP particle-field path against the FP64 oracle (and closed forms).

Tolerances: the int64 fixed-point density grid is compared bit for bit; FP32
FFT / forces against the FP64 oracle within the north-star bound of 1e-4
relative RMS (tighter where the data allow, stated per test).
"""
import numpy as np
import pytest
import torch

from oracle import field as F
from tests.conftest import rel_err, rms_rel

pytestmark = pytest.mark.gpu


def _pm(mesh, box, T):
    from hymd_amd.field import ParticleMesh
    return ParticleMesh(mesh, box, T)


def _dev(a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda().contiguous()


# the row paint owns 4 mesh rows along y per workgroup (fewer in the last
# block when My % 4 != 0; one when My < 4 cannot happen, edges being >= 4):
# My = 4, 6, 10 cover the wrap of the source rows and the partial blocks
@pytest.mark.parametrize("mesh,box", [((32, 48, 40), (7.0, 9.5, 8.25)), ((16, 16, 16), (4.0, 4.0, 4.0)),
                                      ((8, 4, 12), (2.0, 1.0, 3.0)), ((6, 6, 8), (1.5, 1.5, 2.0)),
                                      ((10, 10, 6), (2.5, 2.5, 1.5))])
def test_paint_fixed_bit_exact(hip, mesh, box):
    rng = np.random.default_rng(11)
    n, T = 20000, 3
    x = rng.uniform(-0.5, 1.5, (n, 3)) * np.array(box)           # includes out-of-box positions
    x[:5] = np.array(box)                                         # exactly on the upper face
    x32 = x.astype(np.float32)
    types = rng.integers(0, T, n).astype(np.int32)
    mass = [1.0, 0.75, 2.5]
    pm = _pm(mesh, box, T)
    xd, td = _dev(x32, torch.float32), _dev(types, torch.int32)
    grid = pm.paint_fixed(xd, td, T, mass=mass).cpu().numpy()           # cell sort + gather
    scat = torch.zeros((T,) + tuple(mesh), dtype=torch.int64, device="cuda")
    pm.paint_fixed(xd, td, T, mass=mass, out=scat, accumulate=True)      # 64-bit atomics
    np.testing.assert_array_equal(grid, scat.cpu().numpy())
    for t in range(T):
        sel = types == t
        ref = F.paint_fixed(x32[sel].astype(np.float64), mesh, box, mass=mass[t])
        np.testing.assert_array_equal(grid[t], ref)
        assert int(grid[t].sum()) == int(np.rint(mass[t] * 2 ** 32)) * int(sel.sum())
        np.testing.assert_allclose(grid[t] / 2.0 ** 32, F.paint(x32[sel], mesh, box, mass=mass[t]),
                                   atol=8 * 2.0 ** -33 * mass[t] * 8)


def test_paint_empty(hip):
    pm = _pm((8, 8, 8), (2.0, 2.0, 2.0), 1)
    x = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
    t = torch.zeros(0, dtype=torch.int32, device="cuda")
    assert int(pm.paint_fixed(x, t, 1).abs().sum()) == 0


# (16, 200, 200) and (256, 256, 256): their y-z planes exceed the LDS budget,
# so the separate line passes run; the other meshes use the plane kernels, with the runtime plane
# geometry or, for y-z planes of 48^2, 64^2, 96^2, 128^2, 160^2, the
# compile-time one
@pytest.mark.parametrize("mesh", [(8, 8, 8), (32, 32, 32), (48, 64, 96), (128, 128, 128), (160, 160, 160),
                                  (20, 30, 36), (16, 200, 200), (12, 48, 48), (10, 64, 64), (6, 96, 96),
                                  (256, 256, 256)])
def test_fft_vs_numpy(hip, mesh):
    rng = np.random.default_rng(sum(mesh))
    pm = _pm(mesh, (1.0, 1.0, 1.0), 1)
    B = 2
    x = rng.normal(size=(B,) + mesh).astype(np.float32)
    X = pm.r2c(_dev(x, torch.float32)).cpu().numpy()
    ref = np.fft.rfftn(x.astype(np.float64), axes=(1, 2, 3))
    # FP32 transform error grows ~ sqrt(log G); bound relative to the spectrum's RMS
    assert np.sqrt(np.mean(np.abs(X - ref) ** 2)) / np.sqrt(np.mean(np.abs(ref) ** 2)) < 2e-6
    y = pm.c2r(_dev(ref.astype(np.complex64), torch.complex64)).cpu().numpy()
    G = np.prod(mesh)
    np.testing.assert_allclose(y / G, x, atol=2e-5 * np.abs(x).max())


# 256-point strided passes take 16 lines per workgroup for batches of 4
# grids or more and 8 below: y lines (8 x 256 x 256, line passes) and x lines
# (256 x 16 x 16, after the plane kernels), both directions, both widths
@pytest.mark.parametrize("mesh", [(8, 256, 256), (256, 16, 16)])
@pytest.mark.parametrize("B", [2, 6])
def test_fft_256_lines_batched_vs_numpy(hip, mesh, B):
    rng = np.random.default_rng(B * 1000 + mesh[0])
    pm = _pm(mesh, (1.0, 1.0, 1.0), 1)
    x = rng.normal(size=(B,) + mesh).astype(np.float32)
    X = pm.r2c(_dev(x, torch.float32)).cpu().numpy()
    ref = np.fft.rfftn(x.astype(np.float64), axes=(1, 2, 3))
    for b in range(B):
        assert np.sqrt(np.mean(np.abs(X[b] - ref[b]) ** 2)) / np.sqrt(np.mean(np.abs(ref[b]) ** 2)) < 2e-6, b
    y = pm.c2r(_dev(ref.astype(np.complex64), torch.complex64)).cpu().numpy()
    G = np.prod(mesh)
    np.testing.assert_allclose(y / G, x, atol=2e-5 * np.abs(x).max())


def _diblock(n_chains=600, box=8.0):
    from hymd_amd import synth
    return synth.diblock_melt(n_chains=n_chains, box=box, mesh=32, seed=1000)


def test_update_field_matches_oracle(hip):
    from hymd_amd.synth import hamiltonian_form
    s = _diblock()
    A, b, c0 = hamiltonian_form(s)
    ham = F.Quadratic(A, b, c0)
    x32 = s.positions.astype(np.float32)
    ref = F.update_field(x32.astype(np.float64), s.types, 2, s.mesh, s.box, s.sigma, ham, compute_potential=True)
    f_ref = F.compute_field_force(x32.astype(np.float64), s.types, ref["force_mesh"], s.mesh, s.box)
    e_ref = F.field_energy(ref["phi"], ham, s.mesh, s.box)

    pm = _pm(s.mesh, s.box, 2)
    xd, td = _dev(x32, torch.float32), _dev(s.types, torch.int32)
    fm = pm.create("real", batch=(2, 3))
    phi = pm.create("real", batch=2)
    phif = pm.create("complex", batch=2)
    vext = pm.create("real", batch=2)
    e = pm.field_update(xd, td, 2, A, b, c0, [1.0, 1.0], s.sigma, fm, phi=phi, phi_fourier=phif, v_ext=vext,
                        want_energy=True)
    f = torch.empty((s.n, 3), dtype=torch.float32, device="cuda")
    pm.readout_forces(xd, td, 2, fm, f)
    f = f.cpu().numpy()
    fmg = fm.cpu().numpy()
    for t in range(2):
        for d in range(3):
            assert rms_rel(fmg[t, d], ref["force_mesh"][t][d]) < 1e-5
        assert rms_rel(phi[t].cpu().numpy(), ref["phi"][t]) < 1e-6
        assert rms_rel(phif[t].cpu().numpy(), ref["phi_fourier"][t]) < 1e-6
        assert rms_rel(vext[t].cpu().numpy(), ref["v_ext"][t]) < 1e-6
    assert rms_rel(f, f_ref) < 1e-5
    assert rel_err(f, f_ref) < 1e-4
    assert e == pytest.approx(e_ref, rel=1e-6)
    # real-space energy reduction of the stored filtered densities
    assert pm.field_energy(phi, 2, A, b, c0) == pytest.approx(e_ref, rel=1e-6)


def test_forces_on_nodes_match_pair_forces(hip):
    """Closed form (continuum Gaussian-core pair forces) on the GPU path."""
    M, L, sigma = 64, 10.0, 0.5
    rng = np.random.default_rng(3)
    idx = rng.choice(M ** 3, 12, replace=False)
    x = np.stack(np.unravel_index(idx, (M, M, M)), axis=1) * (L / M)
    types = rng.integers(0, 2, 12).astype(np.int32)
    rho0 = 16.0
    ham = F.default_with_chi(2, 0.05, rho0, rho0, np.array([[0.0, 20.0], [20.0, 0.0]]))
    _, f_ref = F.gaussian_core_pairs(x, types, ham.A, sigma, np.full(3, L))
    pm = _pm((M,) * 3, (L,) * 3, 2)
    xd, td = _dev(x.astype(np.float32), torch.float32), _dev(types, torch.int32)
    fm = pm.create("real", batch=(2, 3))
    pm.field_update(xd, td, 2, ham.A, ham.b, ham.c0, [1.0, 1.0], sigma, fm)
    f = torch.empty((12, 3), dtype=torch.float32, device="cuda")
    pm.readout_forces(xd, td, 2, fm, f)
    np.testing.assert_allclose(f.cpu().numpy(), f_ref, atol=2e-5 * np.abs(f_ref).max())


@pytest.mark.parametrize("case", [0, 1])
def test_gaussian_core_energy_160(hip, goldens, case):
    """160^3 mesh (radix 5), closed-form energy of test/test_hamiltonian.py:172-227."""
    g = goldens["gaussian_core_no_chi"]
    kappa, sigma = g["params"][case]
    r = np.array(g["positions"])
    box = np.array(g["box"])
    N = len(r)
    rho0 = N / np.prod(box)
    ham = F.default_no_chi(1, kappa, rho0, rho0)
    c = 16.0 * np.pi ** 1.5 * kappa * sigma ** 3 * rho0
    E = N / c - 0.5 * N / kappa + sum(2 * np.exp(-np.sum((r[i] - r[j]) ** 2) / (4 * sigma ** 2)) / c
                                      for i in range(N) for j in range(i + 1, N))
    pm = _pm((160,) * 3, box, 1)
    xd = _dev(r.astype(np.float32), torch.float32)
    td = torch.zeros(N, dtype=torch.int32, device="cuda")
    fm = pm.create("real", batch=(1, 3))
    phi = pm.create("real", batch=1)
    e = pm.field_update(xd, td, 1, ham.A, ham.b, ham.c0, [1.0], sigma, fm, phi=phi, want_energy=True)
    assert e == pytest.approx(E, rel=2e-6)
    assert pm.field_energy(phi, 1, ham.A, ham.b, ham.c0) == pytest.approx(E, rel=2e-6)


def test_pme_matches_oracle(hip):
    rng = np.random.default_rng(7)
    n, M, L = 4000, 40, 10.0
    x = rng.uniform(0, L, (n, 3)).astype(np.float32)
    q = np.zeros(n, dtype=np.float32)
    q[: n // 10] = 1.0
    q[n // 10: n // 5] = -1.0
    ke, eps, sigma = 138.935458, 45.0, 0.5
    ref = F.update_field_force_q(x.astype(np.float64), q, M, np.full(3, L), sigma, ke, eps)
    e_self = F.self_energy_q(q, sigma, ke, eps)
    e_ref = F.field_q_energy(ref["phi_q"], ref["psi"], e_self, M, np.full(3, L))
    pm = _pm((M,) * 3, (L,) * 3, 1)
    xd, qd = _dev(x, torch.float32), _dev(q, torch.float32)
    forces = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    phi_q = pm.create("real")
    psi = pm.create("real")
    ef = pm.create("real", batch=3)
    e = pm.pme(xd, qd, sigma, ke / eps, e_self, forces, phi_q=phi_q, psi=psi, elec_field=ef, want_energy=True)
    f = forces.cpu().numpy()
    charged = q != 0
    assert rms_rel(f[charged], ref["forces"][charged]) < 1e-5
    assert np.abs(f[~charged]).max() == 0.0
    assert rms_rel(phi_q.cpu().numpy(), ref["phi_q"]) < 1e-6
    assert rms_rel(psi.cpu().numpy(), ref["psi"]) < 1e-5
    for d in range(3):
        assert rms_rel(ef[d].cpu().numpy(), ref["elec_field"][d]) < 1e-5
    assert e == pytest.approx(e_ref, rel=1e-5)
    assert pm.field_q_energy(phi_q, psi, e_self) == pytest.approx(e_ref, rel=1e-5)


def test_drop_in_field_surface(hip):
    """initialize_pm / update_field / compute_field_force /
    compute_field_and_kinetic_energy with the reference's call signatures."""
    from hymd_amd import field as HF
    from hymd_amd.config import Chi, Config
    from hymd_amd.hamiltonian import DefaultWithChi
    s = _diblock(n_chains=300, box=6.0)
    cfg = Config(n_steps=1, time_step=0.03, mesh_size=32, sigma=0.5, kappa=0.05, box_size=s.box,
                 n_particles=s.n, n_types=2, unique_names=["A", "B"], type_to_name_map={0: "A", 1: "B"},
                 chi=[Chi("A", "B", 15.0)])
    ham = DefaultWithChi(cfg, cfg.unique_names, cfg.type_to_name_map)
    pm, field_list, _, _ = HF.initialize_pm(None, cfg)
    phi, phi_fourier, force_on_grid, v_ext_fourier, v_ext, phi_transfer, phi_laplacian = field_list
    x = _dev(s.positions.astype(np.float32), torch.float32)
    types = _dev(s.types, torch.int32)
    HF.update_field(phi, phi_laplacian, phi_transfer, None, force_on_grid, ham, pm, x, types, cfg, v_ext,
                    phi_fourier, v_ext_fourier, cfg.m, compute_potential=True)
    forces = torch.zeros((s.n, 3), dtype=torch.float32, device="cuda")
    HF.compute_field_force(None, x, force_on_grid, forces, types, 2)
    v = torch.as_tensor(s.velocities).cuda()
    e_field, e_kin, e_q = HF.compute_field_and_kinetic_energy(phi, None, None, v, ham, x, types, v_ext, cfg, None)

    A, b, c0 = ham.quadratic()
    oham = F.Quadratic(A, b, c0)
    xr = s.positions.astype(np.float32).astype(np.float64)
    ref = F.update_field(xr, s.types, 2, s.mesh, s.box, 0.5, oham)
    f_ref = F.compute_field_force(xr, s.types, ref["force_mesh"], s.mesh, s.box)
    assert rms_rel(forces.cpu().numpy(), f_ref) < 1e-5
    assert e_field == pytest.approx(F.field_energy(ref["phi"], oham, s.mesh, s.box), rel=1e-6)
    assert pm.last_field_energy == pytest.approx(e_field, rel=1e-6)
    assert e_kin == pytest.approx(F.kinetic_energy(s.velocities, cfg.mass), rel=1e-12)
    assert e_q == 0.0
    for t in range(2):
        assert rms_rel(phi[t].cpu().numpy(), ref["phi"][t]) < 1e-6


@pytest.mark.parametrize("mesh,T", [((32, 48, 40), 3), ((24, 24, 24), 2), ((16, 20, 12), 5), ((48, 40, 64), 4),
                                    ((128, 96, 80), 6), ((64, 64, 64), 1), ((40, 30, 18), 8),
                                    # compile-time x-line (Mx 128, 160) and plane (128^2, 160^2) geometries
                                    ((128, 40, 24), 4), ((128, 24, 20), 2), ((128, 20, 16), 3), ((128, 16, 12), 1),
                                    ((160, 20, 16), 3), ((160, 16, 24), 1),
                                    ((16, 128, 128), 4), ((8, 160, 160), 2),
                                    # long x lines with many types: 8 lines per workgroup of 1024 threads
                                    ((192, 16, 16), 6), ((256, 12, 10), 5), ((512, 8, 8), 4), ((160, 24, 20), 7)])
def test_fused_xline_pass_equals_unfused(hip, mesh, T):
    """The force-grid-only update (forward x, k-space and inverse x passes in
    one kernel, -i k_y and -i k_z applied in the y-z plane pass) against the
    unfused passes (-i k_d V_t formed in k-space, then three full inverse
    transforms per type), which run when v_ext is requested.  The same
    operator, applied in a different order, so the grids agree to FP32
    rounding, not bit for bit."""
    rng = np.random.default_rng(5)
    box = (6.0, 7.5, 5.0)
    n = 6000
    x = (rng.uniform(0, 1, (n, 3)) * np.array(box)).astype(np.float32)
    types = rng.integers(0, T, n).astype(np.int32)
    A = rng.uniform(0.0, 2.0, (T, T))
    A = A + A.T
    b = rng.uniform(-1.0, 0.0, T)
    pm = _pm(mesh, box, T)
    xd, td = _dev(x, torch.float32), _dev(types, torch.int32)
    f1 = pm.create("real", batch=(T, 3))
    f2 = pm.create("real", batch=(T, 3))
    v = pm.create("real", batch=T)
    e1 = pm.field_update(xd, td, T, A, b, 0.3, [1.0] * T, 0.45, f1, want_energy=True)
    e2 = pm.field_update(xd, td, T, A, b, 0.3, [1.0] * T, 0.45, f2, v_ext=v, want_energy=True)
    a1, a2 = f1.cpu().numpy(), f2.cpu().numpy()
    assert np.abs(a1 - a2).max() <= 1e-6 * np.abs(a2).max()
    assert rms_rel(a1, a2) < 2e-7
    assert e1 == pytest.approx(e2, rel=1e-9)


@pytest.mark.parametrize("mesh,T", [((20, 24, 18), 3), ((16, 16, 16), 2), ((12, 10, 8), 5), ((24, 20, 30), 8),
                                    ((6, 6, 144), 8), ((12, 256, 256), 6), ((8, 200, 256), 4)])
def test_readout_matches_oracle(hip, mesh, T):
    """compute_field_force on arbitrary grids against the oracle's CIC readout
    (field.py:317-330), with types out of range giving zero, on meshes up to
    256 x 256 y-z planes (where two x-planes of the 3T force grids exceed an
    XCD's 4 MB L2)."""
    rng = np.random.default_rng(sum(mesh) + T)
    box = (3.0, 4.5, 2.5)
    n = 5000
    x = (rng.uniform(-0.3, 1.3, (n, 3)) * np.array(box)).astype(np.float32)   # some outside the box
    types = rng.integers(0, T, n).astype(np.int32)
    types[::97] = T + 1                                                        # out of range: zero force
    grids = rng.normal(size=(T, 3) + mesh).astype(np.float32)
    pm = _pm(mesh, box, T)
    f = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    pm.readout_forces(_dev(x, torch.float32), _dev(types, torch.int32), T, _dev(grids, torch.float32), f)
    xr = np.mod(x.astype(np.float64), box)
    ok = types < T
    ref = np.zeros((n, 3))
    ref[ok] = F.compute_field_force(xr[ok], types[ok], [[grids[t, d].astype(np.float64) for d in range(3)]
                                                         for t in range(T)], mesh, box)
    got = f.cpu().numpy()
    np.testing.assert_array_equal(got[~ok], 0.0)
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-6 * np.abs(ref).max())


@pytest.mark.parametrize("mesh,T", [((32, 32, 32), 2), ((24, 20, 30), 5), ((128, 128, 128), 4),
                                    ((128, 24, 24), 6),    # x-line widths on 256, 512 and 1024 threads
                                    ((256, 16, 16), 6)])   # compile-time 256-point lines (8 of 1024, 4 of 512 threads)
def test_tuning_knobs_keep_the_bits(hip, mesh, T, tmp_path):
    """The paint and readout launch shapes the autotuner may choose give the
    same density, force grids, forces and energy bit for bit (integer paint
    sums, unchanged per-element arithmetic); another x-line width may change
    the FP32 rounding of the x transform (1e-6).  The autotuner runs, caches
    its choice and restores it from the cache."""
    from hymd_amd import tune
    rng = np.random.default_rng(sum(mesh) + T)
    box = np.array(mesh, dtype=np.float64) * 0.31
    n = 30_000 if mesh[0] * mesh[1] < 10_000 else 400_000
    x = (rng.uniform(0, 1, (n, 3)) * box).astype(np.float32)
    types = rng.integers(0, T, n).astype(np.int32)
    A = rng.uniform(0.0, 1.0, (T, T))
    A = A + A.T
    b = rng.uniform(-1, 1, T)
    pm = _pm(mesh, box, T)
    xd, td = _dev(x, torch.float32), _dev(types, torch.int32)

    def run():
        fm = pm.create("real", batch=(T, 3))
        e = pm.field_update(xd, td, T, A, b, 0.0, [1.0] * T, 0.5, fm, want_energy=True)
        f = torch.empty((n, 3), dtype=torch.float32, device="cuda")
        pm.readout_forces(xd, td, T, fm, f)
        grid = pm.paint_fixed(xd, td, T)
        return fm.cpu(), f.cpu(), grid.cpu(), e

    ref = run()
    for name, (_, values) in tune.KNOBS.items():
        for v in values:
            tune.apply(pm, dict(tune.DEFAULTS, **{name: v}))
            assert tune.get(pm)[name] == v
            got = run()
            assert torch.equal(got[2], ref[2])   # the integer density grid
            if name == "xline_width":
                # another x-line geometry may round the x transform differently
                for a, c in zip(got[:2], ref[:2]):
                    assert rms_rel(a.numpy(), c.numpy()) < 1e-6, (name, v)
                assert got[3] == pytest.approx(ref[3], rel=1e-6)
            else:
                for a, c in zip(got[:2], ref[:2]):
                    assert torch.equal(a, c), (name, v)
                assert got[3] == ref[3]
    path = str(tmp_path / "tune.json")
    res = tune.autotune(pm, xd, td, T, A, b, 0.0, 0.5, reps=3, path=path)
    assert not res["cached"] and res["tuned_ms"] <= res["default_ms"]
    tune.apply(pm, tune.DEFAULTS)
    again = tune.autotune(pm, xd, td, T, A, b, 0.0, 0.5, reps=3, path=path)
    assert again["cached"] and tune.get(pm) == res["choice"]
    got = run()
    assert rms_rel(got[1].numpy(), ref[1].numpy()) < 1e-6


@pytest.mark.parametrize("T", [1, 3, 6])
@pytest.mark.parametrize("mesh,box", [((16, 200, 200), (3.2, 40.0, 40.0)), ((8, 256, 256), (1.6, 51.2, 51.2))])
def test_line_gradient_path_matches_oracle(hip, T, mesh, box):
    """Meshes whose y-z plane does not fit in LDS (16 x 200 x 200 and
    8 x 256 x 256 here, like 256^3) take the line-pass gradient path: the
    fused x-line pass, the out-of-place y pass with -i k_y on load (for 256-
    point y lines the compile-time kernel whose first and last stages read
    and write global memory) and the z pass with -i k_z on load.  Field
    forces, force grids and energy against the FP64 oracle; PME forces and
    energy against the oracle on the same mesh."""
    from hymd_amd.md import COULOMB_CONSTANT
    rng = np.random.default_rng(40 + T)
    n = 60_000
    x = (rng.uniform(0, 1, (n, 3)) * np.asarray(box)).astype(np.float32)
    types = rng.integers(0, T, n).astype(np.int32)
    A = rng.uniform(0.0, 1.0, (T, T))
    A = A + A.T + T * np.eye(T)
    b = rng.uniform(-1, 1, T)
    ham = F.Quadratic(A, b, 0.3)
    x64 = x.astype(np.float64)
    ref = F.update_field(x64, types, T, mesh, box, 0.5, ham)
    f_ref = F.compute_field_force(x64, types, ref["force_mesh"], mesh, box)
    pm = _pm(mesh, box, T)
    xd, td = _dev(x, torch.float32), _dev(types, torch.int32)
    fm = pm.create("real", batch=(T, 3))
    e = pm.field_update(xd, td, T, A, b, 0.3, [1.0] * T, 0.5, fm, want_energy=True)
    f = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    pm.readout_forces(xd, td, T, fm, f)
    fmg = fm.cpu().numpy()
    for t in range(T):
        for d in range(3):
            assert rms_rel(fmg[t, d], ref["force_mesh"][t][d]) < 1e-5, (t, d)
    assert rms_rel(f.cpu().numpy(), f_ref) < 1e-5
    assert e == pytest.approx(F.field_energy(ref["phi"], ham, mesh, box), rel=1e-6)
    q = np.zeros(n, dtype=np.float32)
    q[: n // 10] = 1.0
    q[n // 10: n // 5] = -1.0
    rq = F.update_field_force_q(x64, q.astype(np.float64), mesh, box, 0.5, COULOMB_CONSTANT, 45.0)
    fq = torch.empty_like(f)
    psi = pm.create("real")
    eq = pm.pme(xd, _dev(q, torch.float32), 0.5, COULOMB_CONSTANT / 45.0, 0.0, fq, psi=psi, want_energy=True)
    assert rms_rel(fq.cpu().numpy(), rq["forces"]) < 1e-5
    assert rms_rel(psi.cpu().numpy(), rq["psi"]) < 1e-5
    assert eq == pytest.approx(F.field_q_energy(rq["phi_q"], rq["psi"], 0.0, mesh, box), rel=1e-5)


@pytest.mark.parametrize("conv", ["max_diff", "csum", "euclidean_norm"])
def test_gpe_matches_oracle(hip, conv):
    """update_field_force_q_GPE on the device (through the reference's call
    surface: initialize_pm's grid set, update_field's filtered densities)
    against the FP64 oracle: dielectric, potential, |E|^2, V_t and forces to
    FP32 rounding of the iteration (1e-4 relative RMS), the same iteration
    count to within one, the energy to 1e-4; the masked cells keep their
    previous values."""
    from hymd_amd import config as C
    from hymd_amd import field as HF
    from oracle import gpe as OG
    rng = np.random.default_rng(17)
    M, L, T = (24, 24, 24), np.array([6.0, 6.0, 6.0]), 3
    n = 9000
    x = rng.uniform(0, 1, (n, 3)) * L
    types = np.where(x[:, 0] < 2.0, 0, np.where(x[:, 0] < 4.0, 1, 2)).astype(np.int32)
    q = np.zeros(n, dtype=np.float32)
    q[:450], q[450:900] = 1.0, -1.0
    cfg = C.Config(n_steps=1, time_step=0.03, mesh_size=M, sigma=0.6, kappa=0.05, box_size=L, n_types=T,
                   coulombtype="PIC_Spectral_GPE", dielectric_type=[4.0, 30.0, 80.0], type_charges=[0.5, 0.0, -0.5],
                   conv_crit={"max_diff": 1e-6, "csum": 1e-3, "euclidean_norm": 1e-9}[conv], convergence_type=conv)
    pm, fields, elec, coul = HF.initialize_pm(None, cfg)
    x32 = x.astype(np.float32)
    ham = F.Quadratic(np.eye(T), np.zeros(T), 0.0)
    A, b, c0 = np.eye(T), np.zeros(T), 0.0
    fm = pm.create("real", batch=(T, 3))
    ph = pm.create("real", batch=T)
    pm.field_update(_dev(x32, torch.float32), _dev(types, torch.int32), T, A, b, c0, [1.0] * T, cfg.sigma, fm, phi=ph)
    phi_q, phi_q_f, psi, efield = elec
    (phi_eps, phi_eps_f, phi_eta, phi_eta_f, phi_pol, phi_pol_prev, elec_dot, contrib, vbar, vbar_f, fme,
     fme_f) = coul
    phi_eps.fill_(7.0)
    contrib.fill_(-3.0)
    forces = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    vb, pe, ed = HF.update_field_force_q_GPE(None, [ph[t] for t in range(T)], _dev(types, torch.int32),
                                             _dev(q, torch.float32), phi_q, phi_q_f, phi_eps, phi_eps_f, phi_eta,
                                             phi_eta_f, phi_pol_prev, phi_pol, efield, forces, contrib, psi, vbar,
                                             vbar_f, fme, fme_f, None, None, None, pm, _dev(x32, torch.float32), cfg)
    e_dev = HF.compute_field_energy_q_GPE(cfg, pe, 0.0, ed)
    x64 = x32.astype(np.float64)
    ref_phi = F.update_field(x64, types, T, M, L, cfg.sigma, ham)["phi"]
    ref = OG.update_field_force_q_gpe(x64, types, q.astype(np.float64), ref_phi, M, L, cfg.sigma, cfg.dielectric_type,
                                      cfg.type_charges, cfg.coulomb_constant, conv_crit=cfg.conv_crit, conv=conv,
                                      phi_eps=np.full(M, 7.0), elec_field_contrib=np.full(M, -3.0))
    assert abs(pm.gpe_iterations - ref["iterations"]) <= 1
    assert rms_rel(pe.cpu().numpy(), ref["phi_eps"]) < 1e-6
    assert rms_rel(psi.cpu().numpy(), ref["psi"]) < 1e-4
    assert rms_rel(ed.cpu().numpy(), ref["elec_dot"]) < 1e-4
    assert rms_rel(phi_pol.cpu().numpy(), ref["phi_pol"]) < 1e-3
    for t in range(T):
        assert rms_rel(vb[t].cpu().numpy(), ref["Vbar"][t]) < 1e-4
    assert rms_rel(forces.cpu().numpy(), ref["forces"]) < 1e-4
    assert e_dev == pytest.approx(ref["energy"], rel=1e-4)
    assert pm.last_field_q_energy == pytest.approx(ref["energy"], rel=1e-4)
    assert torch.all(phi_pol_prev == 0)     # read, not changed
