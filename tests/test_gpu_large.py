"""The large Markov-typed melt: 16,777,216 beads, 100 nm box, 256^3 mesh,
6 types, 5 % charges with PME on the same mesh.

At that size the FP64 oracle is out of reach, so the device path is checked
there by what the mathematics fixes regardless of size:
* the int64 density grid holds exactly round(2^32) per particle of each type;
* the field forces and the PME forces sum to zero (the spectral derivative is
  antisymmetric and the coupling symmetric);
* a 90 degree rotation of the system about z (x -> L - y, y -> x on a cubic
  box, positions on a 2^-17 nm lattice so the map is exact in float32) maps
  the field and PME forces by the same rotation.
The same generator at a size the oracle finishes in seconds is compared with
the oracle (forces, energies, and two MD steps of the fused inner loop with
PME slow forces).
"""
import numpy as np
import pytest
import torch

from tests.conftest import rel_err, rms_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big():
    from hymd_amd import synth
    s = synth.markov_melt()
    # positions on a 2^-17 nm lattice: L - y is exact in float32 for all y
    q = 2.0 ** -17
    s.positions = np.mod(np.round(s.positions / q) * q, s.box)
    return s


def _pm(s):
    from hymd_amd.field import ParticleMesh
    return ParticleMesh(s.mesh, s.box, s.n_types)


def _forces(s, pm, x32):
    from hymd_amd.md import COULOMB_CONSTANT
    from hymd_amd.synth import hamiltonian_form
    A, b, c0 = hamiltonian_form(s)
    T = s.n_types
    xd = torch.as_tensor(x32).cuda()
    td = torch.as_tensor(s.types).cuda()
    qd = torch.as_tensor(s.charges).cuda()
    fm = pm.create("real", batch=(T, 3))
    e = pm.field_update(xd, td, T, A, b, c0, [1.0] * T, s.sigma, fm, want_energy=True)
    f = torch.empty((s.n, 3), dtype=torch.float32, device="cuda")
    pm.readout_forces(xd, td, T, fm, f)
    fq = torch.empty_like(f)
    elec = COULOMB_CONSTANT / s.dielectric
    eq = pm.pme(xd, qd, s.sigma, elec, 0.0, fq, want_energy=True)
    del fm
    return f.cpu().numpy().astype(np.float64), fq.cpu().numpy().astype(np.float64), e, eq


def test_full_size_mass_and_momentum(hip, big):
    from hymd_amd import _lib
    s = big
    T = s.n_types
    assert s.n == 16_777_216 and s.mesh == (256, 256, 256) and T == 6
    pm = _pm(s)
    x32 = s.positions.astype(np.float32)
    xd = torch.as_tensor(x32).cuda()
    td = torch.as_tensor(s.types).cuda()
    grid = pm.paint_fixed(xd, td, T)
    sums = grid.sum(dim=(1, 2, 3)).cpu().numpy()
    np.testing.assert_array_equal(sums, np.bincount(s.types, minlength=T).astype(np.int64) << _lib.HYMD_FIXED_SHIFT)
    del grid
    f, fq, e, eq = _forces(s, pm, x32)
    for F_ in (f, fq):
        assert np.isfinite(F_).all()
        assert np.abs(F_.sum(axis=0)).max() < 1e-6 * np.abs(F_).sum(axis=0).max()
    assert np.isfinite(e) and np.isfinite(eq)


def test_full_size_rotation(hip, big):
    s = big
    L = float(s.box[0])
    pm = _pm(s)
    x32 = s.positions.astype(np.float32)
    xr = np.empty_like(x32)
    xr[:, 0] = np.float32(L) - x32[:, 1]
    xr[:, 0][xr[:, 0] == np.float32(L)] = 0.0
    xr[:, 1] = x32[:, 0]
    xr[:, 2] = x32[:, 2]
    # the map is exact: x' = (L - y) mod L in float64 equals its float32 value
    np.testing.assert_array_equal(xr[:, 0].astype(np.float64), np.mod(L - x32[:, 1].astype(np.float64), L))
    f, fq, e, eq = _forces(s, pm, x32)
    g, gq, er, eqr = _forces(s, pm, xr)
    for a, b in ((f, g), (fq, gq)):
        rot = np.stack([-a[:, 1], a[:, 0], a[:, 2]], axis=1)
        assert rms_rel(b, rot) < 1e-5
        assert rel_err(b, rot) < 1e-4
    assert er == pytest.approx(e, rel=1e-6)
    assert eqr == pytest.approx(eq, rel=1e-6)


def _small():
    from hymd_amd import synth
    n_chains = 1200
    L = 100.0 * (n_chains / 1048576) ** (1.0 / 3.0)
    return synth.markov_melt(n_chains=n_chains, box=L, mesh=24, seed=1004)


def test_small_markov_melt_against_oracle(hip, oracle_built):
    """The same generator at 19,200 beads / 24^3: field and PME forces and
    energies against the FP64 oracle; two outer steps of the library's
    captured step (fused inner loop, PME in the slow forces) against the
    oracle loop."""
    from hymd_amd import synth
    from hymd_amd.md import COULOMB_CONSTANT, RespaMD
    from oracle import field as F
    from oracle import md as OM
    s = _small()
    A, b, c0 = synth.hamiltonian_form(s)
    ham = F.Quadratic(A, b, c0)
    x64 = s.positions.astype(np.float32).astype(np.float64)
    ref = F.update_field(x64, s.types, s.n_types, s.mesh, s.box, s.sigma, ham)
    f_ref = F.compute_field_force(x64, s.types, ref["force_mesh"], s.mesh, s.box)
    q_ref = F.update_field_force_q(x64, s.charges.astype(np.float64), s.mesh, s.box, s.sigma, COULOMB_CONSTANT,
                                   s.dielectric)
    f, fq, e, eq = _forces(s, _pm(s), s.positions.astype(np.float32))
    assert rms_rel(f, f_ref) < 1e-5 and rel_err(f, f_ref) < 1e-4
    assert rms_rel(fq, q_ref["forces"]) < 1e-5 and rel_err(fq, q_ref["forces"]) < 1e-4
    assert e == pytest.approx(F.field_energy(ref["phi"], ham, s.mesh, s.box), rel=1e-6)

    st = OM.State(s, ham)
    md = RespaMD(s)
    assert md.fused and md.charges is not None
    io = md.in_input_order
    assert rms_rel(io(md.f_slow).cpu().numpy(), st.f_field) < 1e-5
    md.run(2, use_graph=True)
    st.outer_step()
    st.outer_step()
    x = io(md.x).cpu().numpy()
    dx = x - st.x
    dx -= s.box * np.round(dx / s.box)
    assert np.abs(dx).max() < 1e-5
    assert rel_err(io(md.v).cpu().numpy(), st.v) < 1e-4
    en = md.run(1, energies=True)
    st.outer_step()
    assert en["kinetic"] == pytest.approx(st.kinetic_energy(), rel=1e-6)
    assert en["field"] == pytest.approx(st.field_energy(), rel=1e-5)


@pytest.mark.parametrize("use_graph", [False, True])
def test_fused_pme_step_equals_python_step(hip, use_graph):
    """With charges and no dihedrals the library keeps the fused inner loop
    and adds the PME forces to the slow forces; its steps equal the Python
    step with the fused kernel's inner loop to FP32 term rounding."""
    from hymd_amd.md import RespaMD
    a = RespaMD(_small(), term_precision="fp64")
    b = RespaMD(_small(), term_precision="fp64")
    assert b.fused
    for _ in range(2):
        a.step()
    b.run(2, use_graph=use_graph)
    torch.cuda.synchronize()
    assert (a.x - b.x).abs().max().item() < 1e-9
    assert rel_err(b.f_slow.cpu().numpy(), a.f_slow.cpu().numpy()) < 1e-5
    ea = a.step(energies=True)
    eb = b.run(1, use_graph=use_graph, energies=True)
    assert eb["field_q"] == pytest.approx(ea["field_q"], rel=1e-6)
    assert eb["field"] == pytest.approx(ea["field"], rel=1e-6)


def test_chi_sweep_small_against_oracle(hip):
    """The frozen-position chi sweep at oracle size: per point the energy and
    its chi derivative from the held moments, the forces from the held
    spectra, and the force loss L = 1/(2N) sum |F - F_t|^2 with dL/dchi,
    dL/dkappa and dL/dsigma through the adjoint chain, against the FP64
    oracle's forces and central differences of its L (h = 1e-3 relative)."""
    from hymd_amd.sweep import ChiSweep, coupling
    from oracle import field as F
    s = _small()
    T = s.n_types
    pm = _pm(s)
    x32 = s.positions.astype(np.float32)
    x64 = x32.astype(np.float64)
    xd = torch.as_tensor(x32).cuda()
    td = torch.as_tensor(s.types).cuda()
    sw = ChiSweep(pm, xd, td, T, s.kappa, s.sigma)
    rng = np.random.default_rng(44)
    target = 0.5 * sw.forces(s.chi).clone()
    tgt = target.cpu().numpy().astype(np.float64)

    def oracle(chi, kappa=s.kappa, sigma=s.sigma):
        A, b, c0 = coupling(chi, kappa, sw.rho0, sw.a)
        ham = F.Quadratic(A, b, c0)
        out = F.update_field(x64, s.types, T, s.mesh, s.box, sigma, ham)
        f = F.compute_field_force(x64, s.types, out["force_mesh"], s.mesh, s.box)
        return F.field_energy(out["phi"], ham, s.mesh, s.box), f, 0.5 * np.sum((f - tgt) ** 2) / s.n

    chis = []
    for _ in range(3):
        c = np.triu(rng.uniform(0.0, 30.0, (T, T)), 1)
        chis.append(c + c.T)
    res = sw.run(chis, target)
    h = 1e-3
    for chi, r in zip(chis, res):
        E0, f0, L0 = oracle(chi)
        assert r["E"] == pytest.approx(E0, rel=1e-6)
        f = sw.forces(chi).cpu().numpy()
        assert rms_rel(f, f0) < 1e-5
        assert r["L"] == pytest.approx(L0, rel=1e-5)
        t, u = 2, 5
        cp, cm = chi.copy(), chi.copy()
        cp[t, u] = cp[u, t] = chi[t, u] * (1 + h)
        cm[t, u] = cm[u, t] = chi[t, u] * (1 - h)
        Ep, _, Lp = oracle(cp)
        Em, _, Lm = oracle(cm)
        d = 2 * chi[t, u] * h
        assert r["dE_dchi"][t, u] == pytest.approx((Ep - Em) / d, rel=1e-5)
        assert r["dL_dchi"][t, u] == pytest.approx((Lp - Lm) / d, rel=1e-4)
    chi = chis[0]
    k, sg = s.kappa, s.sigma
    fd = (oracle(chi, kappa=k * (1 + h))[2] - oracle(chi, kappa=k * (1 - h))[2]) / (2 * k * h)
    assert res[0]["dL_dkappa"] == pytest.approx(fd, rel=1e-4)
    fd = (oracle(chi, sigma=sg * (1 + h))[2] - oracle(chi, sigma=sg * (1 - h))[2]) / (2 * sg * h)
    assert res[0]["dL_dsigma"] == pytest.approx(fd, rel=1e-4)
