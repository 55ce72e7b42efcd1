"""HIP bonded kernels (cbf/caf/cdf, _d twins, redistribution, persistent
topology) against the reference goldens and the FP64 oracle.

Tolerances: FP64 kernels 1e-12 relative to the largest force (the math is the
Fortran's, evaluated per particle in a different summation grouping); FP32
outputs round forces to float32 (relative 1e-6).
"""
import numpy as np
import pytest
import torch

from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _random_terms(seed, n=3000, m=2500, box=(6.0, 7.0, 8.0)):
    rng = np.random.default_rng(seed)
    box = np.array(box)
    r = rng.uniform(0, 1, (n, 3)) * box
    a = rng.integers(0, n, m)
    return rng, box, r, a, (a + 1) % n, (a + 2) % n, (a + 3) % n


def test_cbf_caf_goldens(hip, goldens, oracle_built):
    from hymd_amd import force_kernels as K
    sys = goldens["dppc_single"]
    r = np.array(sys["positions"])
    box = np.array(sys["box"])
    t = np.array(goldens["bonds"]["terms"])
    f = np.zeros((12, 3))
    e, pr = K.cbf_d(f, r, box, t[:, 0].astype(int), t[:, 1].astype(int), t[:, 2], t[:, 3])
    assert e == pytest.approx(sum(goldens["bonds"]["energies"]), rel=1e-14)
    f_ref, e_ref, pr_ref = oracle_built.cbf(r, box, t[:, 0].astype(int), t[:, 1].astype(int), t[:, 2], t[:, 3])
    np.testing.assert_allclose(f, f_ref, rtol=0, atol=1e-12)
    np.testing.assert_allclose(pr, pr_ref, rtol=1e-13, atol=1e-12)
    t = np.array(goldens["angles"]["terms"])
    f = np.zeros((12, 3))
    e, pr = K.caf_d(f, r, box, t[:, 0].astype(int), t[:, 1].astype(int), t[:, 2].astype(int), np.radians(t[:, 3]),
                    t[:, 4])
    assert e == pytest.approx(sum(goldens["angles"]["energies"]), rel=1e-14)
    f_ref, _, pr_ref = oracle_built.caf(r, box, t[:, 0], t[:, 1], t[:, 2], np.radians(t[:, 3]), t[:, 4])
    np.testing.assert_allclose(f, f_ref, rtol=0, atol=1e-12)
    np.testing.assert_allclose(pr, pr_ref, rtol=1e-13, atol=1e-12)


def test_cdf_goldens(hip, goldens):
    from hymd_amd import force_kernels as K
    sys = goldens["alanine_octapeptide"]
    r = np.array(sys["positions"])
    box = np.array(sys["box"])
    g = goldens["dihedrals"]
    for term, E, Fi, Fj, Fk, Fl in zip(g["terms"], g["energies"], g["forces_i"], g["forces_j"], g["forces_k"],
                                       g["forces_l"]):
        coeff = np.zeros((1, 6, 5), dtype=np.float32)
        coeff[0, 0] = 1.0
        f = np.zeros((16, 3))
        dip = np.zeros((1, 4, 3))
        tm = np.zeros((1, 6, 3, 3))
        e = K.cdf_d(f, r, dip, tm, box, [term[0]], [term[1]], [term[2]], [term[3]], coeff, [0], [0], 0)
        assert e == pytest.approx(E, rel=1e-13)
        for idx, Fg in zip(term, (Fi, Fj, Fk, Fl)):  # goldens carry +grad E (see oracle tests)
            np.testing.assert_allclose(f[idx], -np.array(Fg), atol=1e-12)


@pytest.mark.parametrize("single", [False, True])
def test_bonds_angles_vs_oracle(hip, oracle_built, single):
    from hymd_amd import force_kernels as K
    rng, box, r, a, b, c, _ = _random_terms(1)
    if single:
        r = r.astype(np.float32).astype(np.float64)
    r0, k, t0 = rng.uniform(0.3, 0.6, len(a)), rng.uniform(100, 1500, len(a)), rng.uniform(1.6, 3.1, len(a))
    dt = np.float32 if single else np.float64
    tol = 1e-6 if single else 1e-12
    f = torch.zeros((len(r), 3), dtype=torch.float32 if single else torch.float64, device="cuda")
    rd = torch.as_tensor(r.astype(dt)).cuda()
    e, pr = (K.cbf if single else K.cbf_d)(f, rd, box, a, b, r0, k)
    f_ref, e_ref, pr_ref = oracle_built.cbf(r, box, a, b, r0, k)
    assert rel_err(f.cpu().numpy(), f_ref) < tol
    assert e == pytest.approx(e_ref, rel=1e-12)
    pr_tol = 1e-6 if single else 1e-10  # cbf returns bond_pr as real(4), cbf_d as real(8)
    np.testing.assert_allclose(pr, pr_ref, rtol=pr_tol)
    e, pr = (K.caf if single else K.caf_d)(f, rd, box, a, b, c, t0, k)
    f_ref, e_ref, pr_ref = oracle_built.caf(r, box, a, b, c, t0, k)
    assert rel_err(f.cpu().numpy(), f_ref) < tol
    assert e == pytest.approx(e_ref, rel=1e-12)
    np.testing.assert_allclose(pr, pr_ref, rtol=max(pr_tol, 1e-9))


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("dipole_flag", [0, 1])
def test_dihedrals_vs_oracle(hip, oracle_built, single, dipole_flag):
    """cosine series, CBT + dipole reconstruction (incl. last-angle terms) and improper types."""
    from hymd_amd import force_kernels as K
    rng = np.random.default_rng(4)
    box = np.array([9.0, 9.0, 9.0])
    n_chain, L = 40, 20
    r = np.zeros((n_chain, L, 3))
    r[:, 0] = rng.uniform(0, 9, (n_chain, 3))
    for i in range(1, L):
        r[:, i] = r[:, i - 1] + 0.38 * np.array([1.0, 0.6 * (-1) ** i, 0.1]) + rng.normal(scale=0.06, size=(n_chain, 3))
    r = np.mod(r.reshape(-1, 3), box)
    if single:
        r = r.astype(np.float32).astype(np.float64)
    a = (L * np.arange(n_chain)[:, None] + np.arange(L - 3)[None, :]).ravel()
    m = len(a)
    coeff = rng.uniform(-5, 5, (m, 6, 5)).astype(np.float32)
    coeff[:, 4, 0] = 40.0
    coeff[:, 5] = rng.uniform(-0.5, 0.5, (m, 5))
    dtype = rng.integers(0, 3, m)
    coeff[dtype == 2, 0, 1] = np.abs(coeff[dtype == 2, 0, 1])
    last = np.zeros(m, dtype=int)
    last[L - 4::L - 3] = 1
    dt = torch.float32 if single else torch.float64
    f = torch.zeros((len(r), 3), dtype=dt, device="cuda")
    dip = torch.zeros((m, 4, 3), dtype=dt, device="cuda")
    tm = torch.zeros((m, 6, 3, 3), dtype=dt, device="cuda")
    rd = torch.as_tensor(r).to(dt).cuda()
    fn = K.cdf if single else K.cdf_d
    e = fn(f, rd, dip, tm, box, a, a + 1, a + 2, a + 3, coeff, dtype, last, dipole_flag)
    f_ref, d_ref, t_ref, e_ref = oracle_built.cdf(r, box, a, a + 1, a + 2, a + 3, coeff, dtype, last, dipole_flag)
    tol = 2e-6 if single else 1e-11
    assert rel_err(f.cpu().numpy(), f_ref) < tol
    assert e == pytest.approx(e_ref, rel=1e-11)
    np.testing.assert_allclose(dip.cpu().numpy(), d_ref, atol=tol * 10)
    np.testing.assert_allclose(tm.cpu().numpy(), t_ref, atol=tol * 10 * max(1.0, np.abs(t_ref).max()))


def test_redistribution_vs_oracle(hip, oracle_built):
    from hymd_amd.force import dipole_forces_redistribution
    rng = np.random.default_rng(9)
    n, m = 500, 300
    a = rng.integers(0, n - 3, m)
    dtype = rng.integers(0, 3, m)
    last = rng.integers(0, 2, m)
    fd = rng.normal(size=(m, 4, 3))
    tm = rng.normal(size=(m, 6, 3, 3))
    out = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    dipole_forces_redistribution(out, torch.as_tensor(fd).cuda(), torch.as_tensor(tm).cuda(), a, a + 1, a + 2, a + 3,
                                 dtype, last)
    ref = oracle_built.dipole_forces_redistribution(n, fd, tm, a, a + 1, a + 2, a + 3, dtype, last)
    np.testing.assert_allclose(out.cpu().numpy(), ref, atol=1e-12)


def test_topology_forces(hip, oracle_built):
    from hymd_amd.force import BondedTopology
    from hymd_amd import synth
    s = synth.lipid_water(n_chains=2000, n_water=5000, box=8.0, mesh=32, seed=5)
    # random-walk bonds sit exactly at r0; move the beads off equilibrium
    s.positions = np.mod(s.positions + np.random.default_rng(6).normal(scale=0.05, size=s.positions.shape), s.box)
    topo = BondedTopology(s.n, bonds=s.bonds, angles=s.angles)
    x = torch.as_tensor(s.positions).cuda()
    f = torch.zeros((s.n, 3), dtype=torch.float32, device="cuda")
    (eb, ea), (pb, pa) = topo.forces(x, f, s.box, energies=True)
    fb, eb_ref, pb_ref = oracle_built.cbf(s.positions, s.box, *s.bonds)
    fa, ea_ref, pa_ref = oracle_built.caf(s.positions, s.box, *s.angles)
    assert rel_err(f.cpu().numpy(), fb + fa) < 1e-6
    assert eb == pytest.approx(eb_ref, rel=1e-12)
    assert ea == pytest.approx(ea_ref, rel=1e-12)
    np.testing.assert_allclose(pb, pb_ref, rtol=1e-10)
    np.testing.assert_allclose(pa, pa_ref, rtol=1e-10)


def test_empty_terms_zero_output(hip):
    from hymd_amd import force_kernels as K
    f = torch.ones((10, 3), dtype=torch.float32, device="cuda")
    r = torch.zeros((10, 3), dtype=torch.float32, device="cuda")
    e, pr = K.cbf(f, r, [1.0, 1.0, 1.0], np.zeros(0, int), np.zeros(0, int), np.zeros(0), np.zeros(0))
    assert e == 0.0 and float(f.abs().sum()) == 0.0


def test_bad_index_rejected(hip):
    from hymd_amd import force_kernels as K
    from hymd_amd._lib import HymdError
    f = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    r = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(HymdError, match="outside"):
        K.cbf(f, r, [1.0, 1.0, 1.0], [0], [7], [0.4], [100.0])
