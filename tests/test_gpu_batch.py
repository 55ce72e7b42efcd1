"""Batched replicas against each replica run alone, and their gradients
against central finite differences of the FP64 oracle energy.

Parity bar: the batched moments equal the single-replica path's bit for bit
(same kernels, plans, block counts and reduction order); the chi / kappa /
sigma derivatives match central differences (h = 1e-3 relative) of the FP64
oracle energy to 1e-5 relative (sigma: 1e-4; the FP32 spectra limit both).
"""
import numpy as np
import pytest
import torch

from oracle import field as F

pytestmark = pytest.mark.gpu


def _small_set():
    from hymd_amd import synth
    rs = synth.replica_set(n_replicas=7, n_min=2_000, n_max=12_000, meshes=(16, 24, 32), spacing=0.6, seed=21)
    # three mesh shapes (three FFT groups, interleaved in replica order), one
    # replica with a non-cubic mesh and box, one empty replica
    for r, m in enumerate([16, 24, 16, 32, 24, 16, 32]):
        rs.meshes[r] = m
    rs.meshes[2] = (16, 24, 20)
    rs.boxes[2] = rs.boxes[2] * np.array([0.9, 1.1, 1.0])
    x2, _ = rs.replica(2)
    o = np.concatenate([[0], np.cumsum(rs.counts)])
    rs.positions[o[2]:o[3]] = np.mod(x2, rs.boxes[2])
    keep = np.ones(len(rs.types), bool)
    keep[o[4]:o[5]] = False
    rs.positions, rs.types = rs.positions[keep], rs.types[keep]
    rs.counts[4] = 0
    return rs


def _batch(rs):
    from hymd_amd.batch import ReplicaBatch
    return ReplicaBatch(rs.meshes, rs.boxes, rs.counts, 2)


def _alone(rs, r, sigma):
    from hymd_amd.autograd import moments
    from hymd_amd.field import ParticleMesh
    x, t = rs.replica(r)
    pm = ParticleMesh(tuple(int(m) for m in rs.meshes[r]), rs.boxes[r], 2)
    return moments(pm, torch.as_tensor(x.astype(np.float32)).cuda(), torch.as_tensor(t).cuda(), 2, sigma)


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_each_replica_alone(hip, use_graph):
    rs = _small_set()
    b = _batch(rs)
    x = torch.as_tensor(rs.positions.astype(np.float32)).cuda()
    t = torch.as_tensor(rs.types).cuda()
    for sigma in (rs.sigma, rs.sigma * 1.25):   # a graph replay picks up new filter widths
        got = b.moments(x, t, sigma, use_graph=use_graph)
        for r in range(rs.n_replicas):
            P, Q, phi0 = _alone(rs, r, sigma[r])
            np.testing.assert_array_equal(got[r][0], P)
            np.testing.assert_array_equal(got[r][1], Q)
            np.testing.assert_array_equal(got[r][2], phi0)
    assert np.all(got[4][0] == 0.0)    # the empty replica
    b.close()


def _oracle_energy(x, types, mesh, box, chi, kappa, sigma):
    rho0 = len(types) / np.prod(box)
    ham = F.default_with_chi(2, kappa, rho0, rho0, chi)
    out = F.update_field(x, types, 2, mesh, box, sigma, ham)
    return F.field_energy(out["phi"], ham, mesh, box)


def test_batch_gradients_vs_fd(hip):
    rs = _small_set()
    b = _batch(rs)
    x32 = rs.positions.astype(np.float32)
    E, dchi, dk, ds = b.energy_and_gradients(torch.as_tensor(x32).cuda(), torch.as_tensor(rs.types).cuda(),
                                             rs.chi, rs.kappa, rs.sigma)
    # every output finite; the empty replica has no field: zero energy and gradients
    for arr in (E, dchi, dk, ds):
        assert np.all(np.isfinite(arr))
    assert E[4] == 0.0 and dk[4] == 0.0 and ds[4] == 0.0 and np.all(dchi[4] == 0.0)
    h = 1e-3
    for r in (0, 2, 5):
        x, t = rs.replica(r)
        x = x.astype(np.float32).astype(np.float64)
        mesh, box = tuple(int(m) for m in rs.meshes[r]), rs.boxes[r]
        chi, k, sg = rs.chi[r], rs.kappa[r], rs.sigma[r]
        assert E[r] == pytest.approx(_oracle_energy(x, t, mesh, box, chi, k, sg), rel=1e-6)
        c = chi[0, 1]
        cp, cm = chi.copy(), chi.copy()
        cp[0, 1] = cp[1, 0] = c * (1 + h)
        cm[0, 1] = cm[1, 0] = c * (1 - h)
        fd = (_oracle_energy(x, t, mesh, box, cp, k, sg) - _oracle_energy(x, t, mesh, box, cm, k, sg)) / (2 * c * h)
        assert dchi[r, 0, 1] == pytest.approx(fd, rel=1e-5, abs=1e-6 * abs(E[r]))
        fd = (_oracle_energy(x, t, mesh, box, chi, k * (1 + h), sg)
              - _oracle_energy(x, t, mesh, box, chi, k * (1 - h), sg)) / (2 * k * h)
        assert dk[r] == pytest.approx(fd, rel=1e-5)
        fd = (_oracle_energy(x, t, mesh, box, chi, k, sg * (1 + h))
              - _oracle_energy(x, t, mesh, box, chi, k, sg * (1 - h))) / (2 * sg * h)
        assert ds[r] == pytest.approx(fd, rel=1e-4)
    b.close()


def test_batch_rejects_bad_input(hip):
    from hymd_amd._lib import HymdError
    from hymd_amd.batch import ReplicaBatch
    with pytest.raises(HymdError):
        ReplicaBatch([(16, 16, 15)], [(5.0, 5.0, 5.0)], [10], 2)     # odd mesh edge
    b = ReplicaBatch([(16, 16, 16)], [(5.0, 5.0, 5.0)], [10], 2)
    with pytest.raises(HymdError):
        b.moments(torch.zeros((9, 3), device="cuda"), torch.zeros(9, dtype=torch.int32, device="cuda"), 0.5)
    b.close()


def _alone_forces(rs, r, A, b, sigma):
    from hymd_amd.autograd import moments
    from hymd_amd.field import ParticleMesh
    x, t = rs.replica(r)
    pm = ParticleMesh(tuple(int(m) for m in rs.meshes[r]), rs.boxes[r], 2)
    xd = torch.as_tensor(x.astype(np.float32)).cuda()
    td = torch.as_tensor(t).cuda()
    fm = pm.create("real", batch=(2, 3))
    P, Q, phi0 = moments(pm, xd, td, 2, sigma, coupling=A, shift=b, force_mesh=fm)
    f = torch.empty((len(t), 3), dtype=torch.float32, device="cuda")
    pm.readout_forces(xd, td, 2, fm, f)
    return (P, Q, phi0), f


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_forces_equal_each_replica_alone(hip, use_graph):
    """Forces and moments of the batched launch sequence equal the
    single-replica path (hymd_field_moments with force grids, then
    hymd_compute_field_force) bit for bit, replica by replica; a replay with
    new chi and sigma picks them up."""
    rs = _small_set()
    b = _batch(rs)
    x = torch.as_tensor(rs.positions.astype(np.float32)).cuda()
    t = torch.as_tensor(rs.types).cuda()
    o = np.concatenate([[0], np.cumsum(rs.counts)])
    for chi, sigma in ((rs.chi, rs.sigma), (rs.chi * 1.5, rs.sigma * 1.1)):
        E, dchi, dk, ds, f = b.energy_forces_and_gradients(x, t, chi, rs.kappa, sigma, use_graph=use_graph)
        rho0 = rs.counts / np.prod(rs.boxes, axis=1)
        A = b.coupling(chi, rs.kappa, rho0)
        for r in range(rs.n_replicas):
            if rs.counts[r] == 0:
                continue
            (P, Q, phi0), fr = _alone_forces(rs, r, A[r], np.full(2, -1.0 / rs.kappa[r]), sigma[r])
            assert torch.equal(f[o[r]:o[r + 1]], fr), r
        E2, dchi2, _, _ = b.energy_and_gradients(x, t, chi, rs.kappa, sigma, use_graph=use_graph)
        np.testing.assert_array_equal(E, E2)
        np.testing.assert_array_equal(dchi, dchi2)
    b.close()


def test_batch_eight_replicas_48_64_against_oracle(hip):
    """Eight replicas of the batched workload's generator (50k-400k beads, 48^3
    and 64^3 meshes): energies and field forces against the FP64 oracle, and
    dE/dchi, dE/dkappa, dE/dsigma against its central differences (h = 1e-3
    relative), all from one batched call."""
    from hymd_amd import synth
    rs = synth.replica_set(n_replicas=8, seed=1003)
    assert {int(m) for m in rs.meshes[:, 0]} == {48, 64}
    b = _batch(rs)
    x32 = rs.positions.astype(np.float32)
    E, dchi, dk, ds, f = b.energy_forces_and_gradients(torch.as_tensor(x32).cuda(), torch.as_tensor(rs.types).cuda(),
                                                       rs.chi, rs.kappa, rs.sigma)
    f = f.cpu().numpy()
    o = np.concatenate([[0], np.cumsum(rs.counts)])
    h = 1e-3
    F.set_workers(8)
    for r in range(rs.n_replicas):
        x, t = rs.replica(r)
        x = x.astype(np.float32).astype(np.float64)
        mesh, box = tuple(int(m) for m in rs.meshes[r]), rs.boxes[r]
        chi, k, sg = rs.chi[r], rs.kappa[r], rs.sigma[r]
        rho0 = len(t) / np.prod(box)
        ham = F.default_with_chi(2, k, rho0, rho0, chi)
        out = F.update_field(x, t, 2, mesh, box, sg, ham)
        fr = F.compute_field_force(x, t, out["force_mesh"], mesh, box)
        fo = f[o[r]:o[r + 1]]
        assert np.sqrt(np.mean((fo - fr) ** 2)) / np.sqrt(np.mean(fr ** 2)) < 1e-5
        assert E[r] == pytest.approx(F.field_energy(out["phi"], ham, mesh, box), rel=1e-6)
        c = chi[0, 1]
        cp, cm = chi.copy(), chi.copy()
        cp[0, 1] = cp[1, 0] = c * (1 + h)
        cm[0, 1] = cm[1, 0] = c * (1 - h)
        fd = (_oracle_energy(x, t, mesh, box, cp, k, sg) - _oracle_energy(x, t, mesh, box, cm, k, sg)) / (2 * c * h)
        assert dchi[r, 0, 1] == pytest.approx(fd, rel=1e-5, abs=1e-6 * abs(E[r]))
        fd = (_oracle_energy(x, t, mesh, box, chi, k * (1 + h), sg)
              - _oracle_energy(x, t, mesh, box, chi, k * (1 - h), sg)) / (2 * k * h)
        assert dk[r] == pytest.approx(fd, rel=1e-5)
        fd = (_oracle_energy(x, t, mesh, box, chi, k, sg * (1 + h))
              - _oracle_energy(x, t, mesh, box, chi, k, sg * (1 - h))) / (2 * sg * h)
        assert ds[r] == pytest.approx(fd, rel=1e-4)
    b.close()
