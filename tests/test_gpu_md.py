"""rRESPA kernels and the device MD loop.

* kick / drift kernels against the closed-form recurrences of
  test/test_integrator.py:60-298 (FP64, 1e-13 absolute as the reference);
* np.mod wrapping semantics of main.py:837;
* a few outer steps of RespaMD against the FP64 oracle loop.
"""
import ctypes

import numpy as np
import pytest
import torch

from tests.conftest import rel_err, rms_rel

pytestmark = pytest.mark.gpu


def _kick_drift(x, v, f_fast, c_fast, f_slow, c_slow, dt, box):
    from hymd_amd import _lib
    from hymd_amd._lib import call, darr, dptr
    call("hymd_respa_kick_drift", dptr(x), dptr(v), None, x.shape[0], dptr(f_fast), c_fast, dptr(f_slow), c_slow,
         dt, darr(box, 3), _lib.stream_ptr())


def _kick(v, f1, c1, f2=None, c2=0.0):
    from hymd_amd import _lib
    from hymd_amd._lib import call, dptr
    call("hymd_respa_kick", dptr(v), v.shape[0], dptr(f1), c1, dptr(f2), c2, _lib.stream_ptr())


@pytest.mark.parametrize("respa", [1, 2, 3, 4, 5, 10, 20])
def test_respa_constant_force_recurrence(hip, goldens, respa):
    """test/test_integrator.py:225-298 with m = 1 (the kernels take F/m
    coefficients) and float32-representable accelerations."""
    g = goldens["three_atoms"]
    Dt = g["dt"]
    a_l = np.array(goldens["integrator"]["acceleration_long"]).astype(np.float32)
    a_s = np.array(goldens["integrator"]["acceleration_short"]).astype(np.float32)
    x0 = np.array(g["positions"])
    v0 = np.array(g["velocities"])
    box = np.array([1e6, 1e6, 1e6])  # no wrapping
    x = torch.as_tensor(x0.copy()).cuda()
    v = torch.as_tensor(v0.copy()).cuda()
    fl = torch.as_tensor(a_l).cuda()
    fs = torch.as_tensor(a_s).cuda()
    M = respa
    dt = Dt / M
    al, as_ = a_l.astype(np.float64), a_s.astype(np.float64)
    for step in range(1, 10):
        v_o = v.cpu().numpy()
        p_o = x.cpu().numpy()
        for s in range(M):
            if s == 0:
                _kick_drift(x, v, fs, 0.5 * dt, fl, 0.5 * Dt, dt, box)
            else:
                _kick_drift(x, v, fs, 0.5 * dt, fs, 0.5 * dt, dt, box)
        _kick(v, fs, 0.5 * dt, fl, 0.5 * Dt)
        v_exp = v0 + step * (Dt * al + M * dt * as_)
        p_exp = p_o + (dt * M * v_o + dt * M / 2 * Dt * al + dt * M ** 2 / 2 * dt * as_)
        np.testing.assert_allclose(v.cpu().numpy(), v_exp, atol=1e-13, rtol=0)
        np.testing.assert_allclose(x.cpu().numpy(), p_exp, atol=1e-13, rtol=0)


def test_wrap_matches_np_mod(hip):
    box = np.array([3.0, 4.0, 5.0])
    rng = np.random.default_rng(0)
    x0 = rng.uniform(-10, 10, (1000, 3))
    x0[0] = [-1e-20, -1e-17, 0.0]
    x0[1] = [3.0, -4.0, 10.0]          # L, -L, 2L: zero, returned as +0
    x0[2] = [-0.0, 8.0, -5.0]
    x0[3] = [2.9999999999999996, 7.999999999999999, -4.999999999999999]
    v0 = rng.normal(size=(1000, 3))
    v0[:4] = 0.0
    x = torch.as_tensor(x0.copy()).cuda()
    v = torch.as_tensor(v0.copy()).cuda()
    zero = torch.zeros((1000, 3), dtype=torch.float32, device="cuda")
    _kick_drift(x, v, zero, 0.0, None, 0.0, 0.01, box)
    got = x.cpu().numpy()
    # np.mod maps a tiny negative coordinate to L itself (main.py:837)
    np.testing.assert_array_equal(got[0], np.mod(x0[0], box))
    assert got[0, 0] == 3.0 and got[0, 1] == 4.0
    np.testing.assert_array_equal(got[1:4], np.mod(x0[1:4], box))
    assert not np.signbit(got[1:3]).any()
    # elsewhere x + dt v may differ by one rounding (fused multiply-add)
    np.testing.assert_allclose(got, np.mod(x0 + 0.01 * v0, box[None, :]), rtol=0, atol=1e-14)
    assert (got >= 0).all() and (got <= box).all()


def test_respa_md_matches_oracle_loop(hip, oracle_built):
    """Two outer steps (inner = 5) of the diblock melt: device loop (FP32 forces,
    FP64 x/v) against the FP64 oracle loop of main.py."""
    from hymd_amd import synth
    from hymd_amd.md import RespaMD
    from oracle import field as F
    from oracle import md as OM
    s = synth.diblock_melt(n_chains=400, box=7.0, mesh=32, seed=1000)
    A, b, c0 = synth.hamiltonian_form(s)
    ref = OM.State(s, F.Quadratic(A, b, c0))
    md = RespaMD(s)
    io = md.in_input_order
    assert rel_err(io(md.f_field).cpu().numpy(), ref.f_field) < 1e-4
    assert rel_err(io(md.f_fast).cpu().numpy(), ref.f_fast) < 1e-5
    for _ in range(2):
        ref.outer_step()
        md.step()
    x = io(md.x).cpu().numpy()
    dx = x - ref.x
    dx -= s.box * np.round(dx / s.box)
    assert np.abs(dx).max() < 1e-5
    assert rel_err(io(md.v).cpu().numpy(), ref.v) < 1e-4
    e = md.step(energies=True)
    ref.outer_step()
    assert e["kinetic"] == pytest.approx(ref.kinetic_energy(), rel=1e-6)
    assert e["field"] == pytest.approx(ref.field_energy(), rel=1e-5)
    assert e["bond"] == pytest.approx(ref.e_fast["bond"], rel=1e-6)
    assert e["angle"] == pytest.approx(ref.e_fast["angle"], rel=1e-6)


def _small_peptide():
    from hymd_amd import synth
    s = synth.peptide_solvent(n_chains=24, n_res=12, n_solvent=2500, box=7.5, mesh=24, seed=31)
    s.respa_inner = 2          # the dipoles are rebuilt on the last inner step only
    return s


def test_peptide_md_matches_oracle_loop(hip, oracle_built):
    """Peptide/solvent loop (field + PME of the charges + CBT dihedrals with
    dipoles whose charges go through PME and back onto the backbone), two
    outer steps of two inner steps, device against the FP64 oracle loop."""
    from hymd_amd import synth
    from hymd_amd.md import RespaMD
    from oracle import field as F
    from oracle import md as OM
    s = _small_peptide()
    A, b, c0 = synth.hamiltonian_form(s)
    ref = OM.State(s, F.Quadratic(A, b, c0))
    md = RespaMD(s)
    io = md.in_input_order
    assert md.protein and md.charges is not None
    f_slow = io(md.f_slow).cpu().numpy()
    assert rms_rel(f_slow, ref.f_field) < 1e-5
    assert rel_err(f_slow, ref.f_field) < 1e-4
    assert rel_err(io(md.f_fast).cpu().numpy(), ref.f_fast) < 1e-5
    # the dipole part alone (redistributed dipole PME forces)
    assert rel_err(io(md.f_recon).cpu().numpy(), ref.f_slow[2]) < 1e-4
    for _ in range(2):
        ref.outer_step()
        md.step()
    x = io(md.x).cpu().numpy()
    dx = x - ref.x
    dx -= s.box * np.round(dx / s.box)
    assert np.abs(dx).max() < 1e-5
    assert rel_err(io(md.v).cpu().numpy(), ref.v) < 1e-4
    np.testing.assert_allclose(md.dipoles.cpu().numpy(), ref.dipoles, atol=1e-6)
    e = md.step(energies=True)
    ref.outer_step()
    assert e["dihedral"] == pytest.approx(ref.e_fast["dihedral"], rel=1e-6)
    assert e["kinetic"] == pytest.approx(ref.kinetic_energy(), rel=1e-6)


@pytest.mark.parametrize("use_graph", [False, True])
def test_peptide_native_step_equals_python_step(hip, use_graph):
    """The library's outer step with PME, dihedrals and dipoles on the
    per-kernel inner loop (its default with dihedrals; eager and hipGraph
    replay, the last sub-step's bonded kernels on a second stream beside the
    field update) runs the kernels of the Python step: identical results."""
    from hymd_amd.md import RespaMD
    a = RespaMD(_small_peptide())
    b = RespaMD(_small_peptide())
    assert not b.fused
    for _ in range(3):
        a.step()
    b.run(3, use_graph=use_graph)
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x) and torch.equal(a.v, b.v)
    assert torch.equal(a.f_slow, b.f_slow) and torch.equal(a.f_fast, b.f_fast)
    assert torch.equal(a.dipoles, b.dipoles)
    ea = a.step(energies=True)
    eb = b.run(1, energies=True)
    for k in ("field", "bond", "angle", "kinetic", "dihedral", "field_q"):
        assert ea[k] == pytest.approx(eb[k], rel=1e-12), k


@pytest.mark.parametrize("use_graph", [False, True])
def test_fused_peptide_loop_matches_per_kernel_path(hip, use_graph):
    """Dihedrals and CBT in the molecule-blocked fused inner loop against the
    per-kernel inner loop (bond + angle gather, dihedral term kernel and slot
    gather, force sum), FP64 terms: the same kicks and per-term FP64 values;
    the fast force is rounded to float32 once instead of as two partial sums
    (one float32 ulp), so trajectories agree to that rounding.  The dipoles
    of the final positions and the energies, dihedral included, agree."""
    from hymd_amd.md import RespaMD
    a = RespaMD(_small_peptide(), term_precision="fp64")
    b = RespaMD(_small_peptide(), term_precision="fp64")
    b.set_fused(True)
    assert b.fused and not a.fused and b.protein
    a.run(1, use_graph=use_graph)
    b.run(1, use_graph=use_graph)
    torch.cuda.synchronize()
    assert rel_err(b.f_fast.cpu().numpy(), a.f_fast.cpu().numpy()) < 1e-6
    dx = (a.x - b.x).abs().max().item()
    dv = (a.v - b.v).abs().max().item() / a.v.abs().max().item()
    assert dx < 1e-9 and dv < 1e-7, (dx, dv)
    # the fused step's dipoles are those of the dihedral kernel at its final positions
    f = torch.empty_like(b.x)
    dip = torch.empty_like(b.dipoles)
    tm = torch.empty_like(b.tmat)
    b.dihedrals.forces(b.x, f, dip, tm, b.box, 1)
    torch.cuda.synchronize()
    assert (dip - b.dipoles).abs().max().item() < 1e-12
    assert (tm - b.tmat).abs().max().item() < 1e-12 * tm.abs().max().item()
    # and follow the position differences of the two paths
    assert (a.dipoles - b.dipoles).abs().max().item() < 1e-7
    assert (a.tmat - b.tmat).abs().max().item() < 1e-6 * a.tmat.abs().max().item()
    a.run(2, use_graph=use_graph)
    b.run(2, use_graph=use_graph)
    ea = a.run(1, energies=True)
    eb = b.run(1, energies=True)
    torch.cuda.synchronize()
    dx = (a.x - b.x).abs().max().item()
    dv = (a.v - b.v).abs().max().item() / a.v.abs().max().item()
    assert dx < 1e-7 and dv < 1e-5, (dx, dv)
    for k in ("field", "bond", "angle", "kinetic", "dihedral", "field_q"):
        assert ea[k] == pytest.approx(eb[k], rel=1e-7), k
    a.close()
    b.close()


def test_fused_peptide_run_matches_oracle(hip, oracle_built):
    """The library's outer step for the peptide/solvent system on the fused
    inner loop (FP32 bond / angle terms, FP64 dihedral terms; PME of the
    charges and of the dipole charges) against the FP64 oracle loop: two
    outer steps of two inner steps, then a third with energies."""
    from hymd_amd import synth
    from hymd_amd.md import RespaMD
    from oracle import field as F
    from oracle import md as OM
    s = _small_peptide()
    A, b, c0 = synth.hamiltonian_form(s)
    ref = OM.State(s, F.Quadratic(A, b, c0))
    md = RespaMD(s)
    md.set_fused(True)
    assert md.fused and md.protein
    io = md.in_input_order
    md.run(2, use_graph=True)
    for _ in range(2):
        ref.outer_step()
    torch.cuda.synchronize()
    dx = io(md.x).cpu().numpy() - ref.x
    dx -= s.box * np.round(dx / s.box)
    assert np.abs(dx).max() < 1e-5
    assert rel_err(io(md.v).cpu().numpy(), ref.v) < 1e-4
    np.testing.assert_allclose(md.dipoles.cpu().numpy(), ref.dipoles, atol=1e-6)
    e = md.run(1, energies=True)
    ref.outer_step()
    assert e["dihedral"] == pytest.approx(ref.e_fast["dihedral"], rel=1e-6)
    assert e["bond"] == pytest.approx(ref.e_fast["bond"], rel=1e-6)
    assert e["angle"] == pytest.approx(ref.e_fast["angle"], rel=1e-6)
    assert e["kinetic"] == pytest.approx(ref.kinetic_energy(), rel=1e-6)
    md.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_native_step_equals_python_step(hip, use_graph):
    """The library-enqueued outer step (eager and hipGraph replay) runs the same
    kernels in the same order as the Python loop: identical results."""
    from hymd_amd import synth
    from hymd_amd.md import RespaMD
    s = synth.diblock_melt(n_chains=400, box=7.0, mesh=32, seed=1000)
    a = RespaMD(s)
    b = RespaMD(s)
    b.set_fused(False)  # the per-kernel inner loop, as the Python step runs it
    for _ in range(3):
        a.step()
    b.run(3, use_graph=use_graph)
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x) and torch.equal(a.v, b.v)
    assert torch.equal(a.f_field, b.f_field) and torch.equal(a.f_fast, b.f_fast)
    ea = a.step(energies=True)
    eb = b.run(1, use_graph=use_graph, energies=True)
    for k in ("field", "bond", "angle", "kinetic"):
        assert ea[k] == pytest.approx(eb[k], rel=1e-12)
    # replaying the graph again continues the same trajectory
    a.step()
    a.step()
    b.run(2, use_graph=use_graph)
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x)
    a.close()
    b.close()


def test_stage_timing(hip):
    from hymd_amd import synth
    from hymd_amd.md import RespaMD
    s = synth.diblock_melt(n_chains=200, box=6.0, mesh=32, seed=1000)
    md = RespaMD(s)
    assert md.fused
    md.time_stage("bonded")  # the fused kernel: one launch per outer step
    md.run(2, use_graph=False)
    ms, n = md.stage_time()
    assert n == 2 and ms > 0.0
    md.set_fused(False)  # per-kernel path: one bonded launch per inner step
    md.time_stage("bonded")
    md.run(2, use_graph=False)
    ms, n = md.stage_time()
    assert n == 2 * s.respa_inner and ms > 0.0
    md.time_stage(None)
    md.close()


@pytest.mark.parametrize("make", ["lipid", "diblock"])
def test_fused_inner_loop_matches_per_kernel_path(hip, make):
    """Molecule-blocked fused inner loop against the per-kernel inner loop:
    same kicks in the same order and the same per-term math."""
    from hymd_amd import synth
    from hymd_amd.md import RespaMD
    if make == "lipid":
        s = synth.lipid_water(n_chains=3000, n_water=20000, box=12.0, mesh=32, seed=7)
    else:
        s = synth.diblock_melt(n_chains=500, box=7.5, mesh=32, seed=1000)
    a = RespaMD(s, term_precision="fp64")
    b = RespaMD(s, term_precision="fp64")
    a.set_fused(False)
    assert b.fused and not a.fused
    # one outer step: the same additions on the same per-term FP64 values up
    # to the compiler's contraction choices in two different kernels
    a.run(1, use_graph=False)
    b.run(1, use_graph=False)
    torch.cuda.synchronize()
    assert (a.x - b.x).abs().max().item() < 1e-12
    assert (a.v - b.v).abs().max().item() < 1e-11
    assert rel_err(b.f_fast.cpu().numpy(), a.f_fast.cpu().numpy()) < 1e-6
    for use_graph in (False, True):
        a.run(3, use_graph=use_graph)
        b.run(3, use_graph=use_graph)
    ea = a.run(1, energies=True)
    eb = b.run(1, energies=True)
    torch.cuda.synchronize()
    # 8 outer steps later the last-bit differences have grown with the dynamics
    dx = (a.x - b.x).abs().max().item()
    dv = (a.v - b.v).abs().max().item()
    assert dx < 1e-8 and dv < 1e-7, (dx, dv)
    for k in ("field", "bond", "angle", "kinetic"):
        assert ea[k] == pytest.approx(eb[k], rel=1e-10)
    a.close()
    b.close()


def test_fused_rejects_non_contiguous_molecules(hip):
    from hymd_amd import synth
    from hymd_amd._lib import HymdError
    from hymd_amd.md import RespaMD
    s = synth.diblock_melt(n_chains=50, box=5.0, mesh=16, seed=3)
    a, b = s.bonds[0].copy(), s.bonds[1].copy()
    b[0] = 25  # bead 0 bonded into the third chain: a molecule with a gap
    s.bonds = (a, b, s.bonds[2], s.bonds[3])
    md = RespaMD(s, spatial_order=False)
    assert not md.fused
    with pytest.raises(HymdError, match="contiguous"):
        md.set_fused(True)
    md.run(2)  # the per-kernel path still runs
    # the spatial layout keeps every molecule whole and contiguous, so the
    # same input runs fused, with the same trajectory in input order (FP64
    # terms, as the per-kernel path evaluates them)
    sp = RespaMD(s, term_precision="fp64")
    assert sp.fused
    sp.run(2)
    torch.cuda.synchronize()
    dx = sp.in_input_order(sp.x).cpu().numpy() - md.x.cpu().numpy()
    dx -= s.box * np.round(dx / s.box)
    assert np.abs(dx).max() < 1e-9
    md.close()
    sp.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_solvent_reorder_keeps_the_trajectory(hip, use_graph):
    """Putting the one-particle molecules back in mesh-row order every step
    changes only the row order of the state: their rows carry no bonded terms
    and the field path is order-independent (integer paint, per-particle
    readout), so the trajectory in input order is bit-identical to the run
    without reordering; the order maps stay inverse permutations and the
    types follow their particles."""
    from hymd_amd import _lib, synth
    from hymd_amd.md import RespaMD
    s = synth.lipid_water(n_chains=3000, n_water=20000, box=12.0, mesh=32, seed=7)
    a = RespaMD(s, reorder_every=0)
    b = RespaMD(s, reorder_every=1)
    assert a.fused and b.fused
    order0 = b.order.copy()
    for _ in range(3):
        a.run(2, use_graph=use_graph)
        b.run(2, use_graph=use_graph)
    torch.cuda.synchronize()
    for name in ("x", "v", "f_field", "f_fast", "x32"):
        assert torch.equal(a.in_input_order(getattr(a, name)), b.in_input_order(getattr(b, name))), name
    order = b.order
    assert np.array_equal(np.sort(order), np.arange(b.n))
    assert torch.equal(b._order[b._inverse], torch.arange(b.n, device=b._order.device))
    assert np.array_equal(b.in_input_order(b.types).cpu().numpy(), np.asarray(s.types, dtype=np.int32))
    tail = int(_lib.lib().hymd_md_single_tail(b._native_handle()))
    assert 0 < tail < b.n
    assert np.array_equal(order[:tail], order0[:tail])          # molecules keep their rows
    assert not np.array_equal(order[tail:], order0[tail:])      # the solvent was permuted
    ea = a.run(1, energies=True)
    eb = b.run(1, energies=True)
    for k in ("field", "bond", "angle"):
        assert ea[k] == eb[k], k
    assert eb["kinetic"] == pytest.approx(ea["kinetic"], rel=1e-13)   # an FP64 sum in another order
    a.close()
    b.close()


def test_single_step_graph_replays_equal_eager_steps(hip):
    """run(1) completes the velocities, so the next call starts without a
    pending outer kick: its captured step (the no-pending graph) must run the
    same kernels as the eager step."""
    from hymd_amd import synth
    from hymd_amd.md import RespaMD
    s = synth.lipid_water(n_chains=600, n_water=5000, box=8.0, mesh=32, seed=5)
    a = RespaMD(s)
    b = RespaMD(s)
    assert a.fused and b.fused
    for _ in range(4):
        a.run(1, use_graph=True)
        b.run(1, use_graph=False)
    a.run(3, use_graph=True)     # then the pending-kick graph, back to back
    b.run(3, use_graph=False)
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x) and torch.equal(a.v, b.v) and torch.equal(a.f_field, b.f_field)
    a.close()
    b.close()


@pytest.mark.parametrize("make", ["lipid", "diblock"])
def test_fp32_terms_follow_the_fp64_loop(hip, make):
    """FP32 bond / angle term arithmetic (the reference's real(4) kernels
    evaluate the terms in FP64 after an FP32 position difference) in the
    fused loop: fast forces within FP32 rounding of the FP64 loop's after one
    outer step; positions below the float32 resolution the field path sees
    (about 1e-6 nm at these box sizes) after one outer step and within 1e-5
    nm after eight."""
    from hymd_amd import synth
    from hymd_amd.md import RespaMD
    if make == "lipid":
        s = synth.lipid_water(n_chains=3000, n_water=20000, box=12.0, mesh=32, seed=7)
    else:
        s = synth.diblock_melt(n_chains=500, box=7.5, mesh=32, seed=1000)
    a = RespaMD(s, term_precision="fp64")
    b = RespaMD(s, term_precision="fp32")
    assert a.fused and b.fused
    a.run(1, use_graph=False)
    b.run(1, use_graph=False)
    torch.cuda.synchronize()
    assert rms_rel(b.f_fast.cpu().numpy(), a.f_fast.cpu().numpy()) < 1e-6
    assert (a.x - b.x).abs().max().item() < 1e-7
    a.run(7, use_graph=True)
    b.run(7, use_graph=True)
    ea = a.run(1, energies=True)
    eb = b.run(1, energies=True)
    torch.cuda.synchronize()
    dx = (a.x - b.x).abs().max().item()
    assert dx < 1e-5, dx
    for k in ("bond", "angle", "kinetic"):
        assert eb[k] == pytest.approx(ea[k], rel=1e-5)
    a.close()
    b.close()


@pytest.mark.parametrize("precision, tol", [("fp32", 2e-6), ("fp64", 1e-7)])
def test_fused_kernel_forces_against_the_oracle(hip, oracle_built, precision, tol):
    """The fused inner loop's own bond + angle forces (one sub-step with dt = 0
    and zero kicks) on the lipid/water workload, whose head-linker angles have
    theta0 = 120 degrees and whose tail angles are straight (theta0 = 180),
    against the FP64 oracle cbf_d + caf_d (compute_bond_forces__double.f90,
    compute_angle_forces__double.f90): relative RMS 2e-6 for FP32 term
    arithmetic, 1e-7 (the float32 rounding of the force output) for FP64
    terms."""
    from hymd_amd import synth
    from hymd_amd.md import RespaMD
    s = synth.lipid_water(n_chains=4000, n_water=30000, box=14.0, mesh=32, seed=17)
    assert np.any(np.isclose(s.angles[3], np.radians(120.0))) and np.any(np.isclose(s.angles[3], np.pi))
    md = RespaMD(s, term_precision=precision)
    assert md.fused
    f = md.in_input_order(md.fused_fast_forces()).cpu().numpy().astype(np.float64)
    x = md.in_input_order(md.x).cpu().numpy()
    ob, _, _ = oracle_built.cbf(x, s.box, *s.bonds)
    oa, _, _ = oracle_built.caf(x, s.box, *s.angles)
    ref = ob + oa
    assert rms_rel(f, ref) < tol
    assert rel_err(f, ref) < 20 * tol
    # the state did not move
    assert torch.equal(md.in_input_order(md.x).cpu(), torch.as_tensor(x))
    md.close()


def test_captured_step_follows_pm_buffer_growth_and_box(hip):
    """The captured outer step holds the pm handle's buffer pointers and mesh:
    after a call on the exposed ``md.pm`` with more particles (its sort
    buffers are reallocated) and after a new box (``pm.set_box``, as the
    barostats re-create the mesh, barostat.py:163-164), the next graph step
    equals the eager step bit for bit."""
    from hymd_amd import synth
    from hymd_amd.md import RespaMD
    s = synth.lipid_water(n_chains=600, n_water=5000, box=8.0, mesh=32, seed=23)
    a = RespaMD(s)
    b = RespaMD(s)
    for _ in range(2):
        a.run(1, use_graph=True)
        b.run(1, use_graph=False)
    # grow a's pm sort buffers: a field update of twice as many particles
    big = torch.rand((4 * s.n, 3), dtype=torch.float32, device="cuda") * float(s.box[0])
    tb = torch.randint(0, a.T, (4 * s.n,), dtype=torch.int32, device="cuda")
    fm = a.pm.create("real", batch=(a.T, 3))
    a.pm.field_update(big, tb, a.T, a.A, a.b, a.c0, a.type_mass, a.sigma, fm)
    torch.cuda.synchronize()
    del big, tb, fm
    torch.cuda.empty_cache()
    a.run(2, use_graph=True)
    b.run(2, use_graph=False)
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x) and torch.equal(a.v, b.v) and torch.equal(a.f_field, b.f_field)
    # a 3% larger box on both: positions scaled, mesh moved
    f = 1.03
    new_box = np.asarray(s.box, dtype=np.float64) * f
    for md in (a, b):
        md.x.mul_(f)
        md.x32.copy_(md.x.to(torch.float32))
        md.pm.set_box(new_box)
    a.run(2, use_graph=True)
    b.run(2, use_graph=False)
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x) and torch.equal(a.v, b.v) and torch.equal(a.f_field, b.f_field)
    # the wrap follows the new box
    assert float(a.x.max()) < float(new_box[0]) and float(a.x.max()) > float(s.box[0]) * 0.99
    a.close()
    b.close()
