"""Benchmark: particle-steps/s of the rRESPA hot path on one MI355X.

Workload: lipid_water, 1,048,576 beads, 40 nm box, 128^3 mesh, 4 types,
rRESPA inner = 10 (bonds + angles every inner step, particle-field forces
every outer step), synthetic seeded inputs already resident in HBM.
A step is one outer rRESPA step; value = N * steps / wall seconds.

Prints one JSON line on stdout (progress goes to stderr).  The timed region
replays one captured outer step (hipGraph) K times.  ``roofline`` is for one
stage (``--roofline-kernel``): the same outer steps are then launched eagerly
by the library, which records HIP events on the launch stream around every
launch of that stage.  ``cpu_baseline`` times the FP64 oracle (same algorithm:
numpy + scipy.fft + the C restatement of the bonded kernels) on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "particle-steps/s, 1M beads/128³ mesh (fp32 vs fp64 ref) + batched χ-grads/s"
HBM_PEAK_GBS = 8000.0
KERNEL_NAMES = {"bonded": "topo_gather_kernel", "fused": "fused_inner_kernel", "kick_drift": "kick_drift_kernel",
                "field": "update_field (sort + paint + FFT + k-space kernels)", "readout": "readout_kernel"}


def algorithmic_bytes(kernel, n, n_bonds, n_angles, T, mesh):
    """Compulsory HBM bytes of one launch of a stage."""
    G = mesh[0] * mesh[1] * mesh[2]
    if kernel == "fused":       # whole inner loop: x,v f64 read+write, F_slow f32 in, F_fast f32 in+out,
        # x32 out, row keys out (8 B), per-particle term list (4 B offset + 4 B per entry: 2 per bond,
        # 3 per angle), bond (a|b, r0, k) 20 B, angle (a|b|c, theta0, k) 24 B
        return 156 * n + 28 * n_bonds + 36 * n_angles
    if kernel == "bonded":      # x f64 in, F f32 out, bond a,b,r0,k and angle a,b,c,t0,k
        return 36 * n + 24 * n_bonds + 28 * n_angles
    if kernel == "kick_drift":  # v,x f64 read+write, F f32 in, x f32 out
        return 120 * n
    if kernel == "field":       # one update_field: cell sort, gather paint, r2c, k-space pass, c2r of 3T grids
        Gc = mesh[0] * mesh[1] * (mesh[2] // 2 + 1)
        sort = (16 * n) + (16 * G) + (40 * n)       # keys; scan of the cell counts; scatter x, type, index
        paint = 16 * n + 4 * G + 4 * T * G          # sorted x + type, cell starts, float densities out
        fwd = T * (4 * G + 8 * Gc) + 2 * 16 * T * Gc  # z r2c + y, x passes (read + write)
        kspace = 8 * T * Gc + 24 * T * Gc
        inv = 2 * 16 * 3 * T * Gc + 3 * T * (8 * Gc + 4 * G)
        return sort + paint + fwd + kspace + inv
    if kernel == "readout":     # sorted x f32 + type + row index in, 3 grids per type once, F f32 out
        return 32 * n + 12 * T * G
    raise ValueError(kernel)


def profiled(kernel, key):
    """A figure of `kernel` on this workload from the committed counter
    summary in profiles/ (HBM bytes per launch from calibrated FETCH_SIZE +
    WRITE_SIZE, or the VALU issue fraction), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_lipid_water_1M_128_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith(key + ":"):
                return float(line.split(":")[1])
    return None


def cpu_baseline(system, steps, warmup=3):
    """FP64 oracle outer steps on the host cores (bounded sample): `warmup`
    untimed steps, then `steps` steps timed one by one, value from the median
    step; also returns the oracle's initial (field, bonded) forces."""
    from oracle import field as OF
    from oracle import md as OM
    cores = min(os.cpu_count() or 1, 16)
    OF.set_workers(cores)
    A = system.extra["A"], system.extra["b"], system.extra["c0"]
    t0 = time.perf_counter()
    st = OM.State(system, OF.Quadratic(*A))
    t_init = time.perf_counter() - t0
    initial = (st.f_field.copy(), st.f_fast.copy(),
               {"field": st.field_energy(), "bond": st.e_fast["bond"], "angle": st.e_fast["angle"]})
    for _ in range(warmup):
        st.outer_step()
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        st.outer_step()
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return initial, {"value": system.n / med, "unit": "particle-steps/s", "cores": cores, "kind": "port",
                     "ms_per_step_median": round(1e3 * med, 1), "ms_per_step_min_max": [round(1e3 * min(times), 1),
                                                                                     round(1e3 * max(times), 1)],
                     "sample": f"{warmup} warm-up + {steps} timed outer steps (median) of the same 1,048,576-bead "
                               f"workload after a {t_init:.1f} s force initialisation; numpy/scipy.fft FP64 with "
                               f"{cores} FFT workers, single-threaded paint/readout and bonded C loops"}


def batched_chi_gradients(n_replicas, iters):
    """Second metric: replica-gradients/s of the batched path (one captured
    launch sequence over all replicas per evaluation: bin, paint, r2c,
    moments, k-space pass, inverse transforms, readout of every particle's
    field force; then the host energy / chi / kappa / sigma derivatives)
    against the per-replica path doing the same work (per replica one
    field-moments call with force grids and one force readout, on
    preallocated per-replica handles and buffers).  Each path's rate is the
    median over passes."""
    import torch

    from hymd_amd import synth
    from hymd_amd.autograd import energy_and_gradients, moments
    from hymd_amd.batch import ReplicaBatch
    from hymd_amd.field import ParticleMesh
    rs = synth.replica_set(n_replicas=n_replicas, seed=1003)
    x = torch.as_tensor(rs.positions.astype(np.float32)).cuda()
    t = torch.as_tensor(rs.types).cuda()
    f = torch.empty((len(rs.types), 3), dtype=torch.float32, device="cuda")
    b = ReplicaBatch(rs.meshes, rs.boxes, rs.counts, 2)
    for _ in range(2):
        b.energy_forces_and_gradients(x, t, rs.chi, rs.kappa, rs.sigma, forces=f)
    torch.cuda.synchronize()
    times = []
    for _ in range(iters):
        t0 = time.perf_counter()
        b.energy_forces_and_gradients(x, t, rs.chi, rs.kappa, rs.sigma, forces=f)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    batched = n_replicas / float(np.median(times))
    b.close()
    # per-replica path: one handle, force grid and force buffer per replica
    o = np.concatenate([[0], np.cumsum(rs.counts)])
    pms = [ParticleMesh(tuple(int(m) for m in rs.meshes[r]), rs.boxes[r], 2) for r in range(n_replicas)]
    fms = [pm.create("real", batch=(2, 3)) for pm in pms]
    rho0 = rs.counts / np.prod(rs.boxes, axis=1)
    A = [np.full((2, 2), 1.0 / (rs.kappa[r] * rho0[r])) + rs.chi[r] / rho0[r] for r in range(n_replicas)]
    bsh = [np.full(2, -1.0 / rs.kappa[r]) for r in range(n_replicas)]
    xs = [x[o[r]:o[r + 1]] for r in range(n_replicas)]
    ts = [t[o[r]:o[r + 1]] for r in range(n_replicas)]
    fs = [f[o[r]:o[r + 1]] for r in range(n_replicas)]

    def one_pass():
        for r in range(n_replicas):
            P, Q, phi0 = moments(pms[r], xs[r], ts[r], 2, rs.sigma[r], coupling=A[r], shift=bsh[r],
                                 force_mesh=fms[r])
            pms[r].readout_forces(xs[r], ts[r], 2, fms[r], fs[r])
            energy_and_gradients(P, Q, phi0, rs.chi[r], rs.kappa[r], rho0[r], rho0[r], rs.sigma[r],
                                 np.prod(rs.boxes[r]))

    one_pass()
    torch.cuda.synchronize()
    ptimes = []
    for _ in range(max(3, iters // 4)):
        t0 = time.perf_counter()
        one_pass()
        torch.cuda.synchronize()
        ptimes.append(time.perf_counter() - t0)
    single = n_replicas / float(np.median(ptimes))
    for pm in pms:
        pm.close()
    return {"value": batched, "unit": "replica-gradients/s", "replicas": n_replicas,
            "particles": int(rs.counts.sum()), "meshes": sorted({int(m) for m in rs.meshes[:, 0]}),
            "ms_per_batch": 1e3 * n_replicas / batched, "per_replica_path": single, "speedup": batched / single,
            "per_replica_ms_per_pass": round(1e3 * float(np.median(ptimes)), 3),
            "includes": "energy, field forces of every particle, dE/dchi, dE/dkappa, dE/dsigma per replica",
            "data": "synthetic A5B5 melts, seed 1003, chains in mesh order per replica"}


def peptide_solvent_stages(reps, check):
    """The 2,097,152-bead peptide/solvent workload (160^3 mesh shared by the
    density and PME, CBT dihedrals with dipole reconstruction): every force
    stage of one outer step timed on the device (HIP events on the launch
    stream, mean of `reps` calls), and, with `check`, the same forces from the
    FP64 oracle at the same configuration, with its host time."""
    import torch

    from hymd_amd import synth
    from hymd_amd.field import ParticleMesh
    from hymd_amd.force import BondedTopology, compute_dihedral_forces__fortran__double, \
        dipole_forces_redistribution
    s = synth.peptide_solvent(seed=1002)
    A, b, c0 = synth.hamiltonian_form(s)
    T, n = s.n_types, s.n
    ke, sigma = 138.935458, s.sigma
    elec = ke / s.dielectric
    x64 = torch.as_tensor(s.positions).cuda()
    x32 = x64.float().contiguous()
    types = torch.as_tensor(s.types).cuda()
    q = torch.as_tensor(s.charges).cuda()
    pm = ParticleMesh(s.mesh, s.box, T)
    fm = pm.create("real", batch=(T, 3))
    f_field = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    f_q = torch.empty_like(f_field)
    f_fast = torch.empty_like(f_field)
    topo = BondedTopology(n, s.bonds, s.angles)
    da, db, dc, dd, coeff, dty, last = s.extra["dihedrals"]
    m4 = len(da)
    dih = [torch.as_tensor(np.asarray(v, dtype=np.int32)).cuda() for v in (da, db, dc, dd)]
    coeff_d = torch.as_tensor(coeff).cuda()
    dty_d = torch.as_tensor(dty).cuda()
    last_d = torch.as_tensor(last).cuda()
    f_dih = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    dip = torch.zeros((m4, 4, 3), dtype=torch.float64, device="cuda")
    tmat = torch.zeros((m4, 6, 3, 3), dtype=torch.float64, device="cuda")
    f_dip = torch.randn((m4, 4, 3), dtype=torch.float64, device="cuda")
    f_red = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    e_self = float(ke / s.dielectric * np.sqrt(1.0 / (2.0 * np.pi * sigma * sigma)) * np.sum(s.charges.astype(np.float64) ** 2))
    box = s.box
    energies = {}
    stages = {
        "field_update": lambda: pm.field_update(x32, types, T, A, b, c0, [1.0] * T, sigma, fm),
        "field_readout": lambda: pm.readout_forces(x32, types, T, fm, f_field),
        "pme": lambda: pm.pme(x32, q, sigma, elec, e_self, f_q),
        "bonds_angles": lambda: topo.forces(x64, f_fast, box),
        "dihedrals_cbt_dipoles": lambda: compute_dihedral_forces__fortran__double(
            f_dih, x64, dip, tmat, box, *dih, coeff_d, dty_d, last_d, 1),
        "dipole_redistribution": lambda: dipole_forces_redistribution(f_red, f_dip, tmat, *dih, dty_d, last_d),
    }
    energies["field"] = pm.field_update(x32, types, T, A, b, c0, [1.0] * T, sigma, fm, want_energy=True)
    energies["field_q"] = pm.pme(x32, q, sigma, elec, e_self, f_q, want_energy=True)
    energies["dihedral"] = compute_dihedral_forces__fortran__double(f_dih, x64, dip, tmat, box, *dih, coeff_d, dty_d,
                                                                     last_d, 1)
    ms = {}
    for name, fn in stages.items():
        fn()
        torch.cuda.synchronize()
        st = torch.cuda.current_stream()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(st)
        for _ in range(reps):
            fn()
        ev1.record(st)
        torch.cuda.synchronize()
        ms[name] = ev0.elapsed_time(ev1) / reps
    # the whole outer step (respa_inner 1): the library's captured step, with
    # its default inner loop and with the other one
    from hymd_amd.md import RespaMD
    md = RespaMD(s)
    default_fused = md.fused
    paths = {}
    for fused in (default_fused, not default_fused):
        md.set_fused(fused)
        md.run(2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        md.run(reps)
        torch.cuda.synchronize()
        paths["fused" if fused else "per_kernel"] = 1e3 * (time.perf_counter() - t0) / reps
    step_ms = paths["fused" if default_fused else "per_kernel"]
    md.close()
    out = {"workload": "peptide_solvent: %d beads (%d chains x 40 residues x 3 beads + solvent), %.0f nm box, "
                       "%d^3 mesh for density and PME, %d types, %d charges, %d bonds, %d angles, %d CBT "
                       "dihedrals with dipoles" % (n, 8738, s.box[0], s.mesh[0], T, int(np.count_nonzero(s.charges)),
                                                   len(s.bonds[0]), len(s.angles[0]), m4),
           "device_ms": ms, "device_ms_outer_step_forces": sum(ms.values()),
           "md_ms_per_outer_step": step_ms, "md_particle_steps_per_s": n / (step_ms * 1e-3),
           "md_ms_per_outer_step_by_inner_loop": {k: round(v, 4) for k, v in paths.items()},
           "md_path": "RespaMD.run(): one captured outer step replayed as a hipGraph (%s inner loop: %s; PME of "
                      "the charges and of the dipoles)" % (
                          "fused" if default_fused else "per-kernel",
                          "bonds, angles and FP64 dihedrals per sub-step in one kernel, dipoles after it"
                          if default_fused else "the last sub-step's bonded and dihedral kernels on a second "
                                                "stream beside the field update and the charges' PME")}
    if check:
        from oracle import bonded as OB
        from oracle import field as OF
        OF.set_workers(min(os.cpu_count() or 1, 16))
        xo = s.positions.astype(np.float32).astype(np.float64)
        cpu = {}
        t0 = time.perf_counter()
        ref = OF.update_field(xo, s.types, T, s.mesh, s.box, sigma, OF.Quadratic(A, b, c0))
        fr = OF.compute_field_force(xo, s.types, ref["force_mesh"], s.mesh, s.box)
        cpu["field_update_and_readout"] = 1e3 * (time.perf_counter() - t0)
        t0 = time.perf_counter()
        rq = OF.update_field_force_q(xo, s.charges, s.mesh, s.box, sigma, ke, s.dielectric)
        cpu["pme"] = 1e3 * (time.perf_counter() - t0)
        t0 = time.perf_counter()
        fdo, _, _, edo = OB.cdf(s.positions, s.box, da, db, dc, dd, coeff, dty, last, 1)
        cpu["dihedrals_cbt_dipoles"] = 1e3 * (time.perf_counter() - t0)
        e_field_o = OF.field_energy(ref["phi"], OF.Quadratic(A, b, c0), s.mesh, s.box)
        e_q_o = OF.field_q_energy(rq["phi_q"], rq["psi"], OF.self_energy_q(s.charges, sigma, ke, s.dielectric),
                                  s.mesh, s.box)

        def err(got, want):
            return {"rel_rms": float(np.sqrt(np.mean((got - want) ** 2)) / np.sqrt(np.mean(want ** 2))),
                    "rel_max": float(np.abs(got - want).max() / np.abs(want).max())}
        out["accuracy_vs_fp64_oracle"] = {"field_force": err(f_field.cpu().numpy(), fr),
                                          "pme_force": err(f_q.cpu().numpy(), rq["forces"]),
                                          "dihedral_force": err(f_dih.cpu().numpy(), fdo),
                                          "field_energy_rel_err": abs(energies["field"] - e_field_o) / abs(e_field_o),
                                          "pme_energy_rel_err": abs(energies["field_q"] - e_q_o) / abs(e_q_o),
                                          "dihedral_energy_rel_err": abs(energies["dihedral"] - edo) / abs(edo)}
        out["cpu_oracle_ms"] = cpu
        out["cpu_oracle_cores"] = min(os.cpu_count() or 1, 16)
    pm.close()
    topo.close()
    return out


def markov_melt_leg(steps, sweep_points, tune=True):
    """The large single system: 16,777,216 beads, 100 nm box, 256^3 mesh, 6
    Markov-typed types in 16-bead chains, 5 % charges with PME on the same
    mesh, respa_inner 1.  `steps` outer steps of the library's captured step
    (fused inner loop, field + PME slow forces), then a chi sweep of
    `sweep_points` perturbations on the frozen final positions: one moments
    pass, then per point the energy derivatives, the forces from the held
    spectra and the gradient of the force loss through the adjoint chain.
    With `tune` the autotuner chooses the mesh kernels' launch shapes for
    this workload first (the library's normal set-up)."""
    import torch

    from hymd_amd import synth
    from hymd_amd.md import RespaMD
    from hymd_amd.sweep import ChiSweep
    t0 = time.perf_counter()
    s = synth.markov_melt()
    gen_s = time.perf_counter() - t0
    md = RespaMD(s, autotune=tune)
    md.run(3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    md.run(steps)
    torch.cuda.synchronize()
    step_ms = 1e3 * (time.perf_counter() - t0) / steps
    e = md.run(1, energies=True)
    stages = {}
    for name in ("bonded", "field", "readout", "pme"):
        md.time_stage(name)
        md.run(3, use_graph=False)
        torch.cuda.synchronize()
        stages[name] = round(md.stage_time()[0], 4)
        md.time_stage(None)
    out = {"workload": "markov_melt: %d beads (%d chains x 16), %.0f nm box, %d^3 mesh for density and PME, %d "
                       "types, %d charges, %d bonds, %d angles, respa_inner %d"
                       % (s.n, s.n // 16, s.box[0], s.mesh[0], s.n_types, int(np.count_nonzero(s.charges)),
                          len(s.bonds[0]), len(s.angles[0]), s.respa_inner),
           "generate_s": round(gen_s, 1), "fused_inner_loop": md.fused, "steps": steps,
           "ms_per_outer_step": round(step_ms, 4), "particle_steps_per_s": s.n / (step_ms * 1e-3),
           "stage_ms": stages, "energies_last_step": {k: e[k] for k in ("field", "field_q", "bond", "angle",
                                                                          "kinetic")},
           "autotune": None if md.tuning is None else {k: md.tuning[k] for k in ("choice", "default_ms", "tuned_ms")}}
    if sweep_points > 0:
        rng = np.random.default_rng(2004)
        x32 = md.x32.clone()
        types = md.types.clone()
        base = s.chi.copy()
        sw = ChiSweep(md.pm, x32, types, s.n_types, s.kappa, s.sigma)
        target = sw.forces(base).clone()
        chis = []
        for _ in range(sweep_points):
            d = np.triu(rng.uniform(-3.0, 3.0, base.shape), 1)
            chis.append(np.clip(base + d + d.T, 0.0, None))
        sw.run(chis[:2], target)   # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = sw.run(chis, target)
        total = time.perf_counter() - t0
        out["chi_sweep"] = {"points": sweep_points, "moments_pass_ms": round(1e3 * sw.moments_s, 3),
                            "ms_total": round(1e3 * total, 2), "ms_per_point": round(1e3 * total / sweep_points, 3),
                            "per_point": "energy, dE/dchi, dE/dkappa, dE/dsigma from the moments; forces from "
                                         "the held spectra; L = 1/(2N) sum |F - F_base|^2 and dL/dchi, dL/dkappa, "
                                         "dL/dsigma through the adjoint chain",
                            "L_range": [min(r["L"] for r in res), max(r["L"] for r in res)]}
        del sw
    md.close()
    return out


def lipid_water_bond_table(s):
    """The synthetic lipid/water system as an input file holds it: a
    (N, 2) bond table (chain neighbours, -1 padded), names, molecules."""
    n = s.n
    a, b = s.bonds[0], s.bonds[1]
    table = -np.ones((n, 2), dtype=np.int64)
    table[a, 1] = b
    table[b, 0] = a
    names = np.array(s.type_names)[s.types]
    n_chain = len(a) // 11
    molecules = np.r_[np.repeat(np.arange(n_chain), 12), n_chain + np.arange(n - 12 * n_chain)]
    bonds = [[p, q, 0.47, 1250.0] for p, q in (("H", "H"), ("H", "L"), ("L", "L"), ("L", "T"), ("T", "T"))]
    angles = [["H", "H", "L", 180.0, 25.0], ["H", "L", "L", 120.0, 25.0], ["L", "L", "T", 120.0, 25.0],
              ["L", "T", "T", 180.0, 25.0], ["T", "T", "T", 180.0, 25.0]]
    return molecules, names, table, bonds, angles


def topology_build(s, reps, cpu_chains):
    """prepare_bonds (force.py:573-728) of the lipid/water system on the
    device, checked against the system's own term lists; the oracle on the
    first `cpu_chains` chains and as many waters per chain is its CPU
    baseline."""
    import torch
    from hymd_amd.topology import prepare_bonds

    class Cfg:
        pass

    molecules, names, table, bonds, angles = lipid_water_bond_table(s)
    cfg = Cfg()
    cfg.bonds, cfg.angle_bonds, cfg.dihedrals = bonds, angles, []
    idx = np.arange(s.n)
    prepare_bonds(molecules[:24], names[:24], np.where(table[:24] < 24, table[:24], -1), idx[:24], cfg)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = prepare_bonds(molecules, names, table, idx, cfg)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ok = (np.array_equal(out[0].cpu().numpy(), s.bonds[0]) and np.array_equal(out[1].cpu().numpy(), s.bonds[1])
          and np.array_equal(np.stack([t.cpu().numpy() for t in out[4:7]]), np.stack(s.angles[:3])))
    res = {"beads": s.n, "molecules": int(molecules[-1]) + 1, "ms": round(1e3 * min(times), 2),
           "beads_per_s": s.n / min(times), "bonds": len(out[0]), "angles": len(out[4]), "matches_system_terms": ok,
           "includes": "host marshalling (name ids, molecule segments, uploads) + device build + copies"}
    if cpu_chains > 0:
        from oracle import topology as OT          # CPU baseline only
        n_chain = len(s.bonds[0]) // 11
        rows = np.r_[np.arange(12 * cpu_chains), 12 * n_chain + np.arange(8 * cpu_chains)]
        remap = -np.ones(s.n, dtype=np.int64)
        remap[rows] = np.arange(len(rows))
        sub = np.where(table[rows] >= 0, remap[np.maximum(table[rows], 0)], -1)
        t0 = time.perf_counter()
        OT.prepare_bonds(molecules[rows], names[rows], sub, np.arange(len(rows)), bonds, angles, [])
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"kind": "port", "cores": 1, "beads_per_s": len(rows) / dt,
                               "sample": f"{cpu_chains} chains + {8 * cpu_chains} waters ({len(rows)} beads)"}
    return res


def pressure_terms(md, s, reps, check):
    """comp_pressure's field terms (pressure.py:9-179) at the current
    configuration: one update_field writing the filtered spectra, then the
    device Parseval pass (hymd_field_pressure, synchronising); against the
    FP64 oracle's real-space sums over its Laplacian grids on the same
    positions, whose time is the CPU baseline."""
    import torch
    from hymd_amd.pressure import field_pressure
    T = md.T
    pf = md.pm.create("complex", batch=T)
    fm = md.pm.create("real", batch=(T, 3))
    md.pm.field_update(md.x32, md.types, T, md.A, md.b, md.c0, md.type_mass, md.sigma, fm, phi_fourier=pf)
    field_pressure(md.pm, pf, T, md.A, md.b, md.c0, md.sigma)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        got = field_pressure(md.pm, pf, T, md.A, md.b, md.c0, md.sigma)
    ms = 1e3 * (time.perf_counter() - t0) / reps
    res = {"p0_p1_p2xyz": [float(v) for v in got], "ms_per_call": round(ms, 4), "calls": reps,
           "spectra_bytes": int(pf.numel() * 8),
           "includes": "the Parseval kernel, its reduction, the host read-back (the call synchronises)"}
    if check:
        from oracle import field as OF          # CPU baseline and check only
        from oracle import pressure as OP
        cores = min(os.cpu_count() or 1, 16)
        OF.set_workers(cores)
        ham = OF.Quadratic(md.A, md.b, md.c0)
        x = md.x32.cpu().numpy().astype(np.float64)
        ty = md.types.cpu().numpy()
        t0 = time.perf_counter()
        ref = OF.update_field(x, ty, T, s.mesh, s.box, md.sigma, ham)
        t1 = time.perf_counter()
        want = OP.comp_pressure(ref["phi"], ref["phi_fourier"], ham, np.zeros((1, 3)), s.mass, s.mesh, s.box,
                                md.sigma, np.zeros(3), np.zeros(3))[1:6]
        t2 = time.perf_counter()
        scale = max(abs(want[0]), abs(want[1]))
        res["oracle_p0_p1_p2xyz"] = [float(v) for v in want]
        res["max_err_rel_to_p1"] = float(np.max(np.abs(got - want)) / scale)
        res["cpu_baseline"] = {"kind": "port", "cores": cores, "ms_pressure_terms": round(1e3 * (t2 - t1), 1),
                               "ms_update_field": round(1e3 * (t1 - t0), 1),
                               "sample": "the whole configuration (numpy + scipy.fft workers)"}
    return res


def nvt_overhead(md, s, steps, nve_ms_per_step, use_graph):
    """The same outer steps coupled to the CSVR thermostat after every step
    (n_b = 1, one coupling group with COM removal: the reference's defaults,
    main.py:1314-1316) at 310 K, tau 0.1 ps."""
    import torch
    from hymd_amd.config import Config
    cfg = Config(n_steps=steps, time_step=s.time_step, mesh_size=s.mesh, sigma=s.sigma, kappa=s.kappa,
                 mass=s.mass, respa_inner=s.respa_inner, target_temperature=310.0, tau=0.1,
                 unique_names=list(s.type_names), n_types=s.n_types, n_b=1, n_particles=s.n)
    prng = np.random.default_rng(1234)
    md.simulate(3, cfg, prng, use_graph=use_graph)    # warm-up (builds the group tables)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    md.simulate(steps, cfg, prng, use_graph=use_graph)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    return {"ms_per_step": round(ms, 4), "steps": steps, "thermostat_us_per_step": round(1e3 * (ms - nve_ms_per_step), 2),
            "kinetic_temperature_K": round(2.0 * md.kinetic_energy() / (3.0 * s.n * cfg.gas_constant), 2),
            "thermostat_work": md.thermostat_work,
            "coupling": "CSVR every outer step, one group, COM removal, T0 310 K, tau 0.1 ps"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--roofline-kernel", default="bonded", choices=["bonded", "kick_drift", "field", "readout"],
                    help="stage to time; 'bonded' is the fused inner loop when the topology allows it")
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="timed oracle outer steps for cpu_baseline, after 3 warm-up steps (0 = skip)")
    ap.add_argument("--batch-replicas", type=int, default=64, help="replicas of the batched chi-gradient metric (0 = skip)")
    ap.add_argument("--batch-iters", type=int, default=20)
    ap.add_argument("--peptide-reps", type=int, default=10,
                    help="timed calls per force stage of the peptide/solvent workload (0 = skip)")
    ap.add_argument("--small", action="store_true", help="reduced system for quick checks (not a bench line)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 10 device copies of MIB MiB (calibrates HBM byte counters) and exit")
    ap.add_argument("--no-graph", action="store_true", help="launch each step's kernels eagerly (no hipGraph)")
    ap.add_argument("--nvt-steps", type=int, default=30,
                    help="outer steps with the CSVR thermostat after every step (0 = skip)")
    ap.add_argument("--pressure-reps", type=int, default=20,
                    help="device comp_pressure field passes at the bench configuration (0 = skip)")
    ap.add_argument("--topology-reps", type=int, default=3,
                    help="device prepare_bonds builds of the bench system (0 = skip)")
    ap.add_argument("--large-steps", type=int, default=200,
                    help="outer steps of the 16,777,216-bead / 256^3 Markov melt with PME (0 = skip)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes (profiling runs; the bench line is the tuned step)")
    ap.add_argument("--sweep-points", type=int, default=32,
                    help="chi-sweep points on the large system's frozen final positions")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the particle-field path has no multi-GPU implementation; "
              "the value is one GPU's throughput", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe > 0:
        # a plain copy of known size: FETCH_SIZE and WRITE_SIZE per dispatch of
        # this kernel calibrate the counters against bytes moved
        a = torch.ones(args.copy_probe * 2 ** 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": a.numel() * 4, "copies": 10}))
        return

    from hymd_amd import synth
    from hymd_amd.md import RespaMD

    if args.small:
        s = synth.lipid_water(n_chains=5000, n_water=40000, box=16.0, mesh=48, seed=1001)
    else:
        s = synth.lipid_water(seed=1001)
    A, b, c0 = synth.hamiltonian_form(s)
    s.extra.update(A=A, b=b, c0=c0)
    torch.cuda.set_device(0)
    log(f"system ready: {s.n} beads, mesh {s.mesh}")
    md = RespaMD(s)
    log("device state and initial forces ready")
    f_field0 = md.in_input_order(md.f_field).cpu().numpy()
    f_fast0 = md.in_input_order(md.f_fast).cpu().numpy()          # per-kernel FP64 gather
    f_fused0 = (md.in_input_order(md.fused_fast_forces()).cpu().numpy()   # the timed fused kernel's terms
                if md.fused else None)
    md.slow_forces(energy=True)      # energies at the initial configuration (forces recomputed unchanged)
    md.fast_forces(energy=True)
    e0 = dict(md.energies)
    md.run(args.warmup, use_graph=not args.no_graph)
    torch.cuda.synchronize()

    def timed_steps():
        # timed region: K outer steps (one captured step replayed K times)
        t0 = time.perf_counter()
        md.run(args.steps, use_graph=not args.no_graph)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # the default launch shapes, then the autotuner's choice for this
    # workload (the headline is the tuned step, the library's normal set-up)
    default_elapsed = timed_steps()
    autotune = None
    if args.no_autotune:
        elapsed = default_elapsed
    else:
        from hymd_amd import tune
        tuning = tune.autotune(md.pm, md.x32, md.types, md.T, md.A, md.b, md.c0, md.sigma, mass=md.type_mass,
                               use_cache=False)
        md.run(args.warmup, use_graph=not args.no_graph)   # captures the step again with the tuned shapes
        torch.cuda.synchronize()
        elapsed = timed_steps()
        autotune = {"choice": tuning["choice"], "default_ms_per_step": round(1e3 * default_elapsed / args.steps, 4),
                    "tuned_ms_per_step": round(1e3 * elapsed / args.steps, 4),
                    "update_readout_default_ms": tuning["default_ms"], "update_readout_tuned_ms": tuning["tuned_ms"],
                    "trials_ms": tuning["trials"]}
    ms_per_step = 1e3 * elapsed / args.steps
    value = s.n * args.steps / elapsed
    log(f"timed {args.steps} steps: {ms_per_step:.4f} ms/step"
        + ("" if autotune is None else f" tuned ({autotune['default_ms_per_step']:.4f} default)")
        + f", {value:.4g} particle-steps/s")

    # roofline: the same outer steps launched eagerly by the library with HIP
    # events recorded on the launch stream around every launch of one stage
    md.time_stage(args.roofline_kernel)
    rsteps = max(1, min(args.steps, 20))
    t0 = time.perf_counter()
    md.run(rsteps, use_graph=False)
    torch.cuda.synchronize()
    eager_ms = 1e3 * (time.perf_counter() - t0) / rsteps
    dur_ms, launches = md.stage_time()
    md.time_stage(None)
    kname = "fused" if (args.roofline_kernel == "bonded" and md.fused) else args.roofline_kernel
    nbytes = algorithmic_bytes(kname, s.n, len(s.bonds[0]), len(s.angles[0]), s.n_types, s.mesh)
    achieved = nbytes / (dur_ms * 1e-3) / 1e9
    traffic = None if args.small else profiled(KERNEL_NAMES[kname], "traffic_bytes_per_launch")
    traffic = None if traffic is None else int(traffic)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": KERNEL_NAMES[kname], "avg_launch_us": round(dur_ms * 1e3, 2),
                "launches_timed": launches, "algorithmic_bytes": int(nbytes),
                "eager_ms_per_step": round(eager_ms, 4),
                # from the committed counter run: VALU wave-instructions per second
                # against 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles (the kernel's real bound)
                "valu_issue_frac_profiled": None if args.small else profiled(KERNEL_NAMES[kname], "valu_issue_frac")}

    log(f"roofline stage {args.roofline_kernel}: {dur_ms * 1e3:.2f} us/launch over {launches} launches")
    press = None
    if args.pressure_reps > 0:
        press = pressure_terms(md, s, args.pressure_reps, args.cpu_steps > 0 and not args.small)
        log(f"pressure field terms: {press['ms_per_call']:.3f} ms per call")
    nvt = None
    if args.nvt_steps > 0:
        nvt = nvt_overhead(md, s, args.nvt_steps, ms_per_step, not args.no_graph)
        log(f"nvt: {nvt['ms_per_step']:.4f} ms/step with the thermostat every step "
            f"(+{nvt['thermostat_us_per_step']:.1f} us)")
    cpu = accuracy = None
    if args.cpu_steps > 0:
        log(f"cpu baseline: {args.cpu_steps} oracle outer steps")
        f0, cpu = cpu_baseline(s, args.cpu_steps)
        log(f"cpu baseline: {cpu['value']:.4g} particle-steps/s")
        fo_field, fo_fast, eo = f0
        # FP32 device forces vs the FP64 oracle at the same initial configuration
        accuracy = {
            # the synthetic chains start with every bond at r0, so E_bond is
            # rounding-level there: its error is reported with the values
            "energies": {k: {"device": e0[k], "oracle": eo[k], "abs_err": abs(e0[k] - eo[k]),
                             "rel_err": abs(e0[k] - eo[k]) / abs(eo[k]) if eo[k] else None}
                         for k in ("field", "bond", "angle")},
            "field_force_rel_rms": float(np.sqrt(np.mean((f_field0 - fo_field) ** 2)) / np.sqrt(np.mean(fo_field ** 2))),
            "field_force_rel_max": float(np.abs(f_field0 - fo_field).max() / np.abs(fo_field).max()),
            # the fused inner loop's own forces (the timed kernel, its bond and
            # angle terms in the arithmetic the timed step uses)
            "bonded_force_rel_rms": (float(np.sqrt(np.mean((f_fused0 - fo_fast) ** 2)) / np.sqrt(np.mean(fo_fast ** 2)))
                                     if f_fused0 is not None else None),
            "bonded_force_rel_max": (float(np.abs(f_fused0 - fo_fast).max() / np.abs(fo_fast).max())
                                     if f_fused0 is not None else None),
            "bonded_force_kernel": "fused_inner_kernel, %s terms" % md.term_precision,
            # the per-kernel FP64 gather that seeds the first step
            "bonded_force_rel_rms_fp64_gather": float(np.sqrt(np.mean((f_fast0 - fo_fast) ** 2))
                                                      / np.sqrt(np.mean(fo_fast ** 2))),
        }

    topo = None
    if args.topology_reps > 0:
        log("prepare_bonds of the bench system")
        topo = topology_build(s, args.topology_reps, 10000 if args.cpu_steps > 0 else 0)
        log(f"prepare_bonds: {topo['ms']:.1f} ms for {s.n} beads")

    batch = None
    if args.batch_replicas > 0:
        md.close()
        del md
        torch.cuda.empty_cache()
        log(f"batched chi gradients: {args.batch_replicas} replicas")
        batch = batched_chi_gradients(args.batch_replicas, args.batch_iters)
        log(f"batched chi gradients: {batch['value']:.4g} replica-gradients/s "
            f"({batch['speedup']:.2f}x the per-replica path)")

    peptide = None
    if args.peptide_reps > 0 and not args.small:
        torch.cuda.empty_cache()
        log("peptide/solvent force stages")
        peptide = peptide_solvent_stages(args.peptide_reps, check=args.cpu_steps > 0)
        log(f"peptide/solvent: {peptide['device_ms_outer_step_forces']:.3f} ms of device force stages, "
            f"{peptide['md_ms_per_outer_step']:.3f} ms per outer step")

    large = None
    if args.large_steps > 0 and not args.small:
        torch.cuda.empty_cache()
        log("markov melt: 16,777,216 beads, 256^3 mesh, PME")
        large = markov_melt_leg(args.large_steps, args.sweep_points, tune=not args.no_autotune)
        log(f"markov melt: {large['ms_per_outer_step']:.3f} ms per outer step"
            + (f", chi sweep {large['chi_sweep']['ms_per_point']:.2f} ms per point" if "chi_sweep" in large else ""))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "lipid_water rRESPA outer step: %d beads, %.0f nm box, %d^3 mesh, %d types, "
                               "respa_inner %d, bonds %d, angles %d" % (s.n, s.box[0], s.mesh[0], s.n_types,
                                                                      s.respa_inner, len(s.bonds[0]),
                                                                      len(s.angles[0])),
                   "precision": "fp32 paint/FFT/forces; bond and angle terms in fp32 arithmetic (the FP32 "
                                "forces this path allows: the reference's real(4) kernels difference positions "
                                "in fp32 and evaluate the terms in fp64); fp64 positions/velocities, kicks and "
                                "reductions"},
        "roofline": roofline,
        "autotune": autotune,
        "cpu_baseline": cpu,
        "accuracy_vs_fp64_oracle": accuracy,
        "batched_chi_grads": batch,
        "peptide_solvent_forces": peptide,
        "nvt_csvr": nvt,
        "prepare_bonds": topo,
        "pressure_field_terms": press,
        "markov_melt_256": large,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
