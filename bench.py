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
        # x32 out, per-particle term list (4 B offset + 4 B per entry: 2 per bond, 3 per angle),
        # bond (a|b, r0, k) 20 B, angle (a|b|c, theta0, k) 24 B
        return 148 * n + 28 * n_bonds + 36 * n_angles
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


def cpu_baseline(system, steps):
    """FP64 oracle outer steps on the host cores (bounded sample); also
    returns the oracle's initial (field, bonded) forces."""
    from oracle import field as OF
    from oracle import md as OM
    cores = min(os.cpu_count() or 1, 16)
    OF.set_workers(cores)
    A = system.extra["A"], system.extra["b"], system.extra["c0"]
    t0 = time.perf_counter()
    st = OM.State(system, OF.Quadratic(*A))
    t_init = time.perf_counter() - t0
    initial = (st.f_field.copy(), st.f_fast.copy())
    t0 = time.perf_counter()
    for _ in range(steps):
        st.outer_step()
    dt = time.perf_counter() - t0
    return initial, {"value": system.n * steps / dt, "unit": "particle-steps/s", "cores": cores, "kind": "port",
                "sample": f"{steps} outer steps of the same 1,048,576-bead workload after a {t_init:.1f} s "
                          f"force initialisation; numpy/scipy.fft FP64 with {cores} FFT workers, "
                          f"single-threaded paint/readout and bonded C loops"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--roofline-kernel", default="bonded", choices=["bonded", "kick_drift", "field", "readout"],
                    help="stage to time; 'bonded' is the fused inner loop when the topology allows it")
    ap.add_argument("--cpu-steps", type=int, default=4, help="oracle outer steps for cpu_baseline (0 = skip)")
    ap.add_argument("--small", action="store_true", help="reduced system for quick checks (not a bench line)")
    ap.add_argument("--no-graph", action="store_true", help="launch each step's kernels eagerly (no hipGraph)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

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
    f_field0 = md.f_field.cpu().numpy()
    f_fast0 = md.f_fast.cpu().numpy()
    md.run(args.warmup, use_graph=not args.no_graph)
    torch.cuda.synchronize()

    # timed region: K outer steps (one captured step replayed K times)
    t0 = time.perf_counter()
    md.run(args.steps, use_graph=not args.no_graph)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = s.n * args.steps / elapsed
    log(f"timed {args.steps} steps: {ms_per_step:.4f} ms/step, {value:.4g} particle-steps/s")

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
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": KERNEL_NAMES[kname], "avg_launch_us": round(dur_ms * 1e3, 2),
                "launches_timed": launches, "algorithmic_bytes": int(nbytes),
                "eager_ms_per_step": round(eager_ms, 4)}

    log(f"roofline stage {args.roofline_kernel}: {dur_ms * 1e3:.2f} us/launch over {launches} launches")
    cpu = accuracy = None
    if args.cpu_steps > 0:
        log(f"cpu baseline: {args.cpu_steps} oracle outer steps")
        f0, cpu = cpu_baseline(s, args.cpu_steps)
        log(f"cpu baseline: {cpu['value']:.4g} particle-steps/s")
        fo_field, fo_fast = f0
        # FP32 device forces vs the FP64 oracle at the same initial configuration
        accuracy = {
            "field_force_rel_rms": float(np.sqrt(np.mean((f_field0 - fo_field) ** 2)) / np.sqrt(np.mean(fo_field ** 2))),
            "field_force_rel_max": float(np.abs(f_field0 - fo_field).max() / np.abs(fo_field).max()),
            "bonded_force_rel_rms": float(np.sqrt(np.mean((f_fast0 - fo_fast) ** 2)) / np.sqrt(np.mean(fo_fast ** 2))),
        }

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
                   "precision": "fp32 paint/FFT/forces, fp64 positions/velocities and reductions"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy_vs_fp64_oracle": accuracy,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
