"""Autotuner of the particle-mesh kernels' launch shapes.

The knobs (``hymd_pm_set_tuning``) change how the update_field / readout
kernels are launched, not what they compute: the paint and readout shapes
leave every result identical bit for bit; another x-line width may use the
runtime FFT stage geometry instead of a compile-time one, which rounds the
FP32 x transform differently (relative 1e-7):

* ``paint_rows``     mesh rows per paint workgroup: 1 (a workgroup of 256
                     threads per row) or 4 (one wavefront per row);
* ``readout_block``  threads per readout workgroup: 256, 128 or 64;
* ``xline_width``    lines per workgroup of the fused x-line FFT + k-space
                     pass: 0 (the widest that fits, with the compile-time
                     geometries where they exist) or 2, 4, 8, 16.

``autotune`` times one field update and force readout of the caller's
workload per candidate with HIP events on the launch stream (median of
``reps``), one knob at a time (coordinate descent from the defaults), keeps
the fastest, and caches the choice per (device, mesh, types, particle count
rounded to a power of two) in a JSON file (``$HYMD_TUNE_CACHE``, default
``~/.cache/hymd_amd/tune.json``), so later runs of the same shape start
tuned.
"""
import ctypes
import json
import os

import numpy as np
import torch

from . import _lib
from ._lib import call

KNOBS = {
    "paint_rows": (1, (1, 4)),
    "readout_block": (2, (256, 128, 64)),
    "xline_width": (3, (0, 16, 8, 4, 2)),
}
DEFAULTS = {"paint_rows": 1, "readout_block": 256, "xline_width": 0}


def get(pm):
    """The handle's current knob values."""
    out = {}
    v = ctypes.c_int32(0)
    for name, (knob, _) in KNOBS.items():
        call("hymd_pm_get_tuning", pm.handle, knob, ctypes.byref(v))
        out[name] = int(v.value)
    return out


def apply(pm, choice):
    """Set the knobs of ``choice`` (a dict name -> value) on the handle."""
    for name, value in choice.items():
        call("hymd_pm_set_tuning", pm.handle, KNOBS[name][0], int(value))


def cache_path():
    return os.environ.get("HYMD_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "hymd_amd",
                                                             "tune.json")


def cache_key(pm, n, n_types):
    dev = torch.cuda.get_device_name(torch.cuda.current_device()) if torch.cuda.is_available() else "none"
    bucket = 1 << max(0, int(np.ceil(np.log2(max(int(n), 1)))))
    return "%s|abi%d|mesh%s|T%d|N%d" % (dev, _lib.lib().hymd_abi_version(), "x".join(str(int(m)) for m in pm.Nmesh),
                                         int(n_types), bucket)


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(path, key, entry):
    data = _load(path)
    data[key] = entry
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def time_update(pm, x32, types, n_types, A, b, c0, mass, sigma, force_mesh, forces, reps):
    """Median milliseconds of one update_field (force grids) + force readout."""
    st = torch.cuda.current_stream()
    times = []
    for _ in range(reps + 1):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        pm.field_update(x32, types, n_types, A, b, c0, mass, sigma, force_mesh)
        pm.readout_forces(x32, types, n_types, force_mesh, forces)
        e1.record(st)
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    return float(np.median(times[1:]))


def autotune(pm, x32, types, n_types, A, b, c0, sigma, mass=None, reps=10, use_cache=True, path=None):
    """Choose the knobs for this workload; returns {"choice", "default_ms",
    "tuned_ms", "cached"}.  The handle is left tuned."""
    path = path or cache_path()
    n = int(x32.shape[0])
    key = cache_key(pm, n, n_types)
    if use_cache:
        entry = _load(path).get(key)
        if entry and set(entry.get("choice", {})) == set(KNOBS):
            apply(pm, entry["choice"])
            return dict(entry, cached=True)
    mass = [1.0] * int(n_types) if mass is None else mass
    fm = pm.create("real", batch=(int(n_types), 3))
    f = torch.empty((n, 3), dtype=torch.float32, device=x32.device)
    best = dict(DEFAULTS)
    apply(pm, best)
    default_ms = time_update(pm, x32, types, n_types, A, b, c0, mass, sigma, fm, f, reps)
    best_ms = default_ms
    trials = {}
    for name, (_, values) in KNOBS.items():
        for value in values:
            if value == best[name]:
                continue
            cand = dict(best, **{name: value})
            apply(pm, cand)
            ms = time_update(pm, x32, types, n_types, A, b, c0, mass, sigma, fm, f, reps)
            trials["%s=%d" % (name, value)] = round(ms, 4)
            if ms < best_ms * 0.99:   # a change must win by more than the timing noise
                best, best_ms = cand, ms
        apply(pm, best)
    entry = {"choice": best, "default_ms": round(default_ms, 4), "tuned_ms": round(best_ms, 4), "trials": trials}
    if use_cache:
        _store(path, key, entry)
    del fm
    return dict(entry, cached=False)
