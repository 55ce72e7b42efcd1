"""CSVR thermostat and COM momentum removal on device velocities
(thermostat.py:12-15, 111-219 over ``hymd_csvr_thermostat`` /
``hymd_cancel_com_momentum``).

``csvr_thermostat`` and ``cancel_com_momentum`` take the reference's
arguments, with particle *types* (device int32) in place of the names array:
a coupling group's names are mapped to types through
``config.type_to_name_map`` (or the position in ``config.unique_names``).
The random numbers are drawn on the host from ``prng`` through
``random_gaussian`` / ``random_chi_squared`` in the reference's order (per
group: one normal, then chi^2 with 3 N_g - 1 degrees of freedom), so a
seeded generator reproduces the reference's stream.  Without ``work`` the
thermostat work is added to ``config.thermostat_work`` (one device sync, as
the reference's float); ``CsvrThermostat`` keeps it on the device instead
and precomputes the group tables once.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import HymdError, call, dptr


def _random_gaussian(prng):
    """thermostat.py:47-62."""
    return prng.normal()


def _random_chi_squared(prng, M):
    """thermostat.py:65-108."""
    return prng.chisquare(M)


def _type_of_name(config):
    if getattr(config, "type_to_name_map", None):
        return {str(name): int(t) for t, name in config.type_to_name_map.items()}
    if getattr(config, "unique_names", None):
        return {str(name): t for t, name in enumerate(config.unique_names)}
    raise ValueError("config needs type_to_name_map or unique_names to map coupling groups to types")


def _name(x):
    return x.decode("utf-8") if isinstance(x, bytes) else str(x)


def _new_workspace(device):
    """A zero-filled library workspace (it carries the thermostat's COM shifts
    from call to call, so each thermostat owns one)."""
    nbytes = int(_lib.lib().hymd_thermostat_workspace_bytes())
    return torch.zeros(nbytes // 8, dtype=torch.float64, device=device)


class _Workspace:
    """Per-device scratch of cancel_com_momentum (it keeps no state)."""
    _per_device = {}

    @classmethod
    def get(cls, device):
        key = (device.type, device.index)
        ws = cls._per_device.get(key)
        if ws is None:
            ws = cls._per_device[key] = _new_workspace(device)
        return ws


class CsvrThermostat:
    """Group tables, particle counts, a device work accumulator and the
    library workspace for repeated thermostat calls on one particle set (the
    counts are fixed, so a reordering of the rows does not invalidate them).
    The workspace keeps each group's COM velocity of the last call as the
    shift of the next call's one-pass sums; results depend on it only at the
    rounding level, and only through this object's own history."""

    def __init__(self, types, config, n_types=None, remove_center_of_mass_momentum=True):
        _lib.require_gpu()
        if not (isinstance(types, torch.Tensor) and types.is_cuda and types.dtype == torch.int32):
            raise TypeError("types must be a device int32 tensor")
        self.n_types = int(n_types if n_types is not None else (config.n_types or int(types.max()) + 1))
        if not 1 <= self.n_types <= _lib.HYMD_MAX_TYPES:
            raise ValueError(f"n_types must be in [1, {_lib.HYMD_MAX_TYPES}]")
        groups = getattr(config, "thermostat_coupling_groups", None)
        if not groups or not any(groups):        # thermostat.py:176-177
            names = list(config.unique_names) if getattr(config, "unique_names", None) else None
            groups = [names] if names is not None else [None]
        type_of = _type_of_name(config) if groups != [None] else None
        type_group = np.full(self.n_types, -1, dtype=np.int32)
        self.group_types = []
        for g, grp in enumerate(groups):
            if grp is None:                       # every type
                ts = list(range(self.n_types))
            else:
                ts = sorted({type_of[_name(nm)] for nm in grp if _name(nm) in type_of})
            for t in ts:
                if type_group[t] != -1 and remove_center_of_mass_momentum:
                    raise ValueError(f"type {t} is in coupling groups {type_group[t]} and {g}: with COM "
                                     "removal the groups must be disjoint")
                type_group[t] = g
            self.group_types.append(ts)
        if len(groups) > _lib.HYMD_MAX_TYPES:
            raise ValueError(f"at most {_lib.HYMD_MAX_TYPES} coupling groups")
        counts = torch.bincount(types.to(torch.int64), minlength=self.n_types).cpu().numpy()
        self.group_n = np.array([int(counts[ts].sum()) for ts in self.group_types], dtype=np.int64)
        self.n_groups = len(groups)
        self.remove_com = bool(remove_center_of_mass_momentum)
        self.types = types
        self.n = int(types.numel())
        self.config = config
        self._type_group = (ctypes.c_int32 * self.n_types)(*type_group.tolist())
        self._group_n = (ctypes.c_int64 * self.n_groups)(*self.group_n.tolist())
        self.work = torch.zeros(1, dtype=torch.float64, device=types.device)
        self._ws = _new_workspace(types.device)

    def draw(self, prng, random_gaussian=_random_gaussian, random_chi_squared=_random_chi_squared):
        """The reference's draws for one call (thermostat.py:197-198)."""
        R, S = [], []
        for g, n_g in enumerate(self.group_n):
            if n_g == 0:
                raise ValueError(f"thermostat coupling group {g} has no particles")
            R.append(float(random_gaussian(prng)))
            S.append(float(random_chi_squared(prng, 3 * int(n_g) - 1)))
        return R, S

    def __call__(self, velocity, prng, random_gaussian=_random_gaussian, random_chi_squared=_random_chi_squared,
                 stream=None, draws=None):
        """Thermostat ``velocity`` (device float64 (N,3)) in place; the work
        is added to ``self.work`` on the device.  ``draws`` = (R, S) skips
        the draws from ``prng``."""
        if velocity.shape != (self.n, 3):
            raise ValueError(f"velocity must have shape ({self.n}, 3)")
        R, S = draws if draws is not None else self.draw(prng, random_gaussian, random_chi_squared)
        cfg = self.config
        c = float(np.exp(-(cfg.time_step * cfg.respa_inner) / cfg.tau))
        ws = self._ws
        call("hymd_csvr_thermostat", dptr(velocity, torch.float64, "velocity"), dptr(self.types), self.n,
             self.n_types, self._type_group, self.n_groups, self._group_n, _lib.darr(R, self.n_groups),
             _lib.darr(S, self.n_groups), float(cfg.mass), float(cfg.gas_constant), float(cfg.target_temperature),
             c, 1 if self.remove_com else 0, dptr(self.work), dptr(ws), ws.numel() * 8, _lib.stream_ptr(stream))
        return velocity


def csvr_thermostat(velocity, names, config, prng, comm=None, random_gaussian=_random_gaussian,
                    random_chi_squared=_random_chi_squared, remove_center_of_mass_momentum=True, work=None):
    """thermostat.py:111-219 on device velocities; ``names`` is the device
    int32 type array.  The work goes to ``work`` (device float64 tensor,
    accumulated, no sync) or, when None, to ``config.thermostat_work``."""
    th = CsvrThermostat(names, config, remove_center_of_mass_momentum=remove_center_of_mass_momentum)
    th(velocity, prng, random_gaussian, random_chi_squared)
    if work is not None:
        work += th.work
    else:
        config.thermostat_work = float(getattr(config, "thermostat_work", 0.0) or 0.0) + float(th.work.item())
    return velocity


def cancel_com_momentum(velocities, config, comm=None, stream=None):
    """thermostat.py:12-15: v -= sum(v) / config.n_particles, in place."""
    _lib.require_gpu()
    n = int(velocities.shape[0])
    n_particles = int(config.n_particles if getattr(config, "n_particles", None) else n)
    ws = _Workspace.get(velocities.device)
    call("hymd_cancel_com_momentum", dptr(velocities, torch.float64, "velocities"), n, n_particles, dptr(ws),
         ws.numel() * 8, _lib.stream_ptr(stream))
    return velocities


__all__ = ["CsvrThermostat", "csvr_thermostat", "cancel_com_momentum", "HymdError"]
