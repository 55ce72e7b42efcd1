"""Berendsen and stochastic-cell-rescaling barostats on the device
(barostat.py:45-309, barostat_scr.py:34-284).

``isotropic`` / ``semiisotropic`` take the reference's arguments and return
``(pm_stuff, change)``.  Every ``config.n_b``-th step they compute the
pressure with ``hymd_amd.pressure.comp_pressure`` (device), form the scale
factors from it on the host with the reference's formulas, scale
``config.box_size`` and the device positions (``hymd_scale_positions``,
FP64 products), and move the ParticleMesh of ``pm_stuff`` to the new box
(the reference re-creates its pmesh objects, barostat.py:163-164; the grids
are rewritten by the next update_field either way).

The reference's indexing is kept: the isotropic pressure is
np.average(pressure[-3:-1]) (the x and y totals), and the semi-isotropic
`positions[:][0:2] *= alpha_L` / `positions[:][2] *= alpha_N` scale all
coordinates of rows 0-1 / row 2; ``rows_quirk=False`` scales the x, y
columns by alpha_L and the z column by alpha_N instead.  ``x32`` (optional)
is a float32 mirror of the positions kept in step.  ``kind`` selects the
Berendsen (default, config.barostat_type "berendsen") or SCR ("scr")
formulas.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import call, dptr
from .pressure import comp_pressure

BAR = 16.61
BETA = {"berendsen": 4.6e-5, "scr": 7.6e-4}


def _scale(positions, x32, r0, r1, f, stream=None):
    if not (isinstance(positions, torch.Tensor) and positions.is_cuda and positions.is_contiguous()):
        raise TypeError("positions must be a contiguous device tensor")
    prec = _lib.HYMD_F64 if positions.dtype == torch.float64 else _lib.HYMD_F32
    if prec == _lib.HYMD_F32 and positions.dtype != torch.float32:
        raise TypeError("positions must be float32 or float64")
    n = int(positions.shape[0])
    r0, r1 = min(r0, n), min(r1, n)
    call("hymd_scale_positions", dptr(positions), prec, dptr(x32, torch.float32, "x32") if x32 is not None else None,
         n, r0, r1, (ctypes.c_double * 3)(*[float(v) for v in f]), _lib.stream_ptr(stream))


def _scale_lateral(positions, x32, a, rows_quirk):
    if rows_quirk:
        _scale(positions, x32, 0, 2, (a, a, a))
    else:
        _scale(positions, x32, 0, positions.shape[0], (a, a, 1.0))


def _scale_normal(positions, x32, a, rows_quirk):
    if rows_quirk:
        _scale(positions, x32, 2, 3, (a, a, a))
    else:
        _scale(positions, x32, 0, positions.shape[0], (1.0, 1.0, a))


def _move_mesh(pm_stuff, config):
    pm = pm_stuff[0] if isinstance(pm_stuff, (list, tuple)) else pm_stuff
    if pm is not None and hasattr(pm, "set_box"):
        pm.set_box(config.box_size)
    return pm_stuff


def _pressure(phi, phi_q, psi, hamiltonian, velocities, config, phi_fft, phi_laplacian, phi_transfer, positions,
              bond_pr, angle_pr):
    return comp_pressure(phi, phi_q, psi, hamiltonian, velocities, config, phi_fft, phi_laplacian, phi_transfer,
                         positions, bond_pr, angle_pr)


def isotropic(pmesh, pm_stuff, phi, phi_q, psi, hamiltonian, positions, velocities, config, phi_fft, phi_laplacian,
              phi_transfer, bond_pr, angle_pr, step, prng, comm=None, kind=None, x32=None):
    """barostat.py:45-157 (Berendsen) / barostat_scr.py:34-132 (SCR)."""
    kind = (kind or getattr(config, "barostat_type", None) or "berendsen").lower()
    if np.mod(step, config.n_b) != 0:
        return pm_stuff, False
    R = prng.normal() if kind == "scr" else None
    pressure = _pressure(phi, phi_q, psi, hamiltonian, velocities, config, phi_fft, phi_laplacian, phi_transfer,
                         positions, bond_pr, angle_pr)
    P = np.average(pressure[-3:-1]) * BAR
    dt = config.time_step * config.respa_inner
    beta = BETA[kind]
    P_t = config.target_pressure.P_L
    config.box_size = np.asarray(config.box_size, dtype=np.float64)
    if kind == "scr":
        V = np.prod(config.box_size)
        noise = np.sqrt(2.0 * config.n_b * config.gas_constant * config.target_temperature * beta * dt * config.n_b
                        / (V * config.tau_p)) * R
        alpha = np.exp((-config.n_b * dt * beta / config.tau_p * (P_t - P) + noise) / 3.0)
    else:
        alpha = (1.0 - dt * config.n_b / config.tau_p * beta * (P_t - P)) ** (1 / 3)
    config.box_size *= alpha
    _scale(positions, x32, 0, positions.shape[0], (alpha, alpha, alpha))
    return _move_mesh(pm_stuff, config), (kind != "scr")


def semiisotropic(pmesh, pm_stuff, phi, phi_q, psi, hamiltonian, positions, velocities, config, phi_fft,
                  phi_laplacian, phi_transfer, bond_pr, angle_pr, step, prng, comm=None, kind=None, x32=None,
                  rows_quirk=True):
    """barostat.py:160-309 (Berendsen) / barostat_scr.py:135-284 (SCR)."""
    kind = (kind or getattr(config, "barostat_type", None) or "berendsen").lower()
    if np.mod(step, config.n_b) != 0:
        return pm_stuff, False
    if kind == "scr":
        Rxy, Rz = prng.normal(), prng.normal()
    pressure = _pressure(phi, phi_q, psi, hamiltonian, velocities, config, phi_fft, phi_laplacian, phi_transfer,
                         positions, bond_pr, angle_pr)
    PL = (pressure[-3] + pressure[-2]) / 2 * BAR
    PN = pressure[-1] * BAR
    dt = config.time_step * config.respa_inner
    beta = BETA[kind]
    tp = config.target_pressure
    config.box_size = np.asarray(config.box_size, dtype=np.float64)
    box = config.box_size
    if kind == "scr":
        config.surface_tension = box[2] / 2 * (PN - PL)
        if tp.P_L:
            V = np.prod(box)
            noise = np.sqrt(4.0 * config.n_b * config.gas_constant * config.target_temperature * beta * dt
                            * config.n_b / (3 * V * config.tau_p)) * Rxy
            log_a = -2.0 * config.n_b * dt * beta / (3 * config.tau_p) * (
                tp.P_L - PL - config.surface_tension / box[2]) + noise
            a = np.exp(log_a / 2.0)
            box[0:2] *= a
            _scale_lateral(positions, x32, a, rows_quirk)
        if tp.P_N:
            V = np.prod(box)
            noise = np.sqrt(2.0 * config.n_b * config.gas_constant * config.target_temperature * beta * dt
                            * config.n_b / (3 * V * config.tau_p)) * Rz
            log_a = -config.n_b * dt * beta / (3 * config.tau_p) * (tp.P_N - PN) + noise
            a = np.exp(log_a / 1.0)
            box[2] *= a
            _scale_normal(positions, x32, a, rows_quirk)
    else:
        if tp.P_L:
            a = (1.0 - dt * config.n_b / config.tau_p * beta * (tp.P_L - PL)) ** (1 / 3)
            box[0:2] *= a
            _scale_lateral(positions, x32, a, rows_quirk)
        if tp.P_N:
            a = (1.0 - dt * config.n_b / config.tau_p * beta * (tp.P_N - PN)) ** (1 / 3)
            box[2] *= a
            _scale_normal(positions, x32, a, rows_quirk)
    return _move_mesh(pm_stuff, config), True


__all__ = ["isotropic", "semiisotropic"]
