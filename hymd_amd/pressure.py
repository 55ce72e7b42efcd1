"""Internal pressure on the device (pressure.py:9-179 over
``hymd_field_pressure``).

``comp_pressure`` takes the reference's arguments and returns its 18-vector
(numpy float64): kinetic, field p0 / p1 / p2_xyz, bond, angle, dihedral (0)
and total per direction.  The field terms come from the filtered spectra
``phi_fourier`` that update_field wrote, by Parseval sums on the device: for
the quadratic w of every Hamiltonian,

    p0   = -(1/2 S + b.phi~(0) + c0)  [- 1/2 E_q + E_self / V]
    p1   =   S + b.phi~(0)            [+ E_1]
    p2_d = -sigma^2 (S_d [+ E_1d])

with S = sum_k Re(phi~^* A phi~) and S_d its k_d^2-weighted sum (E_* the
same sums against psi), equal to the reference's real-space sums over the
Laplacian grids of comp_laplacian.  ``phi_laplacian`` and ``phi_transfer``
are accepted for the call signature and not written.  The kinetic term is
an FP64 device reduction of the velocities; bond_pr and angle_pr are the
3-vectors the bonded kernels returned.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import call, darr, dptr
from .field import _PM_REGISTRY, _device_tensor, _packed
from .hamiltonian import quadratic_form


def _spectra(phi_fourier):
    flat = _packed(phi_fourier, torch.complex64)
    if flat is not None:
        return flat[0], len(phi_fourier)
    t = phi_fourier if isinstance(phi_fourier, torch.Tensor) else torch.stack(list(phi_fourier))
    return t.contiguous(), t.shape[0]


def _pm_of_spectrum(spec, box):
    """The ParticleMesh of a (..., Mx, My, Mz/2+1) spectrum (even edges), whose
    box must be config.box_size (the pressure depends on it)."""
    mx, my, zh = (int(v) for v in spec.shape[-3:])
    pm = _PM_REGISTRY.get((mx, my, 2 * (zh - 1)))
    if pm is None:
        raise ValueError("no ParticleMesh of this mesh shape exists")
    if not np.allclose(pm.BoxSize, box, rtol=0, atol=0):
        raise ValueError(f"the ParticleMesh box {pm.BoxSize} differs from config.box_size {box}")
    return pm


def field_pressure(pm, phi_fourier, n_types, A, b, c0, sigma, psi=None, phi_q=None, type_charges=None,
                   self_energy=0.0):
    """(p0, p1, p2_x, p2_y, p2_z) from the filtered spectra (device)."""
    out = (ctypes.c_double * 5)()
    T = int(n_types)
    call("hymd_field_pressure", pm.handle, dptr(phi_fourier), T, darr(np.asarray(A, dtype=np.float64).ravel()),
         darr(b, T), float(c0), float(sigma), dptr(psi, torch.float32, "psi") if psi is not None else None,
         dptr(phi_q, torch.float32, "phi_q") if phi_q is not None else None,
         darr(type_charges, T) if psi is not None else None, float(self_energy), out)
    return np.array(out[:])


def comp_pressure(phi, phi_q, psi, hamiltonian, velocities, config, phi_fourier, phi_laplacian, phi_transfer,
                  positions, bond_pr, angle_pr, comm=None):
    """pressure.py:9-179: the 18 pressure components."""
    T = int(config.n_types)
    A, b, c0 = quadratic_form(hamiltonian, T)
    spec, _ = _spectra(phi_fourier)
    box = np.broadcast_to(np.asarray(config.box_size, dtype=np.float64), (3,))
    pm = _pm_of_spectrum(spec, box)
    V = float(np.prod(box))
    elec_w = config.coulombtype in ("PIC_Spectral", "PIC_Spectral_GPE")
    psi_t = None if psi is None else _device_tensor(psi, torch.float32, "psi").contiguous()
    phq_t = None if (psi is None or not elec_w or phi_q is None) else \
        _device_tensor(phi_q, torch.float32, "phi_q").contiguous()
    qt = None
    if psi_t is not None:
        qt = np.asarray(config.type_charges, dtype=np.float64)
    field = field_pressure(pm, spec, T, A, b, c0, config.sigma, psi_t, phq_t, qt,
                           float(config.self_energy or 0.0) if phq_t is not None else 0.0)
    v = _device_tensor(velocities, torch.float64, "velocities")
    ke = ctypes.c_double(0.0)
    call("hymd_kinetic_energy", dptr(v), int(v.shape[0]), _lib.HYMD_F64, float(config.mass), ctypes.byref(ke),
         _lib.stream_ptr())
    p_kin = 2.0 / (3.0 * V) * ke.value
    p0, p1, p2 = field[0], field[1], field[2:5]
    p_bond = np.asarray(bond_pr.cpu() if isinstance(bond_pr, torch.Tensor) else bond_pr, dtype=np.float64) / V
    p_angle = np.asarray(angle_pr.cpu() if isinstance(angle_pr, torch.Tensor) else angle_pr, dtype=np.float64) / V
    p_dih = np.zeros(3)
    p_tot = p_kin + p0 + p1 + p2 + p_bond + p_angle + p_dih
    return np.array([p_kin, p0, p1, *p2, *p_bond, *p_angle, *p_dih, *p_tot])


__all__ = ["comp_pressure", "field_pressure"]
