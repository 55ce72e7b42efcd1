"""Drop-in for the bonded part of HyMD's force.py on the HIP kernels.

The compiled kernels come from ``force_kernels`` (the f2py module's
surface, cbf/caf/cdf and their _d variants), bound under force.py's names
as the reference does (force.py:8-23).  Here also:
``dipole_forces_redistribution`` (force.py:855-877) and the device
topologies that the MD loop uses.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import HYMD_F32, HYMD_F64, call, dptr
from .force_kernels import _InOut, _box, _to_dev
# imported here so that ``from hymd_amd.force import compute_bond_forces__fortran`` works, as in force.py:8-23
from .force_kernels import cbf as compute_bond_forces__fortran  # noqa: F401
from .force_kernels import caf as compute_angle_forces__fortran  # noqa: F401
from .force_kernels import cdf as compute_dihedral_forces__fortran  # noqa: F401
from .force_kernels import cbf_d as compute_bond_forces__fortran__double  # noqa: F401
from .force_kernels import caf_d as compute_angle_forces__fortran__double  # noqa: F401
from .force_kernels import cdf_d as compute_dihedral_forces__fortran__double  # noqa: F401


def dipole_forces_redistribution(f_on_bead, f_dipoles, trans_matrices, a, b, c, d, type_array, last_bb):
    """force.py:855-877: zero ``f_on_bead`` and add the dipole-point forces
    projected through the transfer matrices onto the backbone beads."""
    if isinstance(f_on_bead, torch.Tensor):
        dt = f_on_bead.dtype
    else:
        dt = torch.float32 if np.asarray(f_on_bead).dtype == np.float32 else torch.float64
    prec = HYMD_F32 if dt == torch.float32 else HYMD_F64
    n = f_on_bead.shape[0]
    out = _InOut(f_on_bead, dt, (n, 3), "f_on_bead")
    ai = _to_dev(a, torch.int32)
    m = ai.numel()
    fd = _to_dev(f_dipoles, dt).reshape(m, 4, 3).contiguous()
    tm = _to_dev(trans_matrices, dt).reshape(m, 6, 3, 3).contiguous()
    call("hymd_dipole_redistribution", dptr(out.dev), n, dptr(fd), dptr(tm), dptr(ai), dptr(_to_dev(b, torch.int32)),
         dptr(_to_dev(c, torch.int32)), dptr(_to_dev(d, torch.int32)), dptr(_to_dev(type_array, torch.int32)),
         dptr(_to_dev(last_bb, torch.int32)), m, prec, _lib.stream_ptr())
    out.finish()


class BondedTopology:
    """Bonds + angles with per-particle term lists built once on the device;
    ``forces(x64, f32)`` evaluates cbf + caf for the rRESPA inner loop."""

    def __init__(self, n, bonds=None, angles=None):
        _lib.require_gpu()
        self.n = int(n)
        z_i = torch.zeros(0, dtype=torch.int32, device="cuda")
        z_d = torch.zeros(0, dtype=torch.float64, device="cuda")
        if bonds is not None:
            a2, b2, r0, k2 = bonds
            self._b = [_to_dev(a2, torch.int32), _to_dev(b2, torch.int32), _to_dev(r0, torch.float64),
                       _to_dev(k2, torch.float64)]
        else:
            self._b = [z_i, z_i, z_d, z_d]
        if angles is not None:
            a3, b3, c3, t0, k3 = angles
            self._a = [_to_dev(a3, torch.int32), _to_dev(b3, torch.int32), _to_dev(c3, torch.int32),
                       _to_dev(t0, torch.float64), _to_dev(k3, torch.float64)]
        else:
            self._a = [z_i, z_i, z_i, z_d, z_d]
        h = ctypes.c_void_p()
        call("hymd_topology_create", ctypes.byref(h), self.n, self._b[0].numel(), *[dptr(t) for t in self._b],
             self._a[0].numel(), *[dptr(t) for t in self._a], _lib.stream_ptr())
        self._h = h

    @property
    def n_bonds(self):
        return self._b[0].numel()

    @property
    def n_angles(self):
        return self._a[0].numel()

    def forces(self, x64, f32, box, energies=False):
        """f32 <- bond + angle forces at x64; returns ((E_bond, E_angle),
        (bond_pr, angle_pr)) when ``energies``."""
        e = (ctypes.c_double * 2)() if energies else None
        v = (ctypes.c_double * 6)() if energies else None
        call("hymd_topology_forces", self._h, dptr(x64, torch.float64, "positions"),
             dptr(f32, torch.float32, "forces"), _box(box), e, v, _lib.stream_ptr())
        if energies:
            return (e[0], e[1]), (np.array(v[:3]), np.array(v[3:]))
        return None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().hymd_topology_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DihedralTopology:
    """A fixed dihedral term list on the device (compute_dihedral_forces__double
    + dipole reconstruction, force.py:855-877 redistribution) with per-particle
    lists built once; ``forces`` and ``redistribute`` run without per-call
    list builds or synchronisation (unless an energy is asked for)."""

    def __init__(self, n, a, b, c, d, coeff, dtype, last):
        _lib.require_gpu()
        self.n = int(n)
        self._t = [_to_dev(v, torch.int32) for v in (a, b, c, d)]
        self.m = self._t[0].numel()
        self.coeff = _to_dev(coeff, torch.float32).reshape(self.m, 6, 5).contiguous()
        self.dtype = _to_dev(dtype, torch.int32)
        self.last = _to_dev(last, torch.int32)
        h = ctypes.c_void_p()
        call("hymd_dihedrals_create", ctypes.byref(h), self.n, self.m, *[dptr(t) for t in self._t], dptr(self.coeff),
             dptr(self.dtype), dptr(self.last), _lib.stream_ptr())
        self._h = h

    def forces(self, x64, f64, dipoles, tmat, box, dipole_flag, energy=False):
        """f64 <- dihedral forces at x64; dipoles / tmat written (zero where a
        term has none).  Returns the energy when ``energy``."""
        e = ctypes.c_double(0.0)
        call("hymd_dihedrals_forces", self._h, dptr(x64, torch.float64, "positions"), dptr(f64, torch.float64, "forces"),
             dptr(dipoles, torch.float64, "dipoles"), dptr(tmat, torch.float64, "transfer_matrix"), _box(box),
             int(dipole_flag), ctypes.byref(e) if energy else None, _lib.stream_ptr())
        return e.value if energy else None

    def redistribute(self, f_dipoles, tmat, f64):
        """f64 <- dipole forces (M,4,3) projected through tmat onto the beads."""
        call("hymd_dihedrals_redistribute", self._h, dptr(f_dipoles, torch.float64, "f_dipoles"),
             dptr(tmat, torch.float64, "transfer_matrix"), dptr(f64, torch.float64, "forces"), _lib.stream_ptr())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().hymd_dihedrals_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
