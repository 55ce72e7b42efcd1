"""Drop-in for the compiled bonded kernels of HyMD's force.py.

The reference binds the f2py module ``force_kernels`` (force.py:8-23):

    energy, bond_pr  = cbf(f, r, box, a, b, r0, k)            # + cbf_d
    energy, angle_pr = caf(f, r, box, a, b, c, t0, k)         # + caf_d
    energy           = cdf(force, r, dipoles, transfer_matrix, box, a, b, c, d,
                           coeff, dtype, bb_index, dipole_flag)  # + cdf_d

Same names, argument order and semantics here, on the HIP kernels: outputs
(``f``, ``dipoles``, ``transfer_matrix``) are zeroed and overwritten in place;
float32 (``cbf``) or float64 (``cbf_d``) arrays, exactly as f2py demands.
Arrays may be torch device tensors (used in place) or numpy arrays (copied to
the device and back, like f2py's intent(inout) copies).  The math runs in
float64 per term in both precisions, as the Fortran does; the real(4)
variants' float32 differencing of positions is not reproduced (the HIP path
differences in float64, which is more accurate).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import HYMD_F32, HYMD_F64, call, darr, dptr


def _to_dev(a, dtype):
    if isinstance(a, torch.Tensor):
        t = a if a.is_cuda else a.cuda()
        return t.to(dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda").to(dtype).contiguous()


class _InOut:
    """An in/out array: device tensors of the right dtype are written in place;
    anything else goes through a device temporary copied back at the end."""

    def __init__(self, arr, dtype, shape, name):
        self.arr = arr
        self.np_dtype = np.float32 if dtype == torch.float32 else np.float64
        if isinstance(arr, torch.Tensor):
            if arr.dtype != dtype:
                raise TypeError(f"{name} must be {dtype} (f2py intent(inout) requires the exact type)")
            if tuple(arr.shape) != tuple(shape):
                raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(arr.shape)}")
            self.dev = arr if (arr.is_cuda and arr.is_contiguous()) else torch.empty(shape, dtype=dtype,
                                                                                       device="cuda")
        else:
            a = np.asarray(arr)
            if a.dtype != self.np_dtype:
                raise TypeError(f"{name} must be {self.np_dtype.__name__} (f2py intent(inout) requires the exact type)")
            if a.shape != tuple(shape):
                raise ValueError(f"{name} must have shape {tuple(shape)}, got {a.shape}")
            self.dev = torch.empty(shape, dtype=dtype, device="cuda")

    def finish(self):
        if self.dev is self.arr:
            return
        if isinstance(self.arr, torch.Tensor):
            self.arr.copy_(self.dev)
        else:
            np.copyto(self.arr, self.dev.cpu().numpy())


def _box(box):
    b = np.asarray(box.cpu() if isinstance(box, torch.Tensor) else box, dtype=np.float64).reshape(-1)
    return darr(b, 3)


def _bond(f, r, box, a, b, r0, k, prec):
    dt = torch.float32 if prec == HYMD_F32 else torch.float64
    rr = _to_dev(r, dt)
    n = rr.shape[0]
    fo = _InOut(f, dt, (n, 3), "f")
    ai, bi = _to_dev(a, torch.int32), _to_dev(b, torch.int32)
    r0d, kd = _to_dev(r0, torch.float64), _to_dev(k, torch.float64)
    e = ctypes.c_double(0.0)
    pr = (ctypes.c_double * 3)()
    call("hymd_cbf", dptr(fo.dev), dptr(rr), n, _box(box), dptr(ai), dptr(bi), dptr(r0d), dptr(kd), ai.numel(), prec,
         ctypes.byref(e), pr, _lib.stream_ptr())
    fo.finish()
    return e.value, np.array(list(pr), dtype=np.float32 if prec == HYMD_F32 else np.float64)


def _angle(f, r, box, a, b, c, t0, k, prec):
    dt = torch.float32 if prec == HYMD_F32 else torch.float64
    rr = _to_dev(r, dt)
    n = rr.shape[0]
    fo = _InOut(f, dt, (n, 3), "f")
    ai, bi, ci = _to_dev(a, torch.int32), _to_dev(b, torch.int32), _to_dev(c, torch.int32)
    t0d, kd = _to_dev(t0, torch.float64), _to_dev(k, torch.float64)
    e = ctypes.c_double(0.0)
    pr = (ctypes.c_double * 3)()
    call("hymd_caf", dptr(fo.dev), dptr(rr), n, _box(box), dptr(ai), dptr(bi), dptr(ci), dptr(t0d), dptr(kd),
         ai.numel(), prec, ctypes.byref(e), pr, _lib.stream_ptr())
    fo.finish()
    return e.value, np.array(list(pr), dtype=np.float32 if prec == HYMD_F32 else np.float64)


def _dihedral(force, r, dipoles, transfer_matrix, box, a, b, c, d, coeff, dtype, bb_index, dipole_flag, prec,
              clamps=None):
    dt = torch.float32 if prec == HYMD_F32 else torch.float64
    rr = _to_dev(r, dt)
    n = rr.shape[0]
    ai = _to_dev(a, torch.int32)
    m = ai.numel()
    fo = _InOut(force, dt, (n, 3), "force")
    do = _InOut(dipoles, dt, (m, 4, 3), "dipoles")
    to = _InOut(transfer_matrix, dt, (m, 6, 3, 3), "transfer_matrix")
    bi, ci, di = _to_dev(b, torch.int32), _to_dev(c, torch.int32), _to_dev(d, torch.int32)
    cf = _to_dev(coeff, torch.float32).reshape(m, 6, 5).contiguous()
    ty, bb = _to_dev(dtype, torch.int32), _to_dev(bb_index, torch.int32)
    e = ctypes.c_double(0.0)
    nc = ctypes.c_int64(0)
    call("hymd_cdf", dptr(fo.dev), dptr(rr), dptr(do.dev), dptr(to.dev), n, _box(box), dptr(ai), dptr(bi), dptr(ci),
         dptr(di), dptr(cf), dptr(ty), dptr(bb), m, int(dipole_flag), prec, ctypes.byref(e), ctypes.byref(nc),
         _lib.stream_ptr())
    for o in (fo, do, to):
        o.finish()
    if clamps is not None:
        clamps.append(nc.value)
    return e.value


def compute_bond_forces__fortran(f, r, box, a, b, r0, k):
    """cbf (compute_bond_forces.f90): returns (energy, bond_pr)."""
    return _bond(f, r, box, a, b, r0, k, HYMD_F32)


def compute_bond_forces__fortran__double(f, r, box, a, b, r0, k):
    """cbf_d (compute_bond_forces__double.f90)."""
    return _bond(f, r, box, a, b, r0, k, HYMD_F64)


def compute_angle_forces__fortran(f, r, box, a, b, c, t0, k):
    """caf (compute_angle_forces.f90): returns (energy, angle_pr)."""
    return _angle(f, r, box, a, b, c, t0, k, HYMD_F32)


def compute_angle_forces__fortran__double(f, r, box, a, b, c, t0, k):
    """caf_d (compute_angle_forces__double.f90)."""
    return _angle(f, r, box, a, b, c, t0, k, HYMD_F64)


def compute_dihedral_forces__fortran(force, r, dipoles, transfer_matrix, box, a, b, c, d, coeff, dtype, bb_index,
                                     dipole_flag):
    """cdf (compute_dihedral_forces.f90 + dipole_reconstruction.f90): returns energy."""
    return _dihedral(force, r, dipoles, transfer_matrix, box, a, b, c, d, coeff, dtype, bb_index, dipole_flag,
                     HYMD_F32)


def compute_dihedral_forces__fortran__double(force, r, dipoles, transfer_matrix, box, a, b, c, d, coeff, dtype,
                                             bb_index, dipole_flag):
    """cdf_d (compute_dihedral_forces__double.f90)."""
    return _dihedral(force, r, dipoles, transfer_matrix, box, a, b, c, d, coeff, dtype, bb_index, dipole_flag,
                     HYMD_F64)


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
