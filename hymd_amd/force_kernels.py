"""The f2py module ``force_kernels`` (setup.py:26-38) on the HIP kernels.

HyMD's force.py binds it (force.py:8-23):

    energy, bond_pr  = cbf(f, r, box, a, b, r0, k)            # + cbf_d
    energy, angle_pr = caf(f, r, box, a, b, c, t0, k)         # + caf_d
    energy           = cdf(force, r, dipoles, transfer_matrix, box, a, b, c, d,
                           coeff, dtype, bb_index, dipole_flag)  # + cdf_d

Same names, argument order and semantics here: outputs (``f``, ``dipoles``,
``transfer_matrix``) are zeroed and overwritten in place; float32 (``cbf``)
or float64 (``cbf_d``) arrays, exactly as f2py demands.  Arrays may be torch
device tensors (used in place) or numpy arrays (copied to the device and
back, like f2py's intent(inout) copies).  The math runs in float64 per term
in both precisions, as the Fortran does; the real(4) variants' float32
differencing of positions is not reproduced (the HIP path differences in
float64, which is more accurate).
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


def cbf(f, r, box, a, b, r0, k):
    """cbf (compute_bond_forces.f90): returns (energy, bond_pr)."""
    return _bond(f, r, box, a, b, r0, k, HYMD_F32)


def cbf_d(f, r, box, a, b, r0, k):
    """cbf_d (compute_bond_forces__double.f90)."""
    return _bond(f, r, box, a, b, r0, k, HYMD_F64)


def caf(f, r, box, a, b, c, t0, k):
    """caf (compute_angle_forces.f90): returns (energy, angle_pr)."""
    return _angle(f, r, box, a, b, c, t0, k, HYMD_F32)


def caf_d(f, r, box, a, b, c, t0, k):
    """caf_d (compute_angle_forces__double.f90)."""
    return _angle(f, r, box, a, b, c, t0, k, HYMD_F64)


def cdf(force, r, dipoles, transfer_matrix, box, a, b, c, d, coeff, dtype, bb_index, dipole_flag):
    """cdf (compute_dihedral_forces.f90 + dipole_reconstruction.f90): returns energy."""
    return _dihedral(force, r, dipoles, transfer_matrix, box, a, b, c, d, coeff, dtype, bb_index, dipole_flag,
                     HYMD_F32)


def cdf_d(force, r, dipoles, transfer_matrix, box, a, b, c, d, coeff, dtype, bb_index, dipole_flag):
    """cdf_d (compute_dihedral_forces__double.f90)."""
    return _dihedral(force, r, dipoles, transfer_matrix, box, a, b, c, d, coeff, dtype, bb_index, dipole_flag,
                     HYMD_F64)
