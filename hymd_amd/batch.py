"""Batched replicas: energies, forces and chi / kappa / sigma gradients of
many independent particle-field systems from one captured launch sequence.

Each replica r owns a contiguous particle range of one concatenated
position/type array, its own mesh, box and filter width.  One call paints,
transforms and reduces every replica to the spectral moments of
``hymd_amd.autograd.moments`` (bit-identical to running each replica alone),
from which :func:`hymd_amd.autograd.energy_and_gradients_batched` gives the energy
and the exact parameter derivatives of the DefaultWithChi functional
(hamiltonian.py:393-412).  No reference counterpart: the reference runs one
system per process.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import HymdError, call, dptr
from .autograd import energy_and_gradients_batched


def _unpack(raw, R, T):
    npair = T * (T + 1) // 2
    iu, ju = np.triu_indices(T)
    P = np.empty((R, T, T))
    Q = np.empty((R, T, T))
    P[:, iu, ju] = raw[:, :npair]
    P[:, ju, iu] = raw[:, :npair]
    Q[:, iu, ju] = raw[:, npair:2 * npair]
    Q[:, ju, iu] = raw[:, npair:2 * npair]
    return P, Q, raw[:, 2 * npair:]


class ReplicaBatch:
    """Handle over ``len(counts)`` replicas.

    meshes (R,3) ints, boxes (R,3) floats, counts (R,) particles per replica
    (the replicas' particles are consecutive in the arrays passed to
    :meth:`moments`), n_types <= 8, mass per type (default 1).
    """

    def __init__(self, meshes, boxes, counts, n_types, mass=None):
        _lib.require_gpu()
        self.meshes = np.ascontiguousarray(meshes, dtype=np.int32).reshape(-1, 3)
        self.boxes = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 3)
        counts = np.asarray(counts, dtype=np.int64).ravel()
        self.R = len(counts)
        if self.meshes.shape[0] != self.R or self.boxes.shape[0] != self.R:
            raise ValueError("meshes, boxes and counts must describe the same replicas")
        self.offsets = np.zeros(self.R + 1, dtype=np.int64)
        np.cumsum(counts, out=self.offsets[1:])
        self.n = int(self.offsets[-1])
        self.T = int(n_types)
        m = np.ones(self.T) if mass is None else np.asarray(mass, dtype=np.float64)
        self._mass = (ctypes.c_double * self.T)(*m)
        h = ctypes.c_void_p()
        call("hymd_batch_create", ctypes.byref(h), self.R,
             self.meshes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
             self.boxes.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
             self.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), self.T, self._mass)
        self.handle = h
        self._keep = None

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().hymd_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def moments_raw(self, x32, types, sigma, use_graph=True):
        """(R, 2 T(T+1)/2 + T) moments array (the packed layout of hymd_field_moments)."""
        if x32.shape != (self.n, 3) or types.shape != (self.n,):
            raise HymdError(f"expected positions ({self.n}, 3) and types ({self.n},)")
        sig = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma, dtype=np.float64), (self.R,)))
        nv = self.T * (self.T + 1) + self.T
        out = np.empty((self.R, nv), dtype=np.float64)
        xp = dptr(x32, torch.float32, "positions")
        tp = dptr(types, torch.int32, "types")
        self._keep = (xp, tp)  # a captured graph reads these addresses again
        call("hymd_batch_moments", self.handle, xp, tp, sig.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1 if use_graph else 0, _lib.stream_ptr())
        return out

    def moments_arrays(self, x32, types, sigma, use_graph=True):
        """P (R,T,T), Q (R,T,T), phi0 (R,T)."""
        return _unpack(self.moments_raw(x32, types, sigma, use_graph), self.R, self.T)

    def moments(self, x32, types, sigma, use_graph=True):
        """Lists (P_r, Q_r, phi0_r) per replica, as ``autograd.moments``."""
        P, Q, phi0 = self.moments_arrays(x32, types, sigma, use_graph)
        return [(P[r], Q[r], phi0[r]) for r in range(self.R)]

    def _params(self, chi, kappa, sigma, rho0, a):
        vol = np.prod(self.boxes, axis=1)
        counts = np.diff(self.offsets)
        rho0 = counts / vol if rho0 is None else np.broadcast_to(np.asarray(rho0, dtype=np.float64), (self.R,))
        a = rho0 if a is None else np.broadcast_to(np.asarray(a, dtype=np.float64), (self.R,))
        kappa = np.broadcast_to(np.asarray(kappa, dtype=np.float64), (self.R,))
        sigma = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma, dtype=np.float64), (self.R,)))
        chi = np.asarray(chi, dtype=np.float64).reshape(self.R, self.T, self.T)
        return vol, rho0, a, kappa, sigma, chi

    def coupling(self, chi, kappa, rho0):
        """(R, T, T) quadratic forms A_r = 1/(kappa rho0) + chi/rho0 (zero for
        an empty replica, which has no field)."""
        T = self.T
        chi = np.array(chi, dtype=np.float64).reshape(self.R, T, T)
        chi[:, np.arange(T), np.arange(T)] = 0.0
        A = np.zeros((self.R, T, T))
        ok = rho0 > 0.0
        A[ok] = (1.0 / (kappa[ok] * rho0[ok]))[:, None, None] + chi[ok] / rho0[ok][:, None, None]
        return A

    def energy_forces_and_gradients(self, x32, types, chi, kappa, sigma, forces=None, rho0=None, a=None,
                                    use_graph=True):
        """Per-replica (E, dE/dchi (R,T,T), dE/dkappa, dE/dsigma) arrays and
        the field forces of every particle (device float32 (n,3), written into
        ``forces`` when given), all from one launch sequence."""
        if x32.shape != (self.n, 3) or types.shape != (self.n,):
            raise HymdError(f"expected positions ({self.n}, 3) and types ({self.n},)")
        vol, rho0, a, kappa, sigma, chi = self._params(chi, kappa, sigma, rho0, a)
        A = np.ascontiguousarray(self.coupling(chi, kappa, rho0))
        if forces is None:
            forces = torch.empty((self.n, 3), dtype=torch.float32, device=x32.device)
        nv = self.T * (self.T + 1) + self.T
        raw = np.empty((self.R, nv), dtype=np.float64)
        xp = dptr(x32, torch.float32, "positions")
        tp = dptr(types, torch.int32, "types")
        fp = dptr(forces, torch.float32, "forces")
        self._keep_f = (xp, tp, fp)  # a captured graph reads these addresses again
        call("hymd_batch_forces", self.handle, xp, tp, sigma.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
             A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), raw.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
             fp, 1 if use_graph else 0, _lib.stream_ptr())
        P, Q, phi0 = _unpack(raw, self.R, self.T)
        E, dchi, dk, ds = energy_and_gradients_batched(P, Q, phi0, chi, kappa, rho0, a, sigma, vol)
        return E, dchi, dk, ds, forces

    def energy_and_gradients(self, x32, types, chi, kappa, sigma, rho0=None, a=None, use_graph=True):
        """Per-replica (E, dE/dchi (T,T), dE/dkappa, dE/dsigma) arrays.

        chi (R,T,T), kappa (R,), sigma (R,); rho0 and a default to N_r/V_r.
        """
        vol, rho0, a, kappa, sigma, chi = self._params(chi, kappa, sigma, rho0, a)
        P, Q, phi0 = self.moments_arrays(x32, types, sigma, use_graph)
        return energy_and_gradients_batched(P, Q, phi0, chi, kappa, rho0, a, sigma, vol)
