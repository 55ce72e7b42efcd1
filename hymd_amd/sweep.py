"""chi sweep on frozen positions (DefaultWithChi, hamiltonian.py:369-486).

The densities do not depend on chi, kappa or sigma, so one moments pass
(binning, paint, r2c, spectral moments: ``hymd_field_moments``) holds
everything the energy needs: for every point of the sweep E, dE/dchi,
dE/dkappa and dE/dsigma are closed forms of the moments (``autograd``).
The forces of each point come from the held density spectra
(``hymd_field_forces_held``: k-space pass, inverse transforms, readout), and
the gradient of a force loss

    L(chi) = 1/(2N) sum_i |F_i(chi) - F_i^target|^2

goes through the adjoint chain (``hymd_field_adjoint`` with reuse: CIC paint
of dL/dF, r2c, one reduction against the held spectra), so a point costs one
forward force evaluation and one adjoint pass and nothing is re-binned.
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib
from ._lib import call, darr, dptr
from .autograd import energy_and_gradients, field_adjoint, moments


def coupling(chi, kappa, rho0, a):
    """(A, b, c0) of DefaultWithChi (hamiltonian.py:393-412)."""
    chi = np.array(chi, dtype=np.float64)
    np.fill_diagonal(chi, 0.0)
    T = chi.shape[0]
    inv = 1.0 / (kappa * rho0)
    return np.full((T, T), inv) + chi / rho0, np.full(T, -a * inv), 0.5 * a * a * inv


class ChiSweep:
    """Sweep chi matrices over frozen positions ``x32`` (device float32 (N,3))
    and ``types`` (device int32) on the mesh handle ``pm``."""

    def __init__(self, pm, x32, types, n_types, kappa, sigma, rho0=None, a=None, mass=None):
        self.pm = pm
        self.x32, self.types = x32, types
        self.T = int(n_types)
        self.n = int(x32.shape[0])
        self.kappa, self.sigma = float(kappa), float(sigma)
        self.volume = float(np.prod(pm.BoxSize))
        self.rho0 = float(rho0) if rho0 is not None else self.n / self.volume
        self.a = float(a) if a is not None else self.rho0
        self.mass = mass
        t0 = time.perf_counter()
        self.P, self.Q, self.phi0 = moments(pm, x32, types, self.T, self.sigma, mass=mass)
        self.moments_s = time.perf_counter() - t0
        self.force_mesh = pm.create("real", batch=(self.T, 3))

    def energy(self, chi):
        """(E, dE/dchi (per unordered pair), dE/dkappa, dE/dsigma)."""
        E, dchi, dk, ds, _ = energy_and_gradients(self.P, self.Q, self.phi0, chi, self.kappa, self.rho0, self.a,
                                                  self.sigma, self.volume)
        return E, dchi, dk, ds

    def forces(self, chi, out=None):
        """Field forces (device float32 (N,3)) of ``chi`` from the held spectra."""
        A, b, _ = coupling(chi, self.kappa, self.rho0, self.a)
        f = out if out is not None else torch.empty((self.n, 3), dtype=torch.float32, device=self.x32.device)
        call("hymd_field_forces_held", self.pm.handle, self.T, darr(np.ravel(A)), darr(b), self.sigma,
             dptr(self.force_mesh, torch.float32, "force_mesh"), dptr(f, torch.float32, "force"))
        return f

    def force_loss(self, chi, target):
        """L = 1/(2N) sum |F(chi) - target|^2 and dL/dchi (per unordered pair),
        dL/dkappa, dL/dsigma through the adjoint chain."""
        A, _, _ = coupling(chi, self.kappa, self.rho0, self.a)
        f = self.forces(chi)
        g = (f - target) / self.n
        L = 0.5 * self.n * float(torch.sum(g.double() * g.double()))
        dA, ds = field_adjoint(self.pm, None, None, self.T, self.sigma, A, g.contiguous(), mass=self.mass,
                               reuse=True)
        dchi = (dA + dA.T) / self.rho0
        np.fill_diagonal(dchi, 0.0)
        dk = -float(np.sum(dA)) / (self.kappa * self.kappa * self.rho0)
        return L, dchi, dk, ds

    def run(self, chis, target):
        """Every point: energy terms, force loss and its gradients; per-point
        wall time (synchronised)."""
        out = []
        for chi in chis:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            E, dE, dEk, dEs = self.energy(chi)
            L, dL, dLk, dLs = self.force_loss(chi, target)
            torch.cuda.synchronize()
            out.append(dict(E=E, dE_dchi=dE, dE_dkappa=dEk, dE_dsigma=dEs, L=L, dL_dchi=dL, dL_dkappa=dLk,
                            dL_dsigma=dLs, seconds=time.perf_counter() - t0))
        return out
