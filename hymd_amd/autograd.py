"""Differentiable particle-field energy (the DefaultWithChi functional).

E(x; chi, kappa, sigma) = dV sum_cells w(phi~) with
w = 1/(2 kappa rho0) (sum_t phi~_t - a)^2 + sum_{t<s} chi_ts phi~_t phi~_s / rho0
(hamiltonian.py:393-412), phi~_t = H_sigma * CIC-paint_t / dV (field.py:570-578).

One device pass (paint, r2c, spectral moments) gives, with the full-spectrum
moments P_ts = sum_k Re(phi~_t^* phi~_s) and Q_ts = sum_k |k|^2 Re(...):

    E        = V [ 1/2 sum_ts A_ts P_ts + b . phi~(0) + c0 ]
    dE/dchi  = V P_ts / rho0                       (per unordered pair t != s)
    dE/dkappa= V [ -1/2 sum_ts P_ts + a sum_t phi~_t(0) - a^2/2 ] / (kappa^2 rho0)
    dE/dsigma= -sigma V sum_ts A_ts Q_ts

with A = 1/(kappa rho0) + chi/rho0, b = -a/(kappa rho0), c0 = a^2/(2 kappa rho0).
These are exact derivatives of the grid energy.  The position gradient
returned by backward is minus the reference's forces (the CIC readout of the
spectral gradient, field.py:607-613 + 152-200), which equals the exact
position gradient of the grid energy only up to O(h^2).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import HymdError, call, darr, dptr


def _unpack_sym(vals, T):
    M = np.zeros((T, T))
    q = 0
    for t in range(T):
        for s in range(t, T):
            M[t, s] = M[s, t] = vals[q]
            q += 1
    return M


def moments(pm, x32, types, n_types, sigma, mass=None, coupling=None, shift=None, force_mesh=None):
    """(P, Q, phi0) of the filtered densities; optionally also the force grids."""
    T = int(n_types)
    np_ = T * (T + 1) // 2
    out = (ctypes.c_double * (2 * np_ + T))()
    call("hymd_field_moments", pm.handle, dptr(x32, torch.float32, "positions"), dptr(types, torch.int32, "types"),
         int(x32.shape[0]), T, darr(mass if mass is not None else [1.0] * T), float(sigma), out,
         darr(np.ravel(coupling)) if coupling is not None else None, darr(shift) if shift is not None else None,
         dptr(force_mesh, torch.float32, "force_mesh"))
    vals = np.array(out[:])
    return _unpack_sym(vals[:np_], T), _unpack_sym(vals[np_:2 * np_], T), vals[2 * np_:]


def energy_and_gradients(P, Q, phi0, chi, kappa, rho0, a, sigma, volume):
    """Energy and exact parameter derivatives from the moments (host, FP64).
    Raises HymdError for rho0 <= 0 or kappa <= 0 (W divides by kappa rho0)."""
    if not (rho0 > 0.0 and kappa > 0.0):
        raise HymdError(f"the field energy needs rho0 > 0 and kappa > 0 (got rho0={rho0}, kappa={kappa})")
    T = P.shape[0]
    chi = np.array(chi, dtype=np.float64)
    np.fill_diagonal(chi, 0.0)
    inv = 1.0 / (kappa * rho0)
    A = np.full((T, T), inv) + chi / rho0
    b = -a * inv
    c0 = 0.5 * a * a * inv
    V = float(volume)
    E = V * (0.5 * float(np.sum(A * P)) + b * float(np.sum(phi0)) + c0)
    dchi = V * P / rho0
    np.fill_diagonal(dchi, 0.0)
    dkappa = V * (-0.5 * float(np.sum(P)) + a * float(np.sum(phi0)) - 0.5 * a * a) / (kappa * kappa * rho0)
    dsigma = -sigma * V * float(np.sum(A * Q))
    return E, dchi, dkappa, dsigma, A


def energy_and_gradients_batched(P, Q, phi0, chi, kappa, rho0, a, sigma, volume):
    """energy_and_gradients over R replicas at once: P, Q, chi (R,T,T), phi0
    (R,T), the rest (R,).  Returns arrays E, dchi (R,T,T), dkappa, dsigma.
    An empty replica (rho0 = 0: no particles) has no field and gets zero
    energy and gradients; kappa <= 0 raises HymdError."""
    T = P.shape[1]
    chi = np.array(chi, dtype=np.float64)
    chi[:, np.arange(T), np.arange(T)] = 0.0
    kappa, rho0, a, sigma, V = (np.asarray(z, dtype=np.float64) for z in (kappa, rho0, a, sigma, volume))
    if np.any(~(kappa > 0.0)) or np.any(~(rho0 >= 0.0)):
        raise HymdError("the field energy needs kappa > 0 and rho0 >= 0 for every replica")
    empty = rho0 == 0.0
    rho0 = np.where(empty, 1.0, rho0)
    inv = 1.0 / (kappa * rho0)
    A = inv[:, None, None] + chi / rho0[:, None, None]
    s0 = phi0.sum(axis=1)
    E = V * (0.5 * np.einsum("rts,rts->r", A, P) - a * inv * s0 + 0.5 * a * a * inv)
    dchi = (V / rho0)[:, None, None] * P
    dchi[:, np.arange(T), np.arange(T)] = 0.0
    dkappa = V * (-0.5 * P.sum(axis=(1, 2)) + a * s0 - 0.5 * a * a) / (kappa * kappa * rho0)
    dsigma = -sigma * V * np.einsum("rts,rts->r", A, Q)
    for arr in (E, dchi, dkappa, dsigma):
        arr[empty] = 0.0
    return E, dchi, dkappa, dsigma


class FieldEnergy(torch.autograd.Function):
    """E = FieldEnergy.apply(x, chi, kappa, sigma, pm, types, rho0, a).

    x (N,3) float32/float64 device tensor, chi (T,T) symmetric float64 tensor
    (diagonal ignored), kappa and sigma 0-d float64 tensors.  Backward gives
    d/dx = -F (reference forces), d/dchi with dE/dchi_ts = V P_ts/(2 rho0) per
    entry (a symmetric perturbation of both entries gives V P_ts/rho0),
    d/dkappa, d/dsigma.
    """

    @staticmethod
    def forward(ctx, x, chi, kappa, sigma, pm, types, rho0, a):
        T = chi.shape[0]
        x32 = x.detach().to(torch.float32).contiguous()
        ty = types.to(torch.int32).contiguous()
        k = float(kappa)
        sg = float(sigma)
        chi_np = chi.detach().cpu().numpy().astype(np.float64)
        chi_np = 0.5 * (chi_np + chi_np.T)
        np.fill_diagonal(chi_np, 0.0)
        inv = 1.0 / (k * rho0)
        A = np.full((T, T), inv) + chi_np / rho0
        b = np.full(T, -a * inv)
        fm = pm.create("real", batch=(T, 3))
        P, Q, phi0 = moments(pm, x32, ty, T, sg, coupling=A, shift=b, force_mesh=fm)
        E, dchi, dk, ds, _ = energy_and_gradients(P, Q, phi0, chi_np, k, rho0, a, sg, np.prod(pm.BoxSize))
        F = torch.empty((x32.shape[0], 3), dtype=torch.float32, device=x32.device)
        pm.readout_forces(x32, ty, T, fm, F)
        ctx.save_for_backward(F)
        ctx.grads = (torch.as_tensor(0.5 * dchi, dtype=chi.dtype, device=chi.device),
                     torch.as_tensor(dk, dtype=torch.float64, device=chi.device),
                     torch.as_tensor(ds, dtype=torch.float64, device=chi.device))
        ctx.x_dtype = x.dtype
        return torch.as_tensor(E, dtype=torch.float64, device=x.device)

    @staticmethod
    def backward(ctx, grad_e):
        (F,) = ctx.saved_tensors
        g = grad_e.to(torch.float64)
        dchi, dk, ds = ctx.grads
        dx = (-F.to(torch.float64) * g).to(ctx.x_dtype)
        return dx, dchi * g.to(dchi.device), dk * g, ds * g, None, None, None, None


def field_energy(x, chi, kappa, sigma, pm, types, rho0, a):
    """Differentiable field energy (see FieldEnergy)."""
    return FieldEnergy.apply(x, chi, kappa, sigma, pm, types, rho0, a)
