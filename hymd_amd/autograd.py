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

The forces F_i = readout(c2r(-i k H A H rho^))(x_i) are linear in A, so for
a loss L(E, F) the force part of dL/dA is one adjoint pass
(``hymd_field_adjoint``): the upstream gradient dL/dF is CIC-painted per type
and component (paint is the readout's adjoint), transformed, and reduced
against the density spectra.  With A = 1/(kappa rho0) + chi/rho0:
dL/dchi_ts = (dL/dA_ts + dL/dA_st)/rho0 per unordered pair, dL/dkappa =
-sum_ts dL/dA_ts / (kappa^2 rho0), and dL/dsigma adds the adjoint's sigma
term (both H factors of the force spectrum depend on sigma).
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


def field_adjoint(pm, x32, types, n_types, sigma, coupling, grad_force, mass=None, reuse=False):
    """dL/dA (T, T) and dL/dsigma for the upstream gradient ``grad_force`` =
    dL/dF (device float32 (N, 3)) of the field forces of (coupling, sigma).
    ``reuse``: the binning and density spectra of the last ``moments`` call
    on ``pm`` (same positions and types) are used."""
    T = int(n_types)
    dA = (ctypes.c_double * (T * T))()
    ds = ctypes.c_double(0.0)
    n = int(grad_force.shape[0])
    call("hymd_field_adjoint", pm.handle, None if reuse else dptr(x32, torch.float32, "positions"),
         None if reuse else dptr(types, torch.int32, "types"), n, T,
         darr(mass if mass is not None else [1.0] * T), float(sigma), darr(np.ravel(coupling)),
         dptr(grad_force.contiguous(), torch.float32, "grad_force"), 1 if reuse else 0, dA, ctypes.byref(ds))
    return np.array(dA[:]).reshape(T, T), ds.value


class FieldEnergyForces(torch.autograd.Function):
    """(E, F) = FieldEnergyForces.apply(x, chi, kappa, sigma, pm, types, rho0, a).

    E as FieldEnergy; F (N, 3) float32 the field forces (compute_field_force of
    update_field).  Backward for a loss L(E, F): dL/dchi (per entry, as
    FieldEnergy), dL/dkappa and dL/dsigma combine the energy's moments with
    the adjoint pass over dL/dF; dL/dx is -F dL/dE.  The force output's
    contribution to dL/dx (a Hessian product) is not provided: backward raises
    HymdError when x needs a gradient and dL/dF is non-zero."""

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
        ctx.save_for_backward(F, x32, ty)
        ctx.pm, ctx.T, ctx.A, ctx.sigma, ctx.kappa, ctx.rho0 = pm, T, A, sg, k, rho0
        ctx.grads = (0.5 * dchi, dk, ds)
        ctx.x_dtype = x.dtype
        ctx.chi_meta = (chi.dtype, chi.device)
        return torch.as_tensor(E, dtype=torch.float64, device=x.device), F

    @staticmethod
    def backward(ctx, grad_e, grad_f):
        F, x32, ty = ctx.saved_tensors
        ge = 0.0 if grad_e is None else float(grad_e)
        dchi_e, dk_e, ds_e = ctx.grads
        dchi = ge * dchi_e
        dk = ge * dk_e
        ds = ge * ds_e
        has_f = grad_f is not None and bool(torch.any(grad_f != 0))
        if has_f:
            if ctx.needs_input_grad[0]:
                raise HymdError("dL/dx through the force output (a Hessian product) is not provided")
            dA, dsig = field_adjoint(ctx.pm, x32, ty, ctx.T, ctx.sigma, ctx.A, grad_f.to(torch.float32))
            sym = (dA + dA.T) / ctx.rho0
            np.fill_diagonal(sym, 0.0)
            dchi = dchi + 0.5 * sym
            dk = dk - float(np.sum(dA)) / (ctx.kappa * ctx.kappa * ctx.rho0)
            ds = ds + dsig
        dx = None
        if ctx.needs_input_grad[0]:
            dx = (-F.to(torch.float64) * ge).to(ctx.x_dtype)
        cdt, cdev = ctx.chi_meta
        return (dx, torch.as_tensor(dchi, dtype=cdt, device=cdev), torch.as_tensor(dk, dtype=torch.float64),
                torch.as_tensor(ds, dtype=torch.float64), None, None, None, None)


def field_energy_forces(x, chi, kappa, sigma, pm, types, rho0, a):
    """Differentiable field energy and forces (see FieldEnergyForces)."""
    return FieldEnergyForces.apply(x, chi, kappa, sigma, pm, types, rho0, a)
