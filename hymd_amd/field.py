"""Drop-in for HyMD's field.py over the HIP particle-field kernels.

Same function names, argument order and in-place semantics as the reference
(field.py:10-703); pmesh ``RealField``/``ComplexField`` objects become torch
device tensors (real grids (Mx,My,Mz) float32, spectra (Mx,My,Mz/2+1)
complex64 with pmesh's 1/G r2c normalisation).  ``layouts`` arguments are
accepted and ignored: one GPU holds the whole mesh.  Every call goes through
``libhymd_amd.so``; there is no CPU path.

Differences from the reference, by design:
* ``v_ext_fourier`` and ``phi_transfer`` are scratch in the reference (their
  contents after update_field are the last type's intermediate spectra);
  they are left untouched here.  ``phi_q_fourier``, ``psi_fourier`` and
  ``elec_field_fourier`` likewise.
* W'(phi) is applied in k-space (exact for the quadratic Hamiltonians); the
  filtered densities ``phi`` and spectra ``phi_fourier`` are still written.
* FP32 grids and forces; positions given in float64 are rounded to float32.
"""
import ctypes
import weakref

import numpy as np
import torch

from . import _lib
from ._lib import call, darr, dptr
from .hamiltonian import quadratic_form

# mesh shape -> most recent ParticleMesh (compute_field_force and
# compute_field_and_kinetic_energy take no pm argument in the reference)
_PM_REGISTRY = weakref.WeakValueDictionary()


class ParticleMesh:
    """Mesh + box + FFT plans + scratch on one GPU (replaces pmesh.ParticleMesh,
    field.py:45-47).  ``max_types`` bounds the species handled per call."""

    def __init__(self, mesh, box, max_types, stream=None):
        _lib.require_gpu()
        self.Nmesh = tuple(int(v) for v in np.broadcast_to(np.asarray(mesh), (3,)))
        self.BoxSize = np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)).copy()
        self.max_types = int(max_types)
        self.device = torch.device("cuda", torch.cuda.current_device())
        self._stream = stream
        h = ctypes.c_void_p()
        mesh_arr = (ctypes.c_int32 * 3)(*self.Nmesh)
        call("hymd_pm_create", ctypes.byref(h), mesh_arr, darr(self.BoxSize, 3), self.max_types,
             _lib.stream_ptr(stream))
        self._h = h
        self.last_field_energy = None
        _PM_REGISTRY[self.real_shape] = self

    # ------------------------------------------------------------------
    @property
    def handle(self):
        call("hymd_pm_set_stream", self._h, _lib.stream_ptr(self._stream))
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().hymd_pm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def real_shape(self):
        return self.Nmesh

    @property
    def complex_shape(self):
        return (self.Nmesh[0], self.Nmesh[1], self.Nmesh[2] // 2 + 1)

    @property
    def volume_per_cell(self):
        return float(np.prod(self.BoxSize)) / float(np.prod(self.Nmesh))

    def set_box(self, box):
        self.BoxSize = np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)).copy()
        call("hymd_pm_set_box", self._h, darr(self.BoxSize, 3))

    def create(self, kind, value=0.0, batch=()):
        """Zero (or constant) real / complex grid(s) on the device (pm.create)."""
        batch = tuple(batch) if isinstance(batch, (tuple, list)) else (int(batch),)
        if kind == "real":
            return torch.full(batch + self.real_shape, float(value), dtype=torch.float32, device=self.device)
        if kind == "complex":
            return torch.full(batch + self.complex_shape, complex(value), dtype=torch.complex64,
                              device=self.device)
        raise ValueError(kind)

    # ------------------------------------------------------------------
    # direct entry points (packed (T,...) tensors)
    # ------------------------------------------------------------------
    def field_update(self, x32, types, n_types, A, b, c0, mass, sigma, force_mesh, phi=None,
                     phi_fourier=None, v_ext=None, want_energy=False):
        n = int(x32.shape[0])
        e = ctypes.c_double(0.0)
        call("hymd_update_field", self.handle, dptr(x32, torch.float32, "positions"),
             dptr(types, torch.int32, "types"), n, int(n_types), darr(np.ravel(A)), darr(b), float(c0),
             darr(mass), float(sigma), dptr(force_mesh, torch.float32, "force_mesh"),
             dptr(phi, torch.float32, "phi"), dptr(phi_fourier, torch.complex64, "phi_fourier"),
             dptr(v_ext, torch.float32, "v_ext"), ctypes.byref(e) if want_energy else None)
        return e.value if want_energy else None

    def readout_forces(self, x32, types, n_types, force_mesh, force):
        call("hymd_compute_field_force", self.handle, dptr(x32, torch.float32, "positions"),
             dptr(types, torch.int32, "types"), int(x32.shape[0]), int(n_types),
             dptr(force_mesh, torch.float32, "force_mesh"), dptr(force, torch.float32, "force"))

    def pme(self, x32, charges, sigma, elec_factor, self_energy, forces, phi_q=None, psi=None,
            elec_field=None, want_energy=False):
        e = ctypes.c_double(0.0)
        call("hymd_update_field_force_q", self.handle, dptr(x32, torch.float32, "positions"),
             dptr(charges, torch.float32, "charges"), int(x32.shape[0]), float(sigma), float(elec_factor),
             float(self_energy), dptr(forces, torch.float32, "elec_forces"), dptr(phi_q, torch.float32, "phi_q"),
             dptr(psi, torch.float32, "psi"), dptr(elec_field, torch.float32, "elec_field"),
             ctypes.byref(e) if want_energy else None)
        return e.value if want_energy else None

    def paint_fixed(self, x32, types, n_types, mass=None, weights=None, out=None, accumulate=False):
        """int64 fixed-point CIC grids (value = round(w m 2^32)), (T or 1, Mx, My, Mz);
        ``accumulate`` adds into ``out`` with atomics instead of overwriting."""
        ng = n_types if types is not None else 1
        if out is None:
            out = torch.zeros((ng,) + self.real_shape, dtype=torch.int64, device=self.device)
        call("hymd_paint_fixed", self.handle, dptr(x32, torch.float32, "positions"),
             dptr(types, torch.int32, "types"), dptr(weights, torch.float32, "weights"), int(x32.shape[0]),
             int(n_types), darr(mass if mass is not None else [1.0] * int(n_types)),
             dptr(out, torch.int64, "grid"), 1 if accumulate else 0)
        return out

    def r2c(self, real):
        """Unnormalised batched r2c of (B, Mx, My, Mz) float32."""
        real = real.contiguous()
        B = real.shape[0] if real.dim() == 4 else 1
        out = torch.empty(((B,) if real.dim() == 4 else ()) + self.complex_shape, dtype=torch.complex64,
                          device=self.device)
        call("hymd_fft_r2c", self.handle, dptr(real, torch.float32), dptr(out, torch.complex64), B)
        return out

    def c2r(self, cplx):
        """Unnormalised batched c2r of (B, Mx, My, Mz/2+1) complex64."""
        cplx = cplx.contiguous()
        B = cplx.shape[0] if cplx.dim() == 4 else 1
        out = torch.empty(((B,) if cplx.dim() == 4 else ()) + self.real_shape, dtype=torch.float32,
                          device=self.device)
        call("hymd_fft_c2r", self.handle, dptr(cplx, torch.complex64), dptr(out, torch.float32), B)
        return out

    def readout(self, grid, x32):
        out = torch.empty(x32.shape[0], dtype=torch.float32, device=self.device)
        call("hymd_readout", self.handle, dptr(grid.contiguous(), torch.float32), dptr(x32, torch.float32),
             int(x32.shape[0]), dptr(out, torch.float32))
        return out

    def field_energy(self, phi, n_types, A, b, c0):
        e = ctypes.c_double(0.0)
        call("hymd_field_energy", self.handle, dptr(phi, torch.float32, "phi"), int(n_types),
             darr(np.ravel(A)), darr(b), float(c0), ctypes.byref(e))
        return e.value

    def field_q_energy(self, phi_q, psi, self_energy):
        e = ctypes.c_double(0.0)
        call("hymd_field_q_energy", self.handle, dptr(phi_q, torch.float32, "phi_q"),
             dptr(psi, torch.float32, "psi"), float(self_energy), ctypes.byref(e))
        return e.value


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def _device_tensor(a, dtype, name):
    if isinstance(a, torch.Tensor):
        if not a.is_cuda:
            a = a.cuda()
        return a.to(dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda").to(dtype).contiguous()


def _positions32(positions):
    x = _device_tensor(positions, torch.float32, "positions")
    if x.dim() != 2 or x.shape[1] != 3:
        raise ValueError("positions must have shape (N, 3)")
    return x


def _flatten(lst):
    out = []
    for item in lst:
        if isinstance(item, (list, tuple)):
            out.extend(_flatten(item))
        else:
            out.append(item)
    return out


def _packed(grids, dtype):
    """Pointer of a list of grids that are consecutive slices of one buffer
    (as initialize_pm creates them), else None."""
    flat = _flatten(grids)
    if not flat:
        return None
    t0 = flat[0]
    if not isinstance(t0, torch.Tensor) or not t0.is_cuda or t0.dtype != dtype:
        return None
    step = t0.numel() * t0.element_size()
    base = t0.data_ptr()
    for i, t in enumerate(flat):
        if (not isinstance(t, torch.Tensor) or t.dtype != dtype or not t.is_contiguous()
                or t.shape != t0.shape or t.data_ptr() != base + i * step):
            return None
    return flat


class _Out:
    """A packed output of ``n`` grids of ``elem_shape``: the caller's grids
    when they are consecutive slices of one buffer, else a temporary copied
    back by finish()."""

    def __init__(self, grids, elem_shape, dtype, device):
        self.grids = grids
        self.tensor = None
        self.flat = None
        if grids is None:
            return
        self.flat = _packed(grids, dtype)
        n = len(_flatten(grids))
        if self.flat is not None:
            f0 = self.flat[0]
            self.tensor = torch.as_strided(f0, (n,) + tuple(elem_shape), (f0.numel(),) + tuple(f0.stride()))
        else:
            self.tensor = torch.empty((n,) + tuple(elem_shape), dtype=dtype, device=device)

    def finish(self):
        if self.grids is not None and self.flat is None:
            for i, g in enumerate(_flatten(self.grids)):
                g.copy_(self.tensor[i])


# ---------------------------------------------------------------------------
# reference call surface
# ---------------------------------------------------------------------------
def initialize_pm(pmesh, config, comm=None):
    """field.py:10-149: (pm, field_list, elec_common_list, coulomb_list).

    ``pmesh`` is ignored (kept for the call signature).  Grids of one kind are
    slices of one packed device buffer so the kernels write them in place.
    """
    T = int(config.n_types)
    pm = ParticleMesh(config.mesh, config.box_size, max_types=T)
    phi_all = pm.create("real", batch=T)
    phif_all = pm.create("complex", batch=T)
    force_all = pm.create("real", batch=(T, 3))
    vext_all = pm.create("real", batch=T)
    vf_all = pm.create("complex", batch=4)
    tr_all = pm.create("complex", batch=3)
    lap_all = pm.create("real", batch=(T, 3))
    phi = [phi_all[t] for t in range(T)]
    phi_fourier = [phif_all[t] for t in range(T)]
    force_on_grid = [[force_all[t, d] for d in range(3)] for t in range(T)]
    v_ext_fourier = [vf_all[i] for i in range(4)]
    v_ext = [vext_all[t] for t in range(T)]
    phi_transfer = [tr_all[i] for i in range(3)]
    phi_laplacian = [[lap_all[t, d] for d in range(3)] for t in range(T)]
    elec_common_list = [None, None, None, None]
    coulomb_list = []
    if config.coulombtype == "PIC_Spectral":
        phi_q = pm.create("real")
        phi_q_fourier = pm.create("complex")
        psi = pm.create("real")
        ef_all = pm.create("real", batch=3)
        elec_field = [ef_all[d] for d in range(3)]
        elec_common_list = [phi_q, phi_q_fourier, psi, elec_field]
        eff_all = pm.create("complex", batch=3)
        coulomb_list = [[eff_all[d] for d in range(3)], pm.create("complex")]
    elif config.coulombtype == "PIC_Spectral_GPE":
        phi_q = pm.create("real")
        phi_q_fourier = pm.create("complex")
        psi = pm.create("real")
        ef_all = pm.create("real", batch=3)
        elec_common_list = [phi_q, phi_q_fourier, psi, [ef_all[d] for d in range(3)]]
        eta_all = pm.create("real", batch=3)
        etaf_all = pm.create("complex", batch=3)
        vb_all = pm.create("real", batch=T)
        vbf_all = pm.create("complex", batch=T)
        fme_all = pm.create("real", batch=(T, 3))
        fmef_all = pm.create("complex", batch=(T, 3))
        coulomb_list = [pm.create("real"), pm.create("complex"), [eta_all[d] for d in range(3)],
                        [etaf_all[d] for d in range(3)], pm.create("real"), pm.create("real"), pm.create("real"),
                        pm.create("real"), [vb_all[t] for t in range(T)], [vbf_all[t] for t in range(T)],
                        [[fme_all[t, d] for d in range(3)] for t in range(T)],
                        [[fmef_all[t, d] for d in range(3)] for t in range(T)]]
    elif config.coulombtype is not None:
        raise NotImplementedError(f"coulombtype {config.coulombtype!r} is not on the HIP path")
    field_list = [phi, phi_fourier, force_on_grid, v_ext_fourier, v_ext, phi_transfer, phi_laplacian]
    return pm, field_list, elec_common_list, coulomb_list


def update_field(phi, phi_laplacian, phi_transfer, layouts, force_mesh, hamiltonian, pm, positions, types,
                 config, v_ext, phi_fourier, v_ext_fourier, m, compute_potential=False):
    """field.py:428-616.  Writes ``phi`` (filtered densities), ``phi_fourier``
    (filtered spectra), ``force_mesh`` and, with ``compute_potential``,
    ``v_ext``; with ``compute_potential`` the field energy is also computed
    and kept as ``pm.last_field_energy``."""
    T = int(config.n_types)
    A, b, c0 = quadratic_form(hamiltonian, T)
    x = _positions32(positions)
    ty = _device_tensor(types, torch.int32, "types")
    mass = list(m) if m is not None else [1.0] * T
    dev = pm.device
    fm = _Out(force_mesh, pm.real_shape, torch.float32, dev)
    ph = _Out(phi, pm.real_shape, torch.float32, dev)
    pf = _Out(phi_fourier, pm.complex_shape, torch.complex64, dev)
    ve = _Out(v_ext if compute_potential else None, pm.real_shape, torch.float32, dev)
    e = pm.field_update(x, ty, T, A, b, c0, mass, config.sigma, fm.tensor, phi=ph.tensor,
                        phi_fourier=pf.tensor, v_ext=ve.tensor, want_energy=compute_potential)
    for o in (fm, ph, pf, ve):
        o.finish()
    pm.last_field_energy = e


def compute_field_force(layouts, r, force_mesh, force, types, n_types):
    """field.py:152-200: force[i] = readout of force_mesh[types[i]] at r[i]."""
    x = _positions32(r)
    ty = _device_tensor(types, torch.int32, "types")
    flat = _packed(force_mesh, torch.float32)
    if flat is not None:
        fm = flat[0]
    else:
        fm = torch.stack([g for g in _flatten(force_mesh)]).contiguous()
    if isinstance(force, torch.Tensor) and force.is_cuda and force.dtype == torch.float32 and force.is_contiguous():
        out = force
    else:
        out = torch.empty((x.shape[0], 3), dtype=torch.float32, device=x.device)
    pm = _pm_for(fm)
    pm.readout_forces(x, ty, int(n_types), fm, out)
    if out is not force:
        if isinstance(force, torch.Tensor):
            force.copy_(out)
        else:
            np.copyto(force, out.cpu().numpy().astype(force.dtype))


def _pm_for(grid):
    """The handle of the grid's mesh shape (the readout only needs the mesh)."""
    pm = _PM_REGISTRY.get(tuple(grid.shape[-3:]))
    if pm is None:
        raise ValueError("no ParticleMesh of this mesh shape exists")
    return pm


def compute_self_energy_q(config, charges, comm=None):
    """field.py:203-238."""
    q = _device_tensor(charges, torch.float32, "charges")
    out = ctypes.c_double(0.0)
    call("hymd_self_energy_q", dptr(q, torch.float32), int(q.numel()), float(config.sigma),
         float(config.coulomb_constant / config.dielectric_const), ctypes.byref(out), _lib.stream_ptr())
    return out.value


def update_field_force_q(charges, phi_q, phi_q_fourier, psi, psi_fourier, elec_field_fourier, elec_field,
                         elec_forces, layout_q, hamiltonian, pm, positions, config):
    """field.py:241-403: writes ``phi_q``, ``psi``, ``elec_field`` and ``elec_forces``."""
    x = _positions32(positions)
    q = _device_tensor(charges, torch.float32, "charges")
    ef = _Out(elec_field, pm.real_shape, torch.float32, pm.device)
    out = elec_forces if (isinstance(elec_forces, torch.Tensor) and elec_forces.is_cuda
                          and elec_forces.dtype == torch.float32 and elec_forces.is_contiguous()) else \
        torch.empty((x.shape[0], 3), dtype=torch.float32, device=pm.device)
    pm.pme(x, q, config.sigma, config.coulomb_constant / config.dielectric_const, 0.0, out,
           phi_q=phi_q, psi=psi, elec_field=ef.tensor)
    ef.finish()
    if out is not elec_forces:
        if isinstance(elec_forces, torch.Tensor):
            elec_forces.copy_(out)
        else:
            np.copyto(elec_forces, out.cpu().numpy().astype(elec_forces.dtype))


CONVERGENCE_TYPES = {"max_diff": 0, "csum": 1, "euclidean_norm": 2}


def update_field_force_q_GPE(conv_fun, phi, types, charges, phi_q, phi_q_fourier, phi_eps, phi_eps_fourier,
                             phi_eta, phi_eta_fourier, phi_pol_prev, phi_pol, elec_field, elec_forces,
                             elec_field_contrib, psi, Vbar_elec, Vbar_elec_fourier, force_mesh_elec,
                             force_mesh_elec_fourier, hamiltonian, layout_q, layouts, pm, positions, config,
                             comm=None):
    """field.py:764-1112 on the device (``hymd_update_field_force_q_gpe``):
    the polarisation iteration runs in the library, with the convergence
    measure of ``config.convergence_type`` ("max_diff", the default, "csum"
    or "euclidean_norm"; ``conv_fun`` is accepted for the signature).  Writes
    ``phi_eps``, ``elec_field_contrib``, ``psi``, ``phi_pol`` (the converged
    polarisation; ``phi_pol_prev`` is read, not changed, as the reference's
    loop rebinds the name), ``Vbar_elec`` and ``elec_forces``; returns
    (Vbar_elec, phi_eps, elec_dot), elec_dot a new grid.  The iteration count
    and last convergence measure are kept as ``pm.gpe_iterations`` /
    ``pm.gpe_delta``, the energy of compute_field_energy_q_GPE as
    ``pm.last_field_q_energy``."""
    T = int(config.n_types)
    x = _positions32(positions)
    ty = _device_tensor(types, torch.int32, "types")
    q = _device_tensor(charges, torch.float32, "charges")
    dev = pm.device
    phi_t = _Out(phi, pm.real_shape, torch.float32, dev)
    if phi_t.flat is None:   # an unpacked list: gather it
        phi_t.tensor.copy_(torch.stack([g for g in _flatten(phi)]))
    vb = _Out(Vbar_elec, pm.real_shape, torch.float32, dev)
    out = elec_forces if (isinstance(elec_forces, torch.Tensor) and elec_forces.is_cuda
                          and elec_forces.dtype == torch.float32 and elec_forces.is_contiguous()) else \
        torch.empty((x.shape[0], 3), dtype=torch.float32, device=dev)
    elec_dot = pm.create("real")
    kind = getattr(config, "convergence_type", None) or "max_diff"
    it = ctypes.c_int32(0)
    delta = ctypes.c_double(0.0)
    energy = ctypes.c_double(0.0)
    eps = [float(v) for v in config.dielectric_type]
    tq = [float(v) for v in (config.type_charges if config.type_charges is not None else [0.0] * T)]
    call("hymd_update_field_force_q_gpe", pm.handle, dptr(x, torch.float32), dptr(ty, torch.int32),
         dptr(q, torch.float32), int(x.shape[0]), T, dptr(phi_t.tensor, torch.float32), darr(eps, T), darr(tq, T),
         float(config.sigma), float(config.coulomb_constant), float(config.conv_crit or 1e-6),
         float(config.pol_mixing if config.pol_mixing is not None else 0.6), CONVERGENCE_TYPES[kind], 100,
         dptr(phi_eps, torch.float32, "phi_eps"), dptr(phi_pol_prev, torch.float32, "phi_pol_prev"),
         dptr(phi_pol, torch.float32, "phi_pol"), dptr(elec_field_contrib, torch.float32, "elec_field_contrib"),
         dptr(psi, torch.float32, "psi"), dptr(elec_dot, torch.float32), dptr(vb.tensor, torch.float32),
         dptr(out, torch.float32), ctypes.byref(energy), ctypes.byref(it), ctypes.byref(delta))
    vb.finish()
    if out is not elec_forces:
        if isinstance(elec_forces, torch.Tensor):
            elec_forces.copy_(out)
        else:
            np.copyto(elec_forces, out.cpu().numpy().astype(elec_forces.dtype))
    pm.gpe_iterations, pm.gpe_delta, pm.last_field_q_energy = it.value, delta.value, energy.value
    return Vbar_elec, phi_eps, elec_dot


def compute_field_energy_q_GPE(config, phi_eps, field_q_energy, dot_elec, comm=None):
    """field.py:706-761: dV (eps_0 / 2) sum phi_eps |E|^2 with eps_0 =
    1/(4 pi k_e), an FP64 sum over the cells."""
    V = float(np.prod(config.box_size))
    G = int(np.prod(np.broadcast_to(np.asarray(config.mesh_size), (3,))))
    eps_0 = 1.0 / (config.coulomb_constant * 4.0 * np.pi)
    return V / G * (0.5 * eps_0) * float(torch.sum(phi_eps.double() * dot_elec.double()))


def compute_field_and_kinetic_energy(phi, phi_q, psi, velocity, hamiltonian, positions, types, v_ext, config,
                                     layouts, comm=None):
    """field.py:619-703: (field_energy, kinetic_energy, field_q_energy), FP64
    real-space reductions on the device."""
    T = int(config.n_types)
    A, b, c0 = quadratic_form(hamiltonian, T)
    flat = _packed(phi, torch.float32)
    ph = flat[0] if flat is not None else torch.stack(list(phi)).contiguous()
    pm = _pm_for(ph)
    if flat is not None:
        ph = torch.as_strided(flat[0], (T,) + pm.real_shape, (flat[0].numel(),) + tuple(flat[0].stride()))
    field_energy = pm.field_energy(ph, T, A, b, c0)
    v = _device_tensor(velocity, velocity.dtype if isinstance(velocity, torch.Tensor) else torch.float64, "v")
    prec = _lib.HYMD_F64 if v.dtype == torch.float64 else _lib.HYMD_F32
    if prec == _lib.HYMD_F32:
        v = v.to(torch.float32)
    ke = ctypes.c_double(0.0)
    call("hymd_kinetic_energy", dptr(v), int(v.shape[0]), prec, float(config.mass), ctypes.byref(ke),
         _lib.stream_ptr())
    field_q_energy = 0.0
    if config.coulombtype == "PIC_Spectral":
        field_q_energy = pm.field_q_energy(phi_q, psi, float(config.self_energy or 0.0))
    return field_energy, ke.value, field_q_energy
