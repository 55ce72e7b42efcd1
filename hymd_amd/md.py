"""rRESPA molecular dynamics on one GPU (main.py:764-1169 without IO; the
thermostat and COM removal of main.py:1314-1321 in ``simulate``).

Per outer step, in the reference's order (main.py:802-887, 975-1004,
1143-1169):

    v += 1/2 (n dt) F_slow/m                         outer half kick
    repeat n times:  v += 1/2 dt F_fast/m ; x += dt v ; x = mod(x, L)
                     F_fast = bonds(x) + angles(x)
                     v += 1/2 dt F_fast/m
    F_slow = readout(update_field(x))                slow (particle-field) forces
    v += 1/2 (n dt) F_slow/m                         outer half kick

Adjacent kicks share one kernel with the drift (each half kick is still its
own FP64 addition, in the reference's order).  Positions and velocities are
float64; forces float32; a float32 copy of x feeds paint and readout.

Systems with charges and/or dihedrals (the peptide/solvent workload) add, as
main.py:508-588, 802-887, 1044-1095 do: the dihedral forces (FP64 kernel,
dipoles reconstructed on the last inner step) to the fast forces, and the
PME forces of the charges plus the dipole-charge PME forces redistributed
onto the backbone (force.py:855-877) to the slow forces, both in ``step()``
and in the library's captured outer step (``run()``).  With dihedrals that
step uses the per-kernel inner loop, whose last sub-step's bonded and
dihedral kernels run on a second stream beside the field update; the fused
inner loop with the dihedral terms (FP64 dihedral arithmetic every sub-step,
dipoles reconstructed at the final positions after it) is there through
``set_fused(True)``.  Charges turn the periodic reorder off (it would leave
the charges behind).
"""
import ctypes

import numpy as np
import torch

from . import _lib, layout
from ._lib import HymdError, call, darr, dptr
from .field import ParticleMesh
from .force import BondedTopology, DihedralTopology
from .synth import hamiltonian_form

COULOMB_CONSTANT = 138.935458  # input_parser.py:160


class RespaMD:
    """rRESPA MD of `system` on the device.

    With ``spatial_order`` (default) the state arrays hold the particles in
    the spatial layout of ``hymd_amd.layout`` (whole molecules ordered by mesh
    cell, one-particle molecules last); ``order`` maps state rows to input
    rows and :meth:`in_input_order` returns a per-particle array in the
    input's order.  With ``reorder_every`` > 0 (and the fused inner loop) the
    library puts the one-particle molecules back in mesh-row order every that
    many outer steps, as they diffuse; ``x``, ``v``, ``types``, the forces and
    the order maps are permuted in place, the physics is unchanged (the field
    path's sums are order-independent and those rows carry no bonded terms).
    ``autotune`` chooses the mesh kernels' launch shapes for this workload
    (``hymd_amd.tune``; ``tuning`` holds the choice and its timings).
    ``term_precision`` sets the arithmetic of the bond and angle terms in the
    library's fused inner loop: "fp32" (default: FP32 term arithmetic, the
    FP32 forces this path allows; the reference's real(4) kernels difference
    positions in FP32 but evaluate the terms in FP64) or "fp64" (as its
    --double-precision kernels); kicks, drifts,
    positions, velocities and energies are FP64 either way, and the Python
    ``step()`` path evaluates the terms in FP64.
    """

    def __init__(self, system, stream=None, spatial_order=True, reorder_every=25, term_precision="fp32",
                 autotune=False):
        _lib.require_gpu()
        self.system = system
        if spatial_order:
            order = layout.spatial_order(system.positions, system.box, system.mesh, system.bonds,
                                         system.angles, system.extra.get("dihedrals"))
            s, inverse = layout.permute_system(system, order)
        else:
            order = np.arange(int(system.n))
            s, inverse = system, order
        self.state_system = s
        self.n = int(s.n)
        self.T = int(s.n_types)
        self.box = np.asarray(s.box, dtype=np.float64)
        self.mass = float(s.mass)
        self.dt = float(s.time_step)
        self.n_inner = int(s.respa_inner)
        self.sigma = float(s.sigma)
        self.A, self.b, self.c0 = hamiltonian_form(s)
        self.type_mass = [1.0] * self.T
        dev = torch.device("cuda", torch.cuda.current_device())
        self._order = torch.as_tensor(np.asarray(order, dtype=np.int64), device=dev)
        self._inverse = torch.as_tensor(np.asarray(inverse, dtype=np.int64), device=dev)
        self.reorder_every = int(reorder_every) if spatial_order else 0
        if term_precision not in ("fp32", "fp64"):
            raise ValueError("term_precision must be 'fp32' or 'fp64'")
        self.term_precision = term_precision
        self.x = torch.as_tensor(np.ascontiguousarray(s.positions, dtype=np.float64), device=dev)
        self.x = torch.remainder(self.x, torch.as_tensor(self.box, device=dev)).contiguous()
        self.v = torch.as_tensor(np.ascontiguousarray(s.velocities, dtype=np.float64), device=dev)
        self.x32 = self.x.to(torch.float32)
        self.types = torch.as_tensor(np.ascontiguousarray(s.types, dtype=np.int32), device=dev)
        self.pm = ParticleMesh(s.mesh, self.box, self.T, stream=stream)
        self.topo = BondedTopology(self.n, bonds=s.bonds, angles=s.angles)
        self.force_mesh = self.pm.create("real", batch=(self.T, 3))
        self.f_field = torch.zeros((self.n, 3), dtype=torch.float32, device=dev)
        self.f_fast = torch.zeros((self.n, 3), dtype=torch.float32, device=dev)
        self._box_c = darr(self.box, 3)
        self.energies = {}
        self._setup_peptide_terms(s, dev)
        # initial forces (main.py:331-371, 508-588): the bonded ones first, the
        # dihedrals with their dipoles, which the slow dipole PME then uses
        self.fast_forces(dipole_flag=1)
        self.slow_forces()
        self._native = None
        self._thermostat = None
        self._com_ws = None
        # launch shapes of the mesh kernels for this workload (hymd_amd.tune),
        # from the tuning cache when this shape was tuned before
        self.tuning = None
        if autotune:
            from . import tune
            self.tuning = tune.autotune(self.pm, self.x32, self.types, self.T, self.A, self.b, self.c0, self.sigma,
                                        mass=self.type_mass)

    @property
    def order(self):
        """State row -> input row (numpy), as of the last step."""
        return self._order.cpu().numpy()

    def in_input_order(self, t):
        """Rows of a per-particle state array (x, v, forces) in the input's order."""
        return t[self._inverse]

    # ------------------------------------------------------------------
    def _setup_peptide_terms(self, s, dev):
        """Charges (PME), dihedrals and their dipoles, when the system has them."""
        q = None if s.charges is None else np.asarray(s.charges, dtype=np.float32)
        self.charges = torch.as_tensor(q, device=dev) if q is not None and np.any(q) else None
        dih = s.extra.get("dihedrals") if getattr(s, "extra", None) else None
        self.dihedrals = None
        self.protein = False
        if self.charges is not None:
            self.elec = COULOMB_CONSTANT / float(s.dielectric)
            self.f_elec = torch.zeros((self.n, 3), dtype=torch.float32, device=dev)
            # compute_self_energy_q (field.py:231-238), once
            self.e_self = self.elec * np.sqrt(1.0 / (2.0 * np.pi * self.sigma ** 2)) * float(
                np.sum(np.asarray(q, dtype=np.float64) ** 2))
        if dih is not None and len(dih[0]):
            a, b, c, d, coeff, dtype, last = dih
            self.dihedrals = DihedralTopology(self.n, a, b, c, d, coeff, dtype, last)
            m4 = self.dihedrals.m
            self.f_ba = torch.zeros((self.n, 3), dtype=torch.float32, device=dev)
            self.f_dih = torch.zeros((self.n, 3), dtype=torch.float64, device=dev)
            self.dipoles = torch.zeros((m4, 4, 3), dtype=torch.float64, device=dev)
            self.tmat = torch.zeros((m4, 6, 3, 3), dtype=torch.float64, device=dev)
            self.protein = bool(np.any(np.asarray(dtype) == 1))
            if self.protein:
                qd = np.zeros((m4, 4), dtype=np.float32)
                cbt = np.asarray(dtype) == 1
                both = cbt & (np.asarray(last) == 1)
                qd[cbt, 0], qd[cbt, 1] = 0.25, -0.25      # main.py:475-483
                qd[both, 2], qd[both, 3] = 0.25, -0.25
                self.q_dip = torch.as_tensor(qd.ravel(), device=dev)
                self.f_dip = torch.zeros((4 * m4, 3), dtype=torch.float32, device=dev)
                self.f_dip64 = torch.zeros((m4, 4, 3), dtype=torch.float64, device=dev)
                self.dip32 = torch.zeros((4 * m4, 3), dtype=torch.float32, device=dev)
                self.f_recon = torch.zeros((self.n, 3), dtype=torch.float64, device=dev)
        self.extra_terms = self.charges is not None or self.dihedrals is not None
        # slow forces kicked together: the field forces alone, or their sum with
        # the PME and redistributed dipole forces
        self.f_slow = self.f_field if not self.extra_terms else torch.zeros_like(self.f_field)

    def _native_handle(self):
        if self._native is None:
            h = ctypes.c_void_p()
            call("hymd_md_create", ctypes.byref(h), self.pm._h, self.topo._h, self.n, self.T, dptr(self.x),
                 dptr(self.v), dptr(self.x32), dptr(self.types), dptr(self.f_field), dptr(self.f_fast),
                 dptr(self.force_mesh), darr(np.ravel(self.A)), darr(self.b), float(self.c0),
                 darr(self.type_mass), self.sigma, self.mass, self.dt, self.n_inner, self._box_c)
            self._native = h
            call("hymd_md_set_term_precision", h, _lib.HYMD_F32 if self.term_precision == "fp32" else _lib.HYMD_F64)
            if self.reorder_every > 0 and self.charges is None and _lib.lib().hymd_md_is_fused(h):
                call("hymd_md_set_reorder", h, self.reorder_every, dptr(self.types, torch.int32, "types"),
                     dptr(self._order, torch.int64, "order"), dptr(self._inverse, torch.int64, "where"))
            if self.extra_terms:
                call("hymd_md_set_peptide", h, dptr(self.charges), self.elec if self.charges is not None else 0.0,
                     self.e_self if self.charges is not None else 0.0,
                     self.dihedrals._h if self.dihedrals is not None else None,
                     self.dihedrals.m if self.dihedrals is not None else 0,
                     dptr(self.q_dip) if self.protein else None, dptr(self.f_slow))
        return self._native

    @property
    def fused(self):
        """True when the library runs the molecule-blocked fused inner loop."""
        return bool(_lib.lib().hymd_md_is_fused(self._native_handle()))

    def set_fused(self, enable):
        """Select the fused (True) or per-kernel (False) inner loop; raises
        HymdError if the topology cannot be fused."""
        call("hymd_md_set_fused", self._native_handle(), 1 if enable else 0)

    def run(self, n_steps, use_graph=True, energies=False):
        """``n_steps`` outer steps enqueued by the library (one ctypes call;
        with ``use_graph`` one captured step is replayed).  With ``energies``
        the last step reports field, bond, angle and kinetic energies."""
        e = (ctypes.c_double * 5)() if energies else None
        h = self._native_handle()
        call("hymd_md_step", h, int(n_steps), 1 if use_graph else 0, e, _lib.stream_ptr())
        x2 = (ctypes.c_double * 2)()
        if self.extra_terms:   # the library's dipoles / transfer matrices are the current ones
            call("hymd_md_peptide_state", h, dptr(self.dipoles) if self.dihedrals is not None else None,
                 dptr(self.tmat) if self.dihedrals is not None else None, x2)
        if energies:
            self.energies.update(field=e[0], bond=e[1], angle=e[2], kinetic=e[3])
            if self.dihedrals is not None:
                self.energies["dihedral"] = x2[0]
            if self.charges is not None:
                self.energies["field_q"] = x2[1]
            return dict(self.energies)
        return None

    def simulate(self, n_steps, config, prng=None, use_graph=True, step0=0):
        """``n_steps`` outer steps with the coupling of main.py:1314-1321: after
        outer step ``step`` (counted from ``step0`` + 1), the CSVR thermostat
        when ``config.target_temperature`` is set and step % config.n_b == 0,
        then COM momentum removal when step % config.cancel_com_momentum == 0.
        The thermostat's draws for all those calls are taken from ``prng`` in
        the reference's order first, then the library runs every step and
        coupling in one call (the thermostat folds the pending outer half kick
        into its passes).  The work accumulates on the device in
        ``thermostat_work``."""
        from .thermostat import CsvrThermostat, _new_workspace
        h = self._native_handle()
        n_steps, step0 = int(n_steps), int(step0)
        n_b = int(config.n_b or 1)
        n_com = int(config.cancel_com_momentum or 0)
        dev = self.v.device
        if config.target_temperature:
            if self._thermostat is None:
                self._thermostat = CsvrThermostat(self.types, config, n_types=self.T)
            th = self._thermostat
            c = float(np.exp(-(config.time_step * config.respa_inner) / config.tau))
            call("hymd_md_set_thermostat", h, self.T, th._type_group, th.n_groups, th._group_n,
                 float(config.gas_constant), float(config.target_temperature), c, 1 if th.remove_com else 0, n_b,
                 dptr(th.work), dptr(th._ws), th._ws.numel() * 8)
            calls = (step0 + n_steps) // n_b - step0 // n_b
            if calls and not config.barostat:
                self._push_draws(h, th, prng, calls)
        else:
            call("hymd_md_set_thermostat", h, self.T, None, 0, None, 0.0, 0.0, 0.0, 0, 0, None, None, 0)
        if n_com:
            if self._com_ws is None:
                self._com_ws = _new_workspace(dev)
            n_particles = int(config.n_particles or self.n)
            call("hymd_md_set_com_removal", h, n_com, n_particles, dptr(self._com_ws), self._com_ws.numel() * 8)
        else:
            call("hymd_md_set_com_removal", h, 0, 0, None, 0)
        call("hymd_md_set_step", h, step0)
        try:
            if config.barostat:
                self._npt_steps(h, n_steps, config, prng, step0)
            else:
                self.run(n_steps, use_graph=use_graph)
        finally:   # plain run() calls stay uncoupled
            call("hymd_md_set_thermostat", h, self.T, None, 0, None, 0.0, 0.0, 0.0, 0, 0, None, None, 0)
            call("hymd_md_set_com_removal", h, 0, 0, None, 0)
        return step0 + n_steps

    @staticmethod
    def _push_draws(h, th, prng, calls):
        draws = [th.draw(prng) for _ in range(calls)]
        R = np.ascontiguousarray([d[0] for d in draws], dtype=np.float64)
        S = np.ascontiguousarray([d[1] for d in draws], dtype=np.float64)
        call("hymd_md_push_draws", h, R.ctypes.data_as(_lib._pd), S.ctypes.data_as(_lib._pd), calls)

    def _npt_steps(self, h, n_steps, config, prng, step0):
        """Outer steps with the barostat of main.py:889-932 between the inner
        loop and the slow forces: each step is the library's two halves; in
        between, on every n_b-th step, the barostat (hymd_amd.barostat) takes
        the pressure from the velocities after the inner loop, the bonded
        virials of the last inner step and the filtered spectra of the last
        field update, scales the positions (FP64 state and float32 mirror) and
        config.box_size, and moves the mesh to the new box; the step's field
        update then runs in that box.  Random numbers are taken from ``prng``
        in the reference's order: the barostat's (SCR) within the step, then
        the thermostat's.  The semi-isotropic coupling scales the x, y columns
        and the z column (``rows_quirk=False``: the state's rows are in the
        spatial order, so the reference's rows 0-2 would be other particles).
        With charges, ``config.type_charges`` gives the electrostatic
        pressure terms (the reference's per-type charges)."""
        from .barostat import isotropic, semiisotropic
        from .hamiltonian import Hamiltonian
        kind = config.barostat.lower()
        if kind not in ("isotropic", "semiisotropic"):
            raise ValueError(f"barostat {config.barostat!r} is not isotropic or semiisotropic")
        baro = isotropic if kind == "isotropic" else semiisotropic
        if getattr(self, "_phi_fourier", None) is None:
            self._phi_fourier = self.pm.create("complex", batch=self.T)
            self._psi = self.pm.create("real") if self.charges is not None else None
            self._phi_q = self.pm.create("real") if self.charges is not None else None
            call("hymd_md_set_spectra", h, dptr(self._phi_fourier), dptr(self._psi), dptr(self._phi_q))
        # the spectra of the current positions, as the last field update left them
        fm = self.pm.create("real", batch=(self.T, 3))
        self.pm.field_update(self.x32, self.types, self.T, self.A, self.b, self.c0, self.type_mass, self.sigma, fm,
                             phi_fourier=self._phi_fourier)
        del fm
        if self.charges is not None:
            f = torch.empty_like(self.f_field)
            self.pm.pme(self.x32, self.charges, self.sigma, self.elec, self.e_self, f, phi_q=self._phi_q, psi=self._psi)

        class _Form(Hamiltonian):
            def __init__(self, A, b, c0):
                self._q = (A, b, c0)

            def quadratic(self):
                return self._q

        ham = _Form(self.A, self.b, self.c0)
        n_b = int(config.n_b or 1)
        config.box_size = np.asarray(self.box, dtype=np.float64).copy()
        if self.charges is not None:
            if config.type_charges is None:
                raise ValueError("the pressure of a charged system needs config.type_charges")
            if not config.coulombtype:
                config.coulombtype = "PIC_Spectral"
        stream = _lib.stream_ptr()
        vir = (ctypes.c_double * 6)()
        for step in range(step0 + 1, step0 + n_steps + 1):
            call("hymd_md_step_half", h, 0, vir, None, stream)
            kw = {} if kind == "isotropic" else {"rows_quirk": False}
            baro(None, self.pm, None, self._phi_q, self._psi, ham, self.x, self.v, config, self._phi_fourier, None,
                 None, np.array(vir[:3]), np.array(vir[3:6]), step, prng, x32=self.x32, **kw)
            self.box = np.asarray(config.box_size, dtype=np.float64).copy()
            self._box_c = darr(self.box, 3)
            if config.target_temperature and step % n_b == 0:
                self._push_draws(h, self._thermostat, prng, 1)
            call("hymd_md_step_half", h, 1, None, None, stream)

    @property
    def thermostat_work(self):
        """Thermostat work of ``simulate`` so far (config.thermostat_work)."""
        return 0.0 if self._thermostat is None else float(self._thermostat.work.item())

    STAGES = {"bonded": 1, "kick_drift": 2, "field": 3, "readout": 4, "pme": 5}

    def time_stage(self, name, max_launches=10000):
        """Record HIP events around every launch of one stage in the eager
        (``use_graph=False``) steps that follow."""
        call("hymd_md_time_stage", self._native_handle(), self.STAGES[name] if name else 0, int(max_launches))

    def stage_time(self):
        """(mean milliseconds per timed launch, launches)."""
        tot = ctypes.c_double(0.0)
        cnt = ctypes.c_int32(0)
        call("hymd_md_stage_time", self._native_handle(), ctypes.byref(tot), ctypes.byref(cnt))
        return (tot.value / cnt.value if cnt.value else float("nan")), cnt.value

    # ------------------------------------------------------------------
    def slow_forces(self, energy=False):
        e = self.pm.field_update(self.x32, self.types, self.T, self.A, self.b, self.c0, self.type_mass,
                                 self.sigma, self.force_mesh, want_energy=energy)
        self.pm.readout_forces(self.x32, self.types, self.T, self.force_mesh, self.f_field)
        if energy:
            self.energies["field"] = e
        if not self.extra_terms:
            return
        self.f_slow.copy_(self.f_field)
        if self.charges is not None:     # field.py:241-403 on the particle charges
            eq = self.pm.pme(self.x32, self.charges, self.sigma, self.elec, self.e_self, self.f_elec,
                             want_energy=energy)
            self.f_slow += self.f_elec
            if energy:
                self.energies["field_q"] = eq
        if self.protein:                 # main.py:1060-1095: PME of the dipole charges, redistributed
            self.dip32.copy_(self.dipoles.reshape(-1, 3))
            self.pm.pme(self.dip32, self.q_dip, self.sigma, self.elec, 0.0, self.f_dip)
            self.f_dip64.copy_(self.f_dip.reshape(-1, 4, 3))
            self.dihedrals.redistribute(self.f_dip64, self.tmat, self.f_recon)
            self.f_slow += self.f_recon.to(torch.float32)

    def fast_forces(self, energy=False, dipole_flag=0):
        target = self.f_fast if self.dihedrals is None else self.f_ba
        r = self.topo.forces(self.x, target, self.box, energies=energy)
        if energy:
            (eb, ea), (pb, pa) = r
            self.energies.update(bond=eb, angle=ea, bond_pr=pb, angle_pr=pa)
        if self.dihedrals is not None:   # compute_dihedral_forces__double.f90 (+ dipoles)
            ed = self.dihedrals.forces(self.x, self.f_dih, self.dipoles, self.tmat, self.box,
                                       1 if (dipole_flag and self.protein) else 0, energy=energy)
            torch.add(self.f_ba, self.f_dih.to(torch.float32), out=self.f_fast)
            if energy:
                self.energies["dihedral"] = ed

    def fused_fast_forces(self, term_precision=None):
        """Bond + angle forces (float32, state order) at the current positions
        as the fused inner loop's kernel evaluates them: one sub-step with
        dt = 0 and zero kicks, on copies of x and v (the state is untouched).
        ``term_precision`` "fp32"/"fp64" (default: the handle's).  Raises
        HymdError when the topology cannot be fused."""
        prec = term_precision or self.term_precision
        h = ctypes.c_void_p()
        b, a = self.topo._b, self.topo._a
        call("hymd_fused_create", ctypes.byref(h), self.n, b[0].numel(), *[dptr(t) for t in b], a[0].numel(),
             *[dptr(t) for t in a])
        try:
            call("hymd_fused_set_precision", h, _lib.HYMD_F32 if prec == "fp32" else _lib.HYMD_F64)
            x = self.x.clone()
            v = torch.zeros_like(self.v)
            x32 = torch.empty_like(self.x32)
            f = torch.zeros_like(self.f_fast)
            call("hymd_fused_inner", h, dptr(x), dptr(v), dptr(x32), dptr(f), dptr(self.f_fast), 0.0, 0.0, 0.0,
                 1, 0, self._box_c, None, None, _lib.stream_ptr())
            torch.cuda.synchronize()
        finally:
            _lib.lib().hymd_fused_destroy(h)
        return f

    def kick_drift(self, f_fast, c_fast, f_slow=None, c_slow=0.0):
        call("hymd_respa_kick_drift", dptr(self.x), dptr(self.v), dptr(self.x32), self.n, dptr(f_fast),
             float(c_fast), dptr(f_slow), float(c_slow), self.dt, self._box_c, _lib.stream_ptr())

    def kick(self, f1, c1, f2=None, c2=0.0):
        call("hymd_respa_kick", dptr(self.v), self.n, dptr(f1), float(c1), dptr(f2), float(c2), _lib.stream_ptr())

    def kinetic_energy(self):
        out = ctypes.c_double(0.0)
        call("hymd_kinetic_energy", dptr(self.v), self.n, _lib.HYMD_F64, self.mass, ctypes.byref(out),
             _lib.stream_ptr())
        return out.value

    # ------------------------------------------------------------------
    def step(self, energies=False, timer=None):
        """One outer rRESPA step.  ``timer`` (optional) is called as
        timer(name, fn) around the bonded-force evaluations."""
        half_in = 0.5 * self.dt / self.mass
        half_out = 0.5 * self.n_inner * self.dt / self.mass
        for i in range(self.n_inner):
            if i == 0:
                self.kick_drift(self.f_fast, half_in, self.f_slow, half_out)
            else:
                self.kick_drift(self.f_fast, half_in, self.f_fast, half_in)
            last_inner = i == self.n_inner - 1
            last = energies and last_inner
            if timer is None:
                self.fast_forces(energy=last, dipole_flag=1 if last_inner else 0)
            else:
                timer("bonded", lambda: self.fast_forces(energy=last, dipole_flag=1 if last_inner else 0))
        self.slow_forces(energy=energies)
        self.kick(self.f_fast, half_in, self.f_slow, half_out)
        if energies:
            self.energies["kinetic"] = self.kinetic_energy()
        return self.energies if energies else None

    def close(self):
        if self._native is not None:
            _lib.lib().hymd_md_destroy(self._native)
            self._native = None
        self.topo.close()
        if self.dihedrals is not None:
            self.dihedrals.close()
        self.pm.close()
