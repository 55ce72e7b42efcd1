"""FP64 rRESPA loop restated from main.py (oracle; test infrastructure only).

integrate_velocity / integrate_position: integrator.py:9-73.
outer_step: main.py:802-887 (kicks, inner bonded loop with the dihedral
dipoles reconstructed on the last inner step), 975-1095 (slow forces: field,
PME of the charges, PME of the dipole charges redistributed onto the
backbone), 1143-1169 (second outer kick).  Thermostat, barostat and PLUMED
are not part of this loop.

Pinning: the velocity-Verlet / rRESPA closed-form recurrences of
test/test_integrator.py:60-298 (see tests/test_oracle_md.py).
"""
import numpy as np

from . import bonded, field


def integrate_velocity(velocities, accelerations, time_step):
    """integrator.py:41."""
    return velocities + 0.5 * time_step * accelerations


def integrate_position(positions, velocities, time_step):
    """integrator.py:73."""
    return positions + time_step * velocities


COULOMB_CONSTANT = 138.935458  # input_parser.py:160


def dipole_charges(dtype, last):
    """main.py:475-483: +-0.25 on the dipole pair of every CBT term (type 1),
    on both pairs when it is the chain's last term, zero otherwise."""
    q = np.zeros((len(dtype), 4))
    cbt = np.asarray(dtype) == 1
    q[cbt, 0], q[cbt, 1] = 0.25, -0.25
    both = cbt & (np.asarray(last) == 1)
    q[both, 2], q[both, 3] = 0.25, -0.25
    return q.ravel()


class State:
    def __init__(self, system, ham, single=False, dipoles=True):
        s = system
        self.s = s
        self.ham = ham
        self.single = single
        self.x = np.mod(np.array(s.positions, dtype=np.float64), s.box)
        self.v = np.array(s.velocities, dtype=np.float64)
        self.charges = None if s.charges is None or not np.any(s.charges) else np.asarray(s.charges, np.float64)
        self.dihedrals = s.extra.get("dihedrals")
        self.protein = self.dihedrals is not None and np.any(np.asarray(self.dihedrals[5]) == 1) and dipoles
        self.dipoles = None
        if self.protein:
            self.q_dip = dipole_charges(self.dihedrals[5], self.dihedrals[6])
        self.f_fast, self.e_fast = self.fast(dipole_flag=1)
        self.f_slow = self.slow()

    @property
    def f_field(self):
        """All slow forces summed (field + PME + redistributed dipole forces)."""
        return sum(self.f_slow)

    def slow(self):
        """[field forces, PME forces of the charges?, redistributed dipole PME
        forces?] -- kicked separately, in this order (main.py:802-828, 1143-1169)."""
        s = self.s
        out = field.update_field(self.x, s.types, s.n_types, s.mesh, s.box, s.sigma, self.ham)
        self.phi = out["phi"]
        parts = [field.compute_field_force(self.x, s.types, out["force_mesh"], s.mesh, s.box)]
        if self.charges is not None:
            q = field.update_field_force_q(self.x, self.charges, s.mesh, s.box, s.sigma, COULOMB_CONSTANT,
                                           s.dielectric)
            parts.append(q["forces"])
        if self.protein:
            a, b, c, d, _, dtype, last = self.dihedrals
            pos = self.dipoles.reshape(-1, 3)
            qd = field.update_field_force_q(pos, self.q_dip, s.mesh, s.box, s.sigma, COULOMB_CONSTANT,
                                            s.dielectric)
            parts.append(bonded.dipole_forces_redistribution(len(self.x), qd["forces"].reshape(-1, 4, 3),
                                                             self.tmat, a, b, c, d, dtype, last))
        return parts

    def fast(self, dipole_flag=0):
        s = self.s
        f = np.zeros_like(self.x)
        e = {}
        if s.bonds is not None and len(s.bonds[0]):
            fb, eb, _ = bonded.cbf(self.x, s.box, *s.bonds, single=self.single)
            f += fb
            e["bond"] = eb
        if s.angles is not None and len(s.angles[0]):
            fa, ea, _ = bonded.caf(self.x, s.box, *s.angles, single=self.single)
            f += fa
            e["angle"] = ea
        if self.dihedrals is not None and len(self.dihedrals[0]):
            fd, dip, tm, ed = bonded.cdf(self.x, s.box, *self.dihedrals, int(dipole_flag and self.protein),
                                         single=self.single)
            f += fd
            e["dihedral"] = ed
            if dipole_flag and self.protein:
                self.dipoles, self.tmat = dip, tm
        return f, e

    def outer_step(self):
        s = self.s
        m, dt, n = s.mass, s.time_step, s.respa_inner
        for f in self.f_slow:
            self.v = integrate_velocity(self.v, f / m, n * dt)
        for i in range(n):
            self.v = integrate_velocity(self.v, self.f_fast / m, dt)
            self.x = integrate_position(self.x, self.v, dt)
            self.x = np.mod(self.x, s.box[None, :])
            self.f_fast, self.e_fast = self.fast(dipole_flag=1 if i == n - 1 else 0)
            self.v = integrate_velocity(self.v, self.f_fast / m, dt)
        self.f_slow = self.slow()
        for f in self.f_slow:
            self.v = integrate_velocity(self.v, f / m, n * dt)

    def field_energy(self):
        s = self.s
        return field.field_energy(self.phi, self.ham, s.mesh, s.box)

    def kinetic_energy(self):
        return field.kinetic_energy(self.v, self.s.mass)
