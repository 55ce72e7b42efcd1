"""FP64 rRESPA loop restated from main.py (oracle; test infrastructure only).

integrate_velocity / integrate_position: integrator.py:9-73.
outer_step: main.py:802-887 (kicks, inner bonded loop), 975-1004 (slow
field forces), 1143-1169 (second outer kick).  Thermostat, barostat, PLUMED,
PME and dipoles are not part of this loop.

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


class State:
    def __init__(self, system, ham, single=False):
        self.s = system
        self.ham = ham
        self.single = single
        self.x = np.mod(np.array(system.positions, dtype=np.float64), system.box)
        self.v = np.array(system.velocities, dtype=np.float64)
        self.f_field = self.slow()
        self.f_fast, self.e_fast = self.fast()

    def slow(self):
        s = self.s
        out = field.update_field(self.x, s.types, s.n_types, s.mesh, s.box, s.sigma, self.ham)
        self.phi = out["phi"]
        return field.compute_field_force(self.x, s.types, out["force_mesh"], s.mesh, s.box)

    def fast(self):
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
        return f, e

    def outer_step(self):
        s = self.s
        m, dt, n = s.mass, s.time_step, s.respa_inner
        self.v = integrate_velocity(self.v, self.f_field / m, n * dt)
        for _ in range(n):
            self.v = integrate_velocity(self.v, self.f_fast / m, dt)
            self.x = integrate_position(self.x, self.v, dt)
            self.x = np.mod(self.x, s.box[None, :])
            self.f_fast, self.e_fast = self.fast()
            self.v = integrate_velocity(self.v, self.f_fast / m, dt)
        self.f_field = self.slow()
        self.v = integrate_velocity(self.v, self.f_field / m, n * dt)

    def field_energy(self):
        s = self.s
        return field.field_energy(self.phi, self.ham, s.mesh, s.box)

    def kinetic_energy(self):
        return field.kinetic_energy(self.v, self.s.mass)
