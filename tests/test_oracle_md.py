"""The oracle's rRESPA loop composition (CPU).

With constant forces an outer step of n inner steps is exact kinematics over
T = n dt under the total acceleration: x(kT) = x0 + kT v0 + (kT)^2 a / 2 and
v(kT) = v0 + kT a (the closed forms of test/test_integrator.py:225-298), for
any split of the force into fast and (several, separately kicked) slow parts.
A dropped or doubled kick, a wrong inner count or a kick of the wrong size
breaks it.  The dipole charges follow main.py:475-483.
"""
import numpy as np
import pytest

from oracle import md as OM


class _ConstForces(OM.State):
    """State whose forces are constant: fast a_f, slow parts a_s (per particle)."""

    def __init__(self, system, fast, slow_parts):
        self._fast = fast
        self._slow = slow_parts
        self.s = system
        self.x = np.array(system.positions, dtype=np.float64)
        self.v = np.array(system.velocities, dtype=np.float64)
        self.f_fast, self.e_fast = self.fast()
        self.f_slow = self.slow()

    def fast(self, dipole_flag=0):
        return self._fast.copy(), {}

    def slow(self):
        return [p.copy() for p in self._slow]


class _Sys:
    def __init__(self, n, inner, dt, mass, box, seed):
        rng = np.random.default_rng(seed)
        self.positions = rng.uniform(0.4, 0.6, (n, 3)) * box
        self.velocities = rng.normal(scale=0.05, size=(n, 3))
        self.respa_inner = inner
        self.time_step = dt
        self.mass = mass
        self.box = np.full(3, box)


@pytest.mark.parametrize("inner", [1, 3, 10])
@pytest.mark.parametrize("n_slow", [1, 3])
def test_outer_step_is_exact_kinematics_for_constant_forces(inner, n_slow):
    rng = np.random.default_rng(inner * 10 + n_slow)
    n, dt, m = 50, 0.01, 72.0
    s = _Sys(n, inner, dt, m, 1e4, inner)
    ff = rng.normal(size=(n, 3))
    fs = [rng.normal(size=(n, 3)) for _ in range(n_slow)]
    st = _ConstForces(s, ff, fs)
    a = (ff + sum(fs)) / m
    T = inner * dt
    x0, v0 = st.x.copy(), st.v.copy()
    for k in range(1, 4):
        st.outer_step()
        t = k * T
        np.testing.assert_allclose(st.x, x0 + t * v0 + 0.5 * t * t * a, rtol=0, atol=1e-10)
        np.testing.assert_allclose(st.v, v0 + t * a, rtol=0, atol=1e-12)


def test_outer_step_wraps_positions_into_the_box():
    n, m = 20, 72.0
    s = _Sys(n, 2, 0.5, m, 1.0, 3)
    s.velocities[:] = 3.7          # crosses the box several times per step
    st = _ConstForces(s, np.zeros((n, 3)), [np.zeros((n, 3))])
    x0 = st.x.copy()
    st.outer_step()
    np.testing.assert_allclose(st.x, np.mod(x0 + 1.0 * 3.7, 1.0), atol=1e-12)
    assert (st.x >= 0).all() and (st.x < 1.0).all()


def test_dipole_charges_follow_the_reference_layout():
    dtype = np.array([1, 1, 0, 2, 1])
    last = np.array([0, 1, 1, 0, 1])
    q = OM.dipole_charges(dtype, last).reshape(-1, 4)
    np.testing.assert_array_equal(q, [[0.25, -0.25, 0, 0], [0.25, -0.25, 0.25, -0.25], [0, 0, 0, 0],
                                      [0, 0, 0, 0], [0.25, -0.25, 0.25, -0.25]])
