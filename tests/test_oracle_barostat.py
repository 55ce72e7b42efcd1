"""The barostat oracle (oracle/barostat.py) against fixed points and
closed forms of the Berendsen and SCR updates, and the meaning of the
reference's indexing."""
import numpy as np
import pytest

from oracle import barostat as OB

R_GAS = 0.0083144621


def _pressure(px, py, pz):
    p = np.zeros(18)
    p[15:18] = [px, py, pz]
    return p / OB.BAR            # the formulas take kJ mol^-1 nm^-3


def test_rows_quirk_scales_whole_rows():
    """`positions[:][0:2] *= a` scales rows 0 and 1 (all columns); `[2]` row 2."""
    x0 = np.arange(30, dtype=np.float64).reshape(10, 3) + 1.0
    x = x0.copy()
    box = np.array([5.0, 5.0, 8.0])
    OB.berendsen_semiisotropic(_pressure(3.0, 5.0, -2.0), x, box, 0.3, 2, 1.5, 1.0, 1.0, rows_quirk=True)
    aL = (1 - 0.3 * 2 / 1.5 * OB.BETA_BERENDSEN * (1.0 - 4.0)) ** (1 / 3)
    aN = (1 - 0.3 * 2 / 1.5 * OB.BETA_BERENDSEN * (1.0 + 2.0)) ** (1 / 3)
    want = x0.copy()
    for r in (0, 1):
        for d in range(3):
            want[r, d] = x0[r, d] * aL
    for d in range(3):
        want[2, d] = x0[2, d] * aN
    np.testing.assert_array_equal(x, want)
    np.testing.assert_allclose(box, [5.0 * aL, 5.0 * aL, 8.0 * aN], rtol=1e-15)
    x = x0.copy()
    box = np.array([5.0, 5.0, 8.0])
    OB.berendsen_semiisotropic(_pressure(3.0, 5.0, -2.0), x, box, 0.3, 2, 1.5, 1.0, 1.0, rows_quirk=False)
    np.testing.assert_allclose(x, x0 * np.array([aL, aL, aN]), rtol=1e-15)


def test_isotropic_pressure_is_the_xy_average():
    x = np.ones((4, 3))
    box = np.ones(3) * 3.0
    a = OB.berendsen_isotropic(_pressure(10.0, 20.0, 1e6), x, box, 0.3, 1, 1.0, 15.0)
    assert a == 1.0                       # P = (10 + 20)/2 = target: z ignored
    a = OB.scr_isotropic(_pressure(10.0, 20.0, 1e6), x, box, 0.3, 1, 1.0, 15.0, R_GAS, 300.0, 0.0)
    assert a == 1.0


def test_berendsen_first_order():
    x = np.ones((2, 3))
    box = np.ones(3)
    dP = 7.0
    a = OB.berendsen_isotropic(_pressure(0.0, 0.0, 0.0), x, box, 0.3, 1, 2.0, dP)
    eps = 0.3 / 2.0 * OB.BETA_BERENDSEN * dP
    assert a == pytest.approx(1 - eps / 3 - eps ** 2 / 9, rel=1e-12)


def test_scr_noise_amplitude():
    """With P = P_t, 3 log(alpha) = sqrt(2 n_b^2 R T beta dt / (V tau_p)) R."""
    n_b, dt, tau, T0, R = 3, 0.3, 2.0, 310.0, 1.7
    box = np.array([4.0, 5.0, 6.0])
    x = np.ones((4, 3))
    a = OB.scr_isotropic(_pressure(1.0, 1.0, 0.0), x, box.copy(), dt, n_b, tau, 1.0, R_GAS, T0, R)
    amp = np.sqrt(2 * n_b ** 2 * R_GAS * T0 * OB.BETA_SCR * dt / (120.0 * tau))
    assert 3 * np.log(a) == pytest.approx(amp * R, rel=1e-12)
    aL, aN, gamma = OB.scr_semiisotropic(_pressure(1.0, 1.0, 1.0), x, box.copy(), dt, n_b, tau, 1.0, 1.0,
                                         R_GAS, T0, 0.0, 0.0)
    assert aL == 1.0 and aN == 1.0 and gamma == 0.0
