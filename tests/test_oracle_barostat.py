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


# ---------------------------------------------------------------------------
# Dynamics on an ideal gas: P(V) = N R T / V, known relaxation and fluctuations
# ---------------------------------------------------------------------------
def test_bar_is_the_unit_conversion():
    """1 kJ mol^-1 nm^-3 = 1e3 J / N_A / 1e-27 m^3 = 16.6054 bar (barostat.py:135 uses 16.61)."""
    assert OB.BAR == pytest.approx(1e3 / 6.02214076e23 / 1e-27 / 1e5, rel=5e-4)


def _ideal_gas_pressure(n, T0, box):
    """comp_pressure's 18-vector of an ideal gas: P_xx = P_yy = P_zz = N R T / V (kJ mol^-1 nm^-3)."""
    p = np.zeros(18)
    p[15:18] = n * R_GAS * T0 / float(np.prod(box))
    return p


@pytest.mark.parametrize("semi", [False, True])
def test_berendsen_relaxes_an_ideal_gas(semi):
    """Berendsen coupling (barostat.py:45-157) drives an ideal gas to the volume
    where N R T / V = P_t, from above and from below, with the deviation of
    ln V shrinking by (1 - n_b dt beta P_t / tau_p) per call (P_t in bar, beta
    in 1/bar): the pressure relaxes at rate beta K / tau_p with the ideal
    gas's bulk modulus K = P.  A wrong sign diverges; a wrong unit factor
    changes the fixed point or the rate."""
    n, T0, P_t = 4000, 300.0, 1.0e4           # bar
    dt, n_b, tau = 0.3, 2, 2.0
    v_star = n * R_GAS * T0 * OB.BAR / P_t
    rate = 1.0 - n_b * dt * OB.BETA_BERENDSEN * P_t / tau
    for start in (1.25, 0.8):
        box = np.full(3, (start * v_star) ** (1 / 3))
        x = np.ones((4, 3))
        dev = []
        for _ in range(400):
            p = _ideal_gas_pressure(n, T0, box)
            if semi:
                OB.berendsen_semiisotropic(p, x, box, dt, n_b, tau, P_t, P_t, rows_quirk=False)
            else:
                OB.berendsen_isotropic(p, x, box, dt, n_b, tau, P_t)
            dev.append(np.log(np.prod(box) / v_star))
        dev = np.array(dev)
        assert np.all(np.sign(dev) == np.sign(start - 1.0))  # monotone approach from its side
        assert abs(dev[-1]) < 1e-6
        # linear regime: successive ratios equal the predicted factor
        i = np.argmax(np.abs(dev) < 1e-3)
        ratios = dev[i + 1:i + 20] / dev[i:i + 19]
        np.testing.assert_allclose(ratios, rate, rtol=2e-3)


def test_scr_volume_fluctuations_of_an_ideal_gas():
    """SCR (barostat_scr.py:34-132) on an ideal gas at P_t: ln V performs an
    Ornstein-Uhlenbeck walk, drift rate c P_t with c = n_b dt beta / tau_p and
    noise variance 2 n_b^2 R T beta dt / (V tau_p) per call, so its stationary
    variance is n_b R T / (V P_t) with P_t in bar: 1/(16.61 N) of the NPT
    ensemble's 1/N for n_b = 1 (the noise term multiplies R T in kJ/mol by
    beta in 1/bar without the unit factor, and carries n_b^2 where the drift
    carries n_b).  Recorded as the reference's behaviour; a dropped factor 2
    or sqrt, or a lost n_b, fails."""
    rng = np.random.default_rng(2024)
    n, T0, P_t = 500, 300.0, 100.0
    dt, n_b, tau = 0.3, 2, 0.2
    v_star = n * R_GAS * T0 * OB.BAR / P_t
    box = np.full(3, v_star ** (1 / 3))
    x = np.ones((2, 3))
    c = n_b * dt * OB.BETA_SCR / tau
    assert 0.05 < c * P_t < 0.3
    logs = []
    for step in range(60000):
        OB.scr_isotropic(_ideal_gas_pressure(n, T0, box), x, box, dt, n_b, tau, P_t, R_GAS, T0, rng.normal())
        if step >= 2000:
            logs.append(np.log(np.prod(box)))
    logs = np.array(logs)
    s2 = 2 * n_b ** 2 * R_GAS * T0 * OB.BETA_SCR * dt / (v_star * tau)
    a = 1.0 - c * P_t
    expect = s2 / (1.0 - a * a)                 # discrete OU around the fixed point
    assert expect == pytest.approx(n_b * R_GAS * T0 / (v_star * P_t), rel=0.15)
    assert np.mean(logs) == pytest.approx(np.log(v_star), abs=4 * np.sqrt(expect / 1000))
    assert np.var(logs) == pytest.approx(expect, rel=0.1)
