"""Berendsen and stochastic-cell-rescaling barostats (oracle; test
infrastructure only).

berendsen_isotropic / berendsen_semiisotropic: barostat.py:45-157, 160-309.
scr_isotropic / scr_semiisotropic: barostat_scr.py:34-132, 135-284.

Given the 18-vector of comp_pressure (kJ mol^-1 nm^-3, x 16.61 -> bar),
every n_b-th step they scale config.box_size and the positions:

  Berendsen  alpha = [1 - dt n_inner n_b / tau_p beta (P_t - P)]^(1/3),  beta = 4.6e-5 / bar
  SCR        log alpha = -n_b dt n_inner beta / tau_p (P_t - P) + sqrt(2 n_b R T beta dt n_inner n_b / (V tau_p)) R,
             alpha = exp(log alpha / 3),  beta = 7.6e-4 / bar  (isotropic; the semi-isotropic L and N
             forms of barostat_scr.py:190-282 with the surface tension L_z/2 (P_N - P_L))

with the reference's indexing kept as it is:
* isotropic P is np.average(pressure[-3:-1]): the x and y totals only;
* semi-isotropic `positions[:][0:2] *= alpha_L` scales all three
  coordinates of rows 0 and 1, and `positions[:][2] *= alpha_N` those of
  row 2 (numpy: positions[:] is the whole array, [0:2] its first two rows);
  ``rows_quirk=False`` applies the per-axis scaling the docstrings describe.
SCR draws prng.normal() once (isotropic) or twice, Rxy then Rz
(semi-isotropic), every n_b-th step.

Pinning (tests/test_oracle_barostat.py): numpy's own evaluation of the
reference's indexing expressions on arrays, P = P_t gives alpha = 1 (SCR:
with R = 0), the unit factor against the physical constants, and dynamics on
an ideal gas P = N R T / V: Berendsen coupling relaxes ln V to N R T / P_t at
the rate n_b dt beta P_t / tau_p per call, from either side (isotropic and
semi-isotropic), and SCR's stationary variance of ln V is
n_b R T / (V P_t[bar]) (the reference's noise term carries no unit factor and
n_b^2, so this is 1/(16.61 N) at n_b = 1 instead of the NPT ensemble's 1/N).
"""
import numpy as np

BAR = 16.61              # kJ mol^-1 nm^-3 -> bar (barostat.py:135)
BETA_BERENDSEN = 4.6e-5  # bar^-1 (barostat.py:121)
BETA_SCR = 7.6e-4        # bar^-1 (barostat_scr.py:74)


def _scale_rows(positions, lateral, normal, rows_quirk):
    if rows_quirk:
        if lateral is not None:
            positions[:][0:2] *= lateral
        if normal is not None:
            positions[:][2] *= normal
    else:
        if lateral is not None:
            positions[:, 0:2] *= lateral
        if normal is not None:
            positions[:, 2] *= normal


def berendsen_isotropic(pressure, positions, box, dt_outer, n_b, tau_p, P_target):
    P = np.average(pressure[-3:-1]) * BAR
    alpha = (1.0 - dt_outer * n_b / tau_p * BETA_BERENDSEN * (P_target - P)) ** (1 / 3)
    box *= alpha
    positions *= alpha
    return alpha


def berendsen_semiisotropic(pressure, positions, box, dt_outer, n_b, tau_p, P_L, P_N, rows_quirk=True):
    PL = (pressure[-3] + pressure[-2]) / 2 * BAR
    PN = pressure[-1] * BAR
    aL = aN = 1.0
    if P_L:
        aL = (1.0 - dt_outer * n_b / tau_p * BETA_BERENDSEN * (P_L - PL)) ** (1 / 3)
        box[0:2] *= aL
        _scale_rows(positions, aL, None, rows_quirk)
    if P_N:
        aN = (1.0 - dt_outer * n_b / tau_p * BETA_BERENDSEN * (P_N - PN)) ** (1 / 3)
        box[2] *= aN
        _scale_rows(positions, None, aN, rows_quirk)
    return aL, aN


def scr_isotropic(pressure, positions, box, dt_outer, n_b, tau_p, P_target, gas_constant, T0, R):
    P = np.average(pressure[-3:-1]) * BAR
    V = np.prod(box)
    noise = np.sqrt(2.0 * n_b * gas_constant * T0 * BETA_SCR * dt_outer * n_b / (V * tau_p)) * R
    log_alpha = -n_b * dt_outer * BETA_SCR / tau_p * (P_target - P) + noise
    alpha = np.exp(log_alpha / 3.0)
    box *= alpha
    positions *= alpha
    return alpha


def scr_semiisotropic(pressure, positions, box, dt_outer, n_b, tau_p, P_L, P_N, gas_constant, T0, Rxy, Rz,
                      rows_quirk=True):
    PL = (pressure[-3] + pressure[-2]) / 2 * BAR
    PN = pressure[-1] * BAR
    gamma = box[2] / 2 * (PN - PL)
    aL = aN = 1.0
    if P_L:
        V = np.prod(box)
        noise = np.sqrt(4.0 * n_b * gas_constant * T0 * BETA_SCR * dt_outer * n_b / (3 * V * tau_p)) * Rxy
        log_alpha = -2.0 * n_b * dt_outer * BETA_SCR / (3 * tau_p) * (P_L - PL - gamma / box[2]) + noise
        aL = np.exp(log_alpha / 2.0)
        box[0:2] *= aL
        _scale_rows(positions, aL, None, rows_quirk)
    if P_N:
        V = np.prod(box)
        noise = np.sqrt(2.0 * n_b * gas_constant * T0 * BETA_SCR * dt_outer * n_b / (3 * V * tau_p)) * Rz
        log_alpha = -n_b * dt_outer * BETA_SCR / (3 * tau_p) * (P_N - PN) + noise
        aN = np.exp(log_alpha / 1.0)
        box[2] *= aN
        _scale_rows(positions, None, aN, rows_quirk)
    return aL, aN, gamma
