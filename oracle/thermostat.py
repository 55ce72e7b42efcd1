"""CSVR thermostat and centre-of-mass momentum removal (oracle; test
infrastructure only).

csvr_thermostat: thermostat.py:111-219 (Bussi-Parrinello 2008 form).  Groups
are taken in order; for each, R ~ N(0,1) and S ~ chi2(N_f - 1) are drawn (in
that order) and

    alpha^2 = c + (1-c)(S + R^2) K_t / (N_f K) + 2 R sqrt(c (1-c) K_t / (N_f K)),
    c = exp(-dt n_inner / tau),  K_t = 3/2 R_gas N_g T_0,  N_f = 3 N_g,
    work += K (alpha^2 - 1).

With COM removal and N_g > 1 (thermostat.py:177-181, 207-211) K is the
kinetic energy of the group's velocities less the group's COM velocity, and
only the group's COM-free part is scaled.  Otherwise (thermostat.py:182-183,
212-213) K is the kinetic energy of ALL particles and ALL velocities are
scaled, whatever the group: the reference behaves so and this restatement
keeps it.  A group with no particles is an error (the reference's
prng.chisquare(-1) raises ValueError for it).

cancel_com_momentum: thermostat.py:12-15 (v -= sum(v) / n_particles).

Pinning: the regression values of test/test_thermostat.py:38-137 (mocked
random numbers, kinetic energy and thermostat work for one group and for
four groups, no COM removal), transcribed to
tests/golden/thermostat_goldens.json; the COM branch by the closed forms
tau -> inf (c = 1: alpha = 1, no change), tau -> 0 (c = 0: the group's
COM-free kinetic energy becomes K_t (S + R^2) / N_f exactly) and conservation
of each group's COM velocity (tests/test_oracle_thermostat.py).
"""
import numpy as np

GAS_CONSTANT = 0.0083144621  # input_parser.py:159


def group_masks(names, groups):
    """One boolean row mask per coupling group (None/empty -> one group of
    every name present, thermostat.py:176-177)."""
    names = np.asarray(names)
    if not groups or not any(groups):
        groups = [sorted(set(names.tolist()))]
    return [np.isin(names, list(g)) for g in groups]


def csvr_thermostat(velocity, names, groups, time_step, respa_inner, tau, target_temperature, mass,
                    gaussians, chi_squares, remove_com=True, gas_constant=GAS_CONSTANT):
    """Thermostat `velocity` (N,3) in place; returns the work added.

    `gaussians[g]` and `chi_squares[g]` are the numbers drawn for group g."""
    work = 0.0
    c = np.exp(-(time_step * respa_inner) / tau)
    for g, mask in enumerate(group_masks(names, groups)):
        ind = np.nonzero(mask)[0]
        n_g = len(ind)
        if n_g == 0:
            raise ValueError(f"thermostat coupling group {g} has no particles")
        com = remove_com and n_g > 1
        if com:
            com_v = np.sum(velocity[ind], axis=0)
            clean = velocity[ind] - com_v / n_g
            K = 0.5 * mass * np.sum(clean ** 2)
        else:
            K = 0.5 * mass * np.sum(velocity ** 2)
        K_t = 1.5 * gas_constant * n_g * target_temperature
        n_f = 3 * n_g
        R = float(gaussians[g])
        S = float(chi_squares[g])
        alpha2 = (c + (1 - c) * (S + R ** 2) * K_t / (n_f * K)
                  + 2 * R * np.sqrt(c * (1 - c) * K_t / (n_f * K)))
        work += K * (alpha2 - 1)
        alpha = np.sqrt(alpha2)
        if com:
            clean *= alpha
            velocity[ind] = clean + com_v / n_g
        else:
            velocity *= alpha
    return work


def draw_csvr_numbers(prng, group_sizes):
    """The reference's draws for one thermostat call: per group, in order,
    prng.normal() then prng.chisquare(3 N_g - 1) (thermostat.py:197-198)."""
    R, S = [], []
    for g, n_g in enumerate(group_sizes):
        if n_g == 0:
            raise ValueError(f"thermostat coupling group {g} has no particles")
        R.append(prng.normal())
        S.append(prng.chisquare(3 * n_g - 1))
    return R, S


def cancel_com_momentum(velocity, n_particles):
    """thermostat.py:12-15, in place."""
    velocity -= np.sum(velocity, axis=0) / n_particles
    return velocity
