"""The thermostat oracle against the reference's regression values
(test/test_thermostat.py:38-137, transcribed to golden/thermostat_goldens.json)
and, for the COM-removal branch the reference tests leave open, against the
closed forms of the CSVR update."""
import json
import os

import numpy as np
import pytest

from oracle import thermostat as OT

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "thermostat_goldens.json")))
CFG = GOLD["config"]


def kinetic(v, m=CFG["mass"]):
    return 0.5 * m * np.sum(v ** 2)


def run(v, names, groups, gaussians, chi_squares, remove_com, **over):
    kw = dict(time_step=CFG["time_step"], respa_inner=CFG["respa_inner"], tau=CFG["tau"],
              target_temperature=CFG["target_temperature"], mass=CFG["mass"])
    kw.update(over)
    return OT.csvr_thermostat(v, names, groups, gaussians=gaussians, chi_squares=chi_squares,
                              remove_com=remove_com, **kw)


def test_fixture_kinetic_energies():
    v = np.array(GOLD["velocities"])
    names = np.array(GOLD["names"])
    assert kinetic(v) == pytest.approx(GOLD["kinetic_energy"], abs=1e-12)
    for t, K in GOLD["kinetic_energy_species"].items():
        assert kinetic(v[names == t]) == pytest.approx(K, abs=1e-13)


@pytest.mark.parametrize("case", range(len(GOLD["cases"])))
def test_csvr_matches_reference_regression(case):
    c = GOLD["cases"][case]
    v = np.array(GOLD["velocities"])
    names = np.array(GOLD["names"])
    work = run(v, names, c["groups"], c["gaussian"], c["chi_squared"], c["remove_com"])
    assert work == pytest.approx(c["work"], abs=1e-13)
    if "kinetic_energy_new" in c:
        assert kinetic(v) == pytest.approx(c["kinetic_energy_new"], abs=1e-13)


def test_csvr_no_com_scales_every_particle_per_group():
    """Without COM removal each group scales ALL velocities (thermostat.py:212-213):
    four groups compose to one uniform factor, and the work is the total change
    of the kinetic energy."""
    c = GOLD["cases"][1]
    v0 = np.array(GOLD["velocities"])
    v = v0.copy()
    work = run(v, np.array(GOLD["names"]), c["groups"], c["gaussian"], c["chi_squared"], False)
    ratio = v / v0
    assert np.allclose(ratio, ratio.flat[0], rtol=1e-14, atol=0)
    assert kinetic(v) - kinetic(v0) == pytest.approx(work, abs=1e-12)


def _com_system(seed=3, n=200):
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, 2.0, (n, 3)) + np.array([0.7, -0.3, 0.2])
    names = rng.choice(np.array(["A", "B", "W"]), n)
    return v, names


def test_csvr_com_identity_at_infinite_tau():
    v0, names = _com_system()
    v = v0.copy()
    work = run(v, names, [["A", "B"], ["W"]], [0.4, -1.1], [300.0, 250.0], True, tau=np.inf)
    assert work == 0.0
    np.testing.assert_allclose(v, v0, rtol=0, atol=1e-14)


def test_csvr_com_zero_tau_closed_form():
    """c = 0: each group's COM-free kinetic energy becomes K_t (S + R^2) / N_f,
    the group COM velocities are unchanged, the other rows untouched."""
    v0, names = _com_system()
    groups = [["A"], ["W"]]
    R, S = [0.4, -1.1], [310.0, 150.0]
    v = v0.copy()
    run(v, names, groups, R, S, True, tau=1e-300)
    for g, grp in enumerate(groups):
        m = np.isin(names, grp)
        n_g = m.sum()
        com0 = v0[m].mean(axis=0)
        com1 = v[m].mean(axis=0)
        np.testing.assert_allclose(com1, com0, rtol=0, atol=1e-13)
        K_t = 1.5 * OT.GAS_CONSTANT * n_g * CFG["target_temperature"]
        assert kinetic(v[m] - com1) == pytest.approx(K_t * (S[g] + R[g] ** 2) / (3 * n_g), rel=1e-13)
    rest = ~np.isin(names, ["A", "W"])
    np.testing.assert_array_equal(v[rest], v0[rest])


def test_csvr_com_single_particle_group_falls_back_to_all():
    """A group of one particle takes the no-COM branch even with COM removal on."""
    v0, names = _com_system(n=50)
    names[7] = "S"
    a = v0.copy()
    b = v0.copy()
    wa = run(a, names, [["S"]], [0.3], [2.5], True)
    wb = run(b, names, [["S"]], [0.3], [2.5], False)
    np.testing.assert_array_equal(a, b)
    assert wa == wb


def test_csvr_empty_group_raises():
    v0, names = _com_system(n=20)
    with pytest.raises(ValueError):
        run(v0, names, [["nobody"]], [0.0], [1.0], True)
    with pytest.raises(ValueError):
        OT.draw_csvr_numbers(np.random.default_rng(0), [3, 0])


def test_draws_follow_the_reference_order():
    """normal then chisquare(3 N_g - 1), per group (thermostat.py:197-198)."""
    R, S = OT.draw_csvr_numbers(np.random.default_rng(11), [4, 9])
    rng = np.random.default_rng(11)
    assert R[0] == rng.normal() and S[0] == rng.chisquare(11)
    assert R[1] == rng.normal() and S[1] == rng.chisquare(26)


def test_cancel_com_momentum():
    v, _ = _com_system()
    OT.cancel_com_momentum(v, len(v))
    np.testing.assert_allclose(v.sum(axis=0), 0.0, atol=1e-12)
