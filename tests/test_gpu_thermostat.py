"""Device CSVR thermostat and COM momentum removal against the oracle
(oracle/thermostat.py) and the reference's regression values
(test/test_thermostat.py:38-137, golden/thermostat_goldens.json).

Tolerances (FP64, the device sums in another order than numpy's): kinetic
energies and work to 1e-13 relative, velocities to 1e-13 of max |v|; the
reference's own regression test asserts 1e-13 absolute on values of ~170,
which a different summation order of the same FP64 terms keeps well inside
1e-12 absolute.
"""
import json
import os

import numpy as np
import pytest
import torch

from oracle import thermostat as OT

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "thermostat_goldens.json")))
NAMES = ["A", "B", "C", "D"]


def _config(groups=None, **over):
    from hymd_amd.config import Config
    c = GOLD["config"]
    cfg = Config(n_steps=0, time_step=c["time_step"], mesh_size=[2, 2, 2], sigma=0.5, kappa=0.05,
                 mass=c["mass"], respa_inner=c["respa_inner"], target_temperature=c["target_temperature"],
                 tau=c["tau"], unique_names=list(NAMES), n_types=len(NAMES),
                 thermostat_coupling_groups=groups or [])
    for k, val in over.items():
        setattr(cfg, k, val)
    return cfg


class _Mock:
    """random_gaussian / random_chi_squared returning fixed numbers in turn
    (the RandomMock of test/test_thermostat.py:16-24)."""

    def __init__(self, xs):
        self.xs, self.i = list(xs), 0

    def __call__(self, *args):
        self.i += 1
        return self.xs[self.i - 1]


def _kinetic(v, m=72.0):
    return 0.5 * m * float(np.sum(np.asarray(v, dtype=np.float64) ** 2))


@pytest.mark.parametrize("case", range(len(GOLD["cases"])))
def test_csvr_reference_regression(hip, case):
    """The reference's two asserted cases, through the drop-in function."""
    from hymd_amd.thermostat import csvr_thermostat
    c = GOLD["cases"][case]
    names = np.array(GOLD["names"])
    types = torch.as_tensor(np.searchsorted(NAMES, names).astype(np.int32)).cuda()
    v = torch.as_tensor(np.array(GOLD["velocities"])).cuda()
    cfg = _config(c["groups"])
    print(f"csvr golden case {case}", flush=True)
    csvr_thermostat(v, types, cfg, None, random_gaussian=_Mock(c["gaussian"]),
                    random_chi_squared=_Mock(c["chi_squared"]), remove_center_of_mass_momentum=c["remove_com"])
    assert cfg.thermostat_work == pytest.approx(c["work"], abs=1e-12)
    if "kinetic_energy_new" in c:
        assert _kinetic(v.cpu().numpy()) == pytest.approx(c["kinetic_energy_new"], abs=1e-12)


def _system(n, n_types, seed, drift=(0.7, -0.3, 0.2)):
    rng = np.random.default_rng(seed)
    v = rng.normal(0.0, 1.6, (n, 3)) + np.array(drift)
    types = rng.integers(0, n_types, n).astype(np.int32)
    return v, types


CASES = [
    # (n, n_types, groups (type ids), remove_com)
    (50_000, 4, None, True),
    (50_000, 4, None, False),
    (50_000, 5, [[0, 1], [2], [3, 4]], True),
    (50_000, 5, [[4], [0, 2], [1, 3]], False),
    (30_000, 3, [[0], [1], [2]], True),
    (200_000, 16, [[t] for t in range(16)], True),          # 16 COM groups, no "none" class
    (200_000, 16, [[0, 5, 9], [3], [12, 13, 14, 15]], True),  # some types in no group
    (1_001, 3, [[0, 1, 2]], True),
]


@pytest.mark.parametrize("n,n_types,groups,remove_com", CASES)
def test_csvr_matches_oracle(hip, n, n_types, groups, remove_com):
    from hymd_amd.thermostat import CsvrThermostat
    print(f"csvr n={n} T={n_types} groups={groups} com={remove_com}", flush=True)
    v0, types = _system(n, n_types, seed=n + n_types)
    names = np.array([f"T{t:02d}" for t in range(n_types)])
    cfg = _config(None if groups is None else [[names[t] for t in g] for g in groups],
                  unique_names=list(names), n_types=n_types)
    vd = torch.as_tensor(v0.copy()).cuda()
    td = torch.as_tensor(types).cuda()
    th = CsvrThermostat(td, cfg, remove_center_of_mass_momentum=remove_com)
    prng = np.random.default_rng(7)
    R, S = th.draw(prng)
    th(vd, None, draws=(R, S))
    th(vd, None, draws=([r * 0.5 for r in R], [s * 1.01 for s in S]))  # a second call on the result
    v_ref = v0.copy()
    grp = None if groups is None else [[names[t] for t in g] for g in groups]
    args = dict(time_step=cfg.time_step, respa_inner=cfg.respa_inner, tau=cfg.tau,
                target_temperature=cfg.target_temperature, mass=cfg.mass, remove_com=remove_com)
    w = OT.csvr_thermostat(v_ref, names[types], grp, gaussians=R, chi_squares=S, **args)
    w += OT.csvr_thermostat(v_ref, names[types], grp, gaussians=[r * 0.5 for r in R],
                            chi_squares=[s * 1.01 for s in S], **args)
    got = vd.cpu().numpy()
    err = np.max(np.abs(got - v_ref)) / np.max(np.abs(v_ref))
    assert err < 1e-13, err
    # the work K (alpha^2 - 1) loses digits when alpha^2 ~ 1: bound it by K
    assert float(th.work.item()) == pytest.approx(w, rel=1e-13, abs=1e-13 * _kinetic(v0))


@pytest.mark.parametrize("remove_com", [True, False])
def test_csvr_one_particle_group_between_com_groups(hip, remove_com):
    """A one-particle group takes the all-particle branch even with COM
    removal (thermostat.py:177, 182-183), between two COM groups."""
    from hymd_amd.thermostat import CsvrThermostat
    v0, types = _system(20_000, 3, seed=9)
    types[1234] = 3
    names = np.array(NAMES)
    groups = [["A"], ["D"], ["B", "C"]]
    cfg = _config(groups)
    vd = torch.as_tensor(v0.copy()).cuda()
    th = CsvrThermostat(torch.as_tensor(types).cuda(), cfg, remove_center_of_mass_momentum=remove_com)
    assert list(th.group_n)[1] == 1
    R, S = th.draw(np.random.default_rng(4))
    th(vd, None, draws=(R, S))
    v_ref = v0.copy()
    w = OT.csvr_thermostat(v_ref, names[types], groups, cfg.time_step, cfg.respa_inner, cfg.tau,
                           cfg.target_temperature, cfg.mass, R, S, remove_com=remove_com)
    err = np.max(np.abs(vd.cpu().numpy() - v_ref)) / np.max(np.abs(v_ref))
    assert err < 1e-13, err
    assert float(th.work.item()) == pytest.approx(w, rel=1e-13, abs=1e-13 * _kinetic(v0))


def test_csvr_draws_follow_the_reference_order(hip):
    from hymd_amd.thermostat import CsvrThermostat
    v0, types = _system(4000, 3, seed=5)
    cfg = _config([["A", "B"], ["C"]], unique_names=["A", "B", "C"], n_types=3)
    th = CsvrThermostat(torch.as_tensor(types).cuda(), cfg)
    R, S = th.draw(np.random.default_rng(3))
    counts = np.bincount(types, minlength=3)
    Ro, So = OT.draw_csvr_numbers(np.random.default_rng(3), [counts[0] + counts[1], counts[2]])
    assert R == Ro and S == So


def test_csvr_full_size_closed_form(hip):
    """At the lipid/water size (1,048,576 beads): with tau -> 0 (c = 0) each
    COM group's COM-free kinetic energy becomes K_t (S + R^2) / N_f exactly and
    its COM velocity is unchanged; the device work is the change of the total
    kinetic energy."""
    from hymd_amd.thermostat import CsvrThermostat
    n = 1_048_576
    v0, types = _system(n, 4, seed=11)
    cfg = _config([["A", "B"], ["C"], ["D"]], tau=1e-300)
    vd = torch.as_tensor(v0).cuda()
    th = CsvrThermostat(torch.as_tensor(types).cuda(), cfg)
    R, S = th.draw(np.random.default_rng(1))
    th(vd, None, draws=(R, S))
    v = vd.cpu().numpy()
    for g, ts in enumerate(th.group_types):
        m = np.isin(types, ts)
        n_g = int(m.sum())
        np.testing.assert_allclose(v[m].mean(axis=0), v0[m].mean(axis=0), rtol=0, atol=1e-12)
        K_t = 1.5 * cfg.gas_constant * n_g * cfg.target_temperature
        K = _kinetic(v[m] - v[m].mean(axis=0))
        assert K == pytest.approx(K_t * (S[g] + R[g] ** 2) / (3 * n_g), rel=1e-12)
    assert float(th.work.item()) == pytest.approx(_kinetic(v) - _kinetic(v0), rel=1e-10)


def test_csvr_empty_group_raises(hip):
    from hymd_amd.thermostat import CsvrThermostat
    v0, types = _system(100, 2, seed=2)     # only types 0, 1 present
    cfg = _config([["A"], ["C"]])
    th = CsvrThermostat(torch.as_tensor(types).cuda(), cfg)
    with pytest.raises(ValueError):
        th(torch.as_tensor(v0).cuda(), np.random.default_rng(0))


@pytest.mark.parametrize("n", [1, 777, 1_048_576])
def test_cancel_com_momentum(hip, n):
    from hymd_amd.thermostat import cancel_com_momentum
    print(f"cancel_com n={n}", flush=True)
    v0, _ = _system(n, 1, seed=n)
    vd = torch.as_tensor(v0.copy()).cuda()
    cfg = _config(n_particles=n)
    cancel_com_momentum(vd, cfg)
    ref = OT.cancel_com_momentum(v0.copy(), n)
    got = vd.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref)) + 1e-15
    assert np.max(np.abs(got.sum(axis=0))) < 1e-9 * np.sqrt(n)


@pytest.mark.parametrize("use_graph", [False, True])
def test_simulate_couples_like_the_main_loop(hip, use_graph):
    """RespaMD.simulate = outer steps with the thermostat after every n_b-th
    step and COM removal after every cancel_com_momentum-th step
    (main.py:1314-1321), here against the same calls made by hand."""
    from hymd_amd import synth
    from hymd_amd.config import Config
    from hymd_amd.md import RespaMD
    from hymd_amd.thermostat import CsvrThermostat, cancel_com_momentum
    s = synth.diblock_melt(n_chains=300, box=6.5, mesh=32, seed=1000)
    cfg = Config(n_steps=7, time_step=s.time_step, mesh_size=s.mesh, sigma=s.sigma, kappa=s.kappa,
                 mass=s.mass, respa_inner=s.respa_inner, target_temperature=310.0, tau=0.1,
                 unique_names=list(s.type_names), n_types=s.n_types, n_b=2, cancel_com_momentum=3,
                 n_particles=s.n)
    a = RespaMD(s)
    b = RespaMD(s)
    a.simulate(7, cfg, np.random.default_rng(21), use_graph=use_graph)
    th = CsvrThermostat(b.types, cfg, n_types=b.T)
    prng = np.random.default_rng(21)
    for step in range(1, 8):
        b.run(1, use_graph=use_graph)
        if step % 2 == 0:
            th(b.v, prng)
        if step % 3 == 0:
            cancel_com_momentum(b.v, cfg)
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x) and torch.equal(a.v, b.v)
    assert a.thermostat_work == float(th.work.item()) != 0.0
    a.close()
    b.close()
