"""Device barostats (hymd_amd.barostat) against the oracle's formulas
(oracle/barostat.py) fed the same pressure 18-vector: box and positions
bit for bit (FP64 products), the float32 mirror, the mesh's box, the
n_b gate and the SCR draws."""
import numpy as np
import pytest
import torch

from oracle import barostat as OB

pytestmark = pytest.mark.gpu


def _setup(seed=3, n=6000, T=2, mesh=(24, 24, 32), box=(4.0, 4.0, 5.0)):
    from hymd_amd.config import Config, Target_pressure
    from hymd_amd.field import initialize_pm, update_field
    from oracle import field as OF
    rng = np.random.default_rng(seed)
    box = np.array(box)
    x = rng.uniform(0.0, 1.0, (n, 3)) * box
    types = rng.integers(0, T, n)
    rho0 = n / np.prod(box)
    ham = OF.default_with_chi(T, 0.05, rho0, rho0, np.array([[0, 10.0], [10.0, 0]]))
    cfg = Config(n_steps=0, time_step=0.03, mesh_size=list(mesh), sigma=0.5, kappa=0.05, box_size=box.copy(),
                 n_types=T, respa_inner=5, n_b=2, tau_p=1.0, target_temperature=300.0,
                 target_pressure=Target_pressure(P_L=1.0, P_N=1.0))

    class Ham:
        def quadratic(self):
            return ham.A, ham.b, ham.c0

    pm_stuff = initialize_pm(None, cfg)
    pm, fl = pm_stuff[0], pm_stuff[1]
    xd = torch.as_tensor(x).cuda()
    update_field(fl[0], fl[6], fl[5], None, fl[2], Ham(), pm, xd, torch.as_tensor(types, dtype=torch.int32).cuda(),
                 cfg, fl[4], fl[1], fl[3], None)
    v = torch.as_tensor(rng.normal(size=(n, 3))).cuda()
    bp, ap = rng.normal(size=3) * 50, rng.normal(size=3) * 50
    return cfg, Ham(), pm_stuff, fl, xd, v, bp, ap


# the isotropic form scales every coordinate alike: it has no row indexing to vary
CASES = [(g, k, q) for g in ("isotropic", "semiisotropic") for k in ("berendsen", "scr")
         for q in ((True,) if g == "isotropic" else (True, False))]


@pytest.mark.parametrize("geom,kind,rows_quirk", CASES)
def test_barostat_matches_oracle(hip, geom, kind, rows_quirk):
    from hymd_amd import barostat
    from hymd_amd.pressure import comp_pressure
    print(f"barostat {geom} {kind} rows_quirk={rows_quirk}", flush=True)
    cfg, H, pm_stuff, fl, xd, v, bp, ap = _setup()
    cfg.barostat_type = kind
    p = comp_pressure(fl[0], None, None, H, v, cfg, fl[1], fl[6], fl[5], xd, bp, ap)
    x_ref = xd.cpu().numpy().copy()
    box_ref = cfg.box_size.copy()
    x32 = xd.to(torch.float32)
    dt = cfg.time_step * cfg.respa_inner
    # the gate: nothing happens off the n_b grid
    out, change = getattr(barostat, geom)(None, pm_stuff, fl[0], None, None, H, xd, v, cfg, fl[1], fl[6], fl[5],
                                          bp, ap, 1, np.random.default_rng(9), x32=x32)
    assert change is False and np.array_equal(cfg.box_size, box_ref)
    kw = {} if geom == "isotropic" else {"rows_quirk": rows_quirk}
    out, change = getattr(barostat, geom)(None, pm_stuff, fl[0], None, None, H, xd, v, cfg, fl[1], fl[6], fl[5],
                                          bp, ap, 4, np.random.default_rng(9), x32=x32, **kw)
    rng = np.random.default_rng(9)
    tp = cfg.target_pressure
    if geom == "isotropic" and kind == "berendsen":
        OB.berendsen_isotropic(p, x_ref, box_ref, dt, cfg.n_b, cfg.tau_p, tp.P_L)
    elif geom == "isotropic":
        OB.scr_isotropic(p, x_ref, box_ref, dt, cfg.n_b, cfg.tau_p, tp.P_L, cfg.gas_constant, cfg.target_temperature,
                         rng.normal())
    elif kind == "berendsen":
        OB.berendsen_semiisotropic(p, x_ref, box_ref, dt, cfg.n_b, cfg.tau_p, tp.P_L, tp.P_N, rows_quirk=rows_quirk)
    else:
        rxy, rz = rng.normal(), rng.normal()
        OB.scr_semiisotropic(p, x_ref, box_ref, dt, cfg.n_b, cfg.tau_p, tp.P_L, tp.P_N, cfg.gas_constant,
                             cfg.target_temperature, rxy, rz, rows_quirk=rows_quirk)
    np.testing.assert_array_equal(cfg.box_size, box_ref)
    got = xd.cpu().numpy()
    np.testing.assert_array_equal(got, x_ref)
    np.testing.assert_array_equal(x32.cpu().numpy(), x_ref.astype(np.float32))
    assert np.array_equal(out[0].BoxSize, cfg.box_size)
    assert not np.array_equal(box_ref, np.array([4.0, 4.0, 5.0]))     # the box did move
    assert change is (not (geom == "isotropic" and kind == "scr"))  # barostat_scr.py:132 returns False


@pytest.mark.parametrize("geom", ["isotropic", "semiisotropic"])
def test_field_after_box_move_matches_oracle(hip, geom):
    """After a barostat step the moved mesh (same grid, new box) gives the
    oracle's filtered densities and force grids of the rescaled positions in
    the new box: cell size, wave numbers, filter and cell volume all follow
    the box."""
    from hymd_amd import barostat
    from hymd_amd.field import update_field
    from tests.conftest import rms_rel
    from oracle import field as OF
    print(f"field after box move {geom}", flush=True)
    cfg, H, pm_stuff, fl, xd, v, bp, ap = _setup(seed=11)
    cfg.barostat_type = "berendsen"
    box0 = cfg.box_size.copy()
    kw = {} if geom == "isotropic" else {"rows_quirk": False}
    getattr(barostat, geom)(None, pm_stuff, fl[0], None, None, H, xd, v, cfg, fl[1], fl[6], fl[5], bp, ap, 2,
                            np.random.default_rng(1), **kw)
    assert not np.array_equal(cfg.box_size, box0)
    T = cfg.n_types
    rng = np.random.default_rng(11)      # the types of _setup(seed=11): drawn after the positions
    rng.uniform(0.0, 1.0, (xd.shape[0], 3))
    types = rng.integers(0, T, xd.shape[0])
    ty = torch.as_tensor(types, dtype=torch.int32).cuda()
    pm = pm_stuff[0]
    update_field(fl[0], fl[6], fl[5], None, fl[2], H, pm, xd, ty, cfg, fl[4], fl[1], fl[3], None)
    x_new = xd.to(torch.float32).cpu().numpy().astype(np.float64)
    A, b, c0 = H.quadratic()
    ref = OF.update_field(x_new, types, T, cfg.mesh, cfg.box_size, cfg.sigma, OF.Quadratic(A, b, c0))
    for t in range(T):
        assert rms_rel(fl[0][t].cpu().numpy(), ref["phi"][t]) < 1e-6
        for d in range(3):
            assert rms_rel(fl[2][t][d].cpu().numpy(), ref["force_mesh"][t][d]) < 1e-5
