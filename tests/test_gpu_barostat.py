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


@pytest.mark.parametrize("btype", ["berendsen", "scr"])
@pytest.mark.parametrize("kind", ["isotropic", "semiisotropic"])
def test_npt_steps_match_the_oracle(hip, oracle_built, kind, btype):
    """RespaMD.simulate with a barostat every outer step, called between the
    inner loop and the field update as main.py:889-932, against the same
    steps composed from the FP64 oracle: kicks and inner loop, the pressure
    from the velocities after the inner loop, the last inner step's bonded
    virials and the previous field update's spectra, the barostat's box and
    position scaling, the field update in the new box, the second kick.  SCR
    takes its normal draws from the same generator; its target temperature
    switches the CSVR thermostat on, here with tau = 1e30 (c = 1: the
    identity, up to rounding), whose draws follow the barostat's in the
    reference's order.  Box to 1e-6 relative, positions to 1e-5 nm,
    velocities to 1e-4."""
    from hymd_amd import synth
    from hymd_amd.config import Config, Target_pressure
    from hymd_amd.md import RespaMD
    from oracle import bonded as OBd
    from oracle import field as F
    from oracle import md as OM
    from oracle import pressure as OP
    s = synth.diblock_melt(n_chains=300, box=6.5, mesh=24, seed=1000)
    s.respa_inner = 3
    A, b, c0 = synth.hamiltonian_form(s)
    ham = F.Quadratic(A, b, c0)
    T, m, n_in, dt = s.n_types, s.mass, s.respa_inner, s.time_step
    # target below the current pressure: the box grows by about 0.5 % over the three steps
    st = OM.State(s, ham)
    out = F.update_field(st.x, s.types, T, s.mesh, s.box, s.sigma, ham)
    _, _, bp = OBd.cbf(st.x, s.box, *s.bonds)
    _, _, ap = OBd.caf(st.x, s.box, *s.angles)
    p0 = OP.comp_pressure(out["phi"], out["phi_fourier"], ham, st.v, m, s.mesh, s.box, s.sigma, bp, ap)
    P_t = float(np.mean(p0[-3:])) * OB.BAR * 0.7
    cfg = Config(n_steps=3, time_step=dt, mesh_size=list(s.mesh), sigma=s.sigma, kappa=s.kappa, mass=m,
                 respa_inner=n_in, n_types=T, n_particles=s.n, box_size=s.box.copy(), barostat=kind,
                 barostat_type=btype, tau_p=1.0, n_b=1, target_pressure=Target_pressure(P_L=P_t, P_N=P_t))
    if btype == "scr":
        cfg.target_temperature = 300.0
        cfg.tau = 1e30
    md = RespaMD(s)
    md.simulate(3, cfg, np.random.default_rng(0))
    torch.cuda.synchronize()
    from oracle import thermostat as OT
    rng = np.random.default_rng(0)
    # the oracle's steps
    box = s.box.copy()
    last = out
    for _ in range(3):
        for f in st.f_slow:
            st.v = OM.integrate_velocity(st.v, f / m, n_in * dt)
        for i in range(n_in):
            st.v = OM.integrate_velocity(st.v, st.f_fast / m, dt)
            st.x = np.mod(OM.integrate_position(st.x, st.v, dt), box[None, :])
            fb, _, bp = OBd.cbf(st.x, box, *s.bonds)
            fa, _, ap = OBd.caf(st.x, box, *s.angles)
            st.f_fast = fb + fa
            st.v = OM.integrate_velocity(st.v, st.f_fast / m, dt)
        p = OP.comp_pressure(last["phi"], last["phi_fourier"], ham, st.v, m, s.mesh, box, s.sigma, bp, ap)
        if btype == "berendsen" and kind == "isotropic":
            OB.berendsen_isotropic(p, st.x, box, n_in * dt, 1, 1.0, P_t)
        elif btype == "berendsen":
            OB.berendsen_semiisotropic(p, st.x, box, n_in * dt, 1, 1.0, P_t, P_t, rows_quirk=False)
        elif kind == "isotropic":
            OB.scr_isotropic(p, st.x, box, n_in * dt, 1, 1.0, P_t, cfg.gas_constant, 300.0, rng.normal())
        else:
            rxy, rz = rng.normal(), rng.normal()
            OB.scr_semiisotropic(p, st.x, box, n_in * dt, 1, 1.0, P_t, P_t, cfg.gas_constant, 300.0, rxy, rz,
                                 rows_quirk=False)
        last = F.update_field(st.x, s.types, T, s.mesh, box, s.sigma, ham)
        f_field = F.compute_field_force(st.x, s.types, last["force_mesh"], s.mesh, box)
        st.v = OM.integrate_velocity(st.v, f_field / m, n_in * dt)
        st.f_slow = [f_field]
        if btype == "scr":
            OT.draw_csvr_numbers(rng, [s.n])     # the thermostat's draws (tau = 1e30: no change)
    assert np.max(np.abs(box / s.box - 1.0)) > (2e-3 if btype == "berendsen" else 1e-5)   # the box moved
    np.testing.assert_allclose(md.box, box, rtol=1e-6)
    np.testing.assert_allclose(cfg.box_size, box, rtol=1e-6)
    np.testing.assert_allclose(md.pm.BoxSize, box, rtol=1e-6)
    io = md.in_input_order
    dx = io(md.x).cpu().numpy() - st.x
    dx -= box * np.round(dx / box)
    assert np.abs(dx).max() < 1e-5
    v = io(md.v).cpu().numpy()
    assert np.abs(v - st.v).max() < 1e-4 * np.abs(st.v).max()
