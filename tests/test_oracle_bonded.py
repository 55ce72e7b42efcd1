"""Pin the bonded-kernel oracle (oracle/bonded.c) before trusting it.

* reference goldens: test/test_force.py:47-273 (energies and forces);
* central finite differences of the energies (forces = -grad E);
* the cosine series and the CBT energy 1/2 K(phi) (gamma - gamma0(phi))^2 on
  hand-built geometries of known gamma and phi, with phases that fix the sign
  convention of phi (the goldens use zero phases);
* the dipole pair of reconstruct (dipole_reconstruction.f90:148-166): length
  delta, angle psi to the bond, the B-C midpoint, and theta(gamma) as an angle
  measured in the (n, m) frame against known values of the fit;
* the transfer matrices against central finite differences of the dipole
  displacement, and the redistribution (force.py:855-877) by virtual work;
* translation invariance of the dipole transfer matrices.
"""
import numpy as np
import pytest


def _match(goldens_E, goldens_F, terms, evaluate):
    """Each golden row must be produced by exactly one term (the reference
    tests index terms in prepare_bonds order, which is not stored)."""
    used = set()
    for E, F in zip(goldens_E, goldens_F):
        hit = None
        for j, t in enumerate(terms):
            if j in used:
                continue
            e, forces = evaluate(t)
            if abs(e - E) <= 1e-12 and any(np.allclose(f, F, atol=1e-11, rtol=0) for f in forces):
                hit = j
                break
        assert hit is not None, f"no term reproduces golden energy {E}"
        used.add(hit)
    assert len(used) == len(goldens_E)


def test_bond_goldens(oracle_built, goldens):
    B = oracle_built
    g, sys = goldens["bonds"], goldens["dppc_single"]
    r = np.array(sys["positions"])
    box = np.array(sys["box"])

    def ev(t):
        i, j, r0, k = t
        f, e, _ = B.cbf(r, box, [i], [j], [r0], [k])
        return e, [f[i], f[j]]

    _match(g["energies"], g["forces_i"], g["terms"], ev)


def test_angle_goldens(oracle_built, goldens):
    B = oracle_built
    g, sys = goldens["angles"], goldens["dppc_single"]
    r = np.array(sys["positions"])
    box = np.array(sys["box"])
    used = []
    for E, Fi, Fj, Fk in zip(g["energies"], g["forces_i"], g["forces_j"], g["forces_k"]):
        ok = False
        for t in g["terms"]:
            a, b, c, deg, k = t
            for (p, q) in ((a, c), (c, a)):
                f, e, _ = B.caf(r, box, [p], [b], [q], [np.radians(deg)], [k])
                if (abs(e - E) < 1e-12 and np.allclose(f[p], Fi, atol=1e-11) and np.allclose(f[b], Fj, atol=1e-11)
                        and np.allclose(f[q], Fk, atol=1e-11)):
                    ok = True
                    used.append(tuple(t))
        assert ok, f"no angle reproduces golden energy {E}"
    assert len(set(used)) == len(g["terms"])


@pytest.mark.parametrize("single", [False, True])
def test_dihedral_goldens(oracle_built, goldens, single):
    """Energies equal the goldens; forces are the goldens negated: the reference's
    test kernel (force.py:805-852) returns +grad E, the Fortran -grad E."""
    B = oracle_built
    g, sys = goldens["dihedrals"], goldens["alanine_octapeptide"]
    r = np.array(sys["positions"])
    box = np.array(sys["box"])
    coeff = np.zeros((1, 6, 5), dtype=np.float32)
    coeff[0, 0, :] = 1.0
    tol = 2e-5 if single else 1e-12
    for term, E, Fi, Fj, Fk, Fl in zip(g["terms"], g["energies"], g["forces_i"], g["forces_j"], g["forces_k"],
                                       g["forces_l"]):
        a, b, c, d = term
        f, _, _, e = B.cdf(r, box, [a], [b], [c], [d], coeff, [0], [0], 0, single=single)
        assert e == pytest.approx(E, abs=tol * max(1, abs(E)))
        for idx, Fg in zip(term, (Fi, Fj, Fk, Fl)):
            np.testing.assert_allclose(f[idx], -np.array(Fg), atol=tol * 100)


def _random_chain_system(seed, n=60, m=40, box=(4.0, 5.0, 6.0)):
    rng = np.random.default_rng(seed)
    box = np.array(box)
    r = rng.uniform(0, 1, (n, 3)) * box
    a = rng.choice(n - 3, m, replace=False)
    return rng, box, r, a


def _fd_forces(energy, r, atoms, h=1e-6):
    out = {}
    for p in atoms:
        g = np.zeros(3)
        for d in range(3):
            rp = r.copy()
            rm = r.copy()
            rp[p, d] += h
            rm[p, d] -= h
            g[d] = (energy(rp) - energy(rm)) / (2 * h)
        out[p] = -g
    return out


def test_bond_angle_fd(oracle_built):
    B = oracle_built
    rng, box, r, a = _random_chain_system(3)
    b, c = a + 1, a + 2
    r0 = rng.uniform(0.3, 0.6, len(a))
    k = rng.uniform(100, 1500, len(a))
    t0 = rng.uniform(1.6, 3.1, len(a))
    f, _, _ = B.cbf(r, box, a, b, r0, k)
    fd = _fd_forces(lambda x: B.cbf(x, box, a, b, r0, k)[1], r, range(len(r)))
    for p, v in fd.items():
        np.testing.assert_allclose(f[p], v, atol=1e-5 * max(1.0, np.abs(f).max()))
    f, _, _ = B.caf(r, box, a, b, c, t0, k)
    fd = _fd_forces(lambda x: B.caf(x, box, a, b, c, t0, k)[1], r, range(len(r)))
    for p, v in fd.items():
        np.testing.assert_allclose(f[p], v, atol=1e-5 * max(1.0, np.abs(f).max()))


@pytest.mark.parametrize("dtype", [0, 1, 2])
def test_dihedral_fd(oracle_built, dtype):
    """cosine-series, CBT (combined bending-torsion, dipole_reconstruction.f90)
    and improper dihedrals: forces = -grad E by central differences."""
    B = oracle_built
    rng = np.random.default_rng(10 + dtype)
    box = np.array([6.0, 6.0, 6.0])
    # a zig-zag backbone with angles well away from 0 and pi (no sin(gamma) clamp)
    n = 12
    r = np.zeros((n, 3))
    for i in range(1, n):
        r[i] = r[i - 1] + 0.38 * np.array([1.0, 0.6 * (-1) ** i, 0.0]) + rng.normal(scale=0.06, size=3)
    r = np.mod(r + 1.0, box)
    m = n - 3
    a = np.arange(m)
    coeff = np.zeros((m, 6, 5), dtype=np.float32)
    coeff[:, 0, 1:] = rng.uniform(-5, 5, (m, 4))
    coeff[:, 1, :] = rng.uniform(-1, 1, (m, 5))
    if dtype == 1:
        coeff[:, 4, 0] = 40.0
        coeff[:, 4, 1] = 5.0
        coeff[:, 5, 1] = 0.3
    if dtype == 2:
        coeff[:, 0, 0] = 0.4
        coeff[:, 0, 1] = 30.0
    dt = np.full(m, dtype)
    last = np.zeros(m, dtype=int)
    last[-1] = 1
    f, _, _, _ = B.cdf(r, box, a, a + 1, a + 2, a + 3, coeff, dt, last, 0)
    fd = _fd_forces(lambda x: B.cdf(x, box, a, a + 1, a + 2, a + 3, coeff, dt, last, 0)[3], r, range(n))
    scale = max(1.0, np.abs(f).max())
    for p, v in fd.items():
        np.testing.assert_allclose(f[p], v, atol=2e-5 * scale)


def test_transfer_matrices_translation_invariant(oracle_built):
    """The three transfer matrices of one reconstruction sum to zero: moving
    all beads rigidly does not move the dipole pair relative to them
    (dipole_reconstruction.f90:121-127,181-213)."""
    B = oracle_built
    rng = np.random.default_rng(5)
    box = np.array([8.0, 8.0, 8.0])
    n = 10
    r = np.zeros((n, 3))
    for i in range(1, n):
        r[i] = r[i - 1] + 0.38 * np.array([1.0, 0.6 * (-1) ** i, 0.1]) + rng.normal(scale=0.05, size=3)
    r += 2.0
    m = n - 3
    a = np.arange(m)
    coeff = np.zeros((m, 6, 5), dtype=np.float32)
    coeff[:, 4, 0] = 40.0
    last = np.zeros(m, dtype=int)
    last[-1] = 1
    _, dip, tm, _ = B.cdf(r, box, a, a + 1, a + 2, a + 3, coeff, np.ones(m), last, 1)
    s = tm[:, 0:3].sum(axis=1)
    np.testing.assert_allclose(s, 0.0, atol=1e-12)
    s2 = tm[-1, 3:6].sum(axis=0)
    np.testing.assert_allclose(s2, 0.0, atol=1e-12)
    # the dipole pair straddles the B-C bond midpoint (both wrapped with NINT)
    d0 = dip[:, 1] + box * np.round((dip[:, 0] - dip[:, 1]) / box)
    mid = 0.5 * (dip[:, 0] + d0) - 0.5 * (r[a + 1] + r[a + 2])
    np.testing.assert_allclose(mid - box * np.round(mid / box), 0.0, atol=1e-12)


def test_redistribution_conserves_total_force(oracle_built):
    """force.py:855-877: sum over beads equals sum over dipole points."""
    B = oracle_built
    rng = np.random.default_rng(8)
    box = np.array([8.0, 8.0, 8.0])
    n = 12
    r = np.zeros((n, 3))
    for i in range(1, n):
        r[i] = r[i - 1] + 0.38 * np.array([1.0, 0.6 * (-1) ** i, 0.1]) + rng.normal(scale=0.05, size=3)
    r += 2.0
    m = n - 3
    a = np.arange(m)
    coeff = np.zeros((m, 6, 5), dtype=np.float32)
    coeff[:, 4, 0] = 40.0
    last = np.zeros(m, dtype=int)
    last[-1] = 1
    dt = np.ones(m, dtype=int)
    _, dip, tm, _ = B.cdf(r, box, a, a + 1, a + 2, a + 3, coeff, dt, last, 1)
    fd = rng.normal(size=(m, 4, 3))
    fd[last == 0, 2:] = 0.0
    out = B.dipole_forces_redistribution(n, fd, tm, a, a + 1, a + 2, a + 3, dt, last)
    np.testing.assert_allclose(out.sum(axis=0), fd.reshape(-1, 3).sum(axis=0), atol=1e-12)


# ---------------------------------------------------------------------------
# Hand-built geometries: gamma and phi known by construction
# ---------------------------------------------------------------------------
DELTA = 0.3        # dipole_reconstruction.f90:86
PSI = 1.392947     # dipole_reconstruction.f90:86 (cos, sin taken in real(4))
# theta(gamma) of the fit at dipole_reconstruction.f90:148-150, evaluated once
# at these angles (radians)
THETA_KNOWN = {1.45: -0.35317574815910624, 1.70: -1.1907678326713496, 1.73: -1.7446100000000002,
               1.76: -2.2984521673286507, 1.95: -3.0393662123957736, 2.30: -3.6020999997639866}


def _frame(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 2] *= -1
    return q[:, 0], q[:, 1], q[:, 2]


def _four_beads(g1, g2, p, seed=0, lengths=(0.37, 0.38, 0.36), centre=(4.0, 4.1, 3.9)):
    """A-B-C-D with the angle at B (between BA and BC) = g1, the angle at C
    (between CB and CD) = g2, and D = C + d(-cos g2 e1 + sin g2 (cos p e2 +
    sin p e3)).  For this construction the reference's dihedral
    (f = rA - rB, g = rB - rC, h = rD - rC, phi = atan2((h x g).f |g|,
    (f x g).(h x g)), compute_dihedral_forces.f90:35-50) equals -p."""
    e1, e2, e3 = _frame(seed)
    a, b, d = lengths
    B = np.array(centre)
    C = B + b * e1
    A = B + a * (np.cos(g1) * e1 + np.sin(g1) * e2)
    D = C + d * (-np.cos(g2) * e1 + np.sin(g2) * (np.cos(p) * e2 + np.sin(p) * e3))
    return np.stack([A, B, C, D])


def _cdf1(B, r, coeff, dtype, last, flag=0):
    box = np.array([9.0, 9.5, 10.0])
    return B.cdf(r, box, [0], [1], [2], [3], coeff[None], [dtype], [last], flag)


@pytest.mark.parametrize("p", [-2.5, -0.7, 0.4, 1.9, 3.0])
def test_cosine_series_energy_and_phi_sign(oracle_built, p):
    """V = sum_n c_n (1 + cos(n phi - d_n)) (dipole_reconstruction.f90:37-48)
    with nonzero phases, at phi = -p: the phases make V odd-sensitive, so a
    dihedral of the opposite sign fails."""
    B = oracle_built
    coeff = np.zeros((6, 5), dtype=np.float32)
    coeff[0] = [0.7, -2.0, 3.5, 1.25, -0.8]
    coeff[1] = [0.0, 0.6, -1.1, 2.0, 0.3]
    c, d = coeff[0].astype(np.float64), coeff[1].astype(np.float64)
    r = _four_beads(1.9, 2.0, p, seed=1)
    phi = -p
    expect = float(np.sum(c * (1 + np.cos(np.arange(5) * phi - d))))
    e = _cdf1(B, r, coeff, 0, 0)[3]
    assert e == pytest.approx(expect, rel=1e-12, abs=1e-12)
    wrong = float(np.sum(c * (1 + np.cos(np.arange(5) * p - d))))
    assert abs(wrong - expect) > 1e-3


@pytest.mark.parametrize("g1, g2, p, last", [(1.9, 2.1, 0.8, 0), (1.6, 2.4, -1.3, 1), (2.2, 1.75, 2.6, 1)])
def test_cbt_energy_closed_form(oracle_built, g1, g2, p, last):
    """CBT (dtype 1, dipole_reconstruction.f90:90-136): with the torsion series
    zero, E = 1/2 K(phi) (gamma - gamma0(phi))^2 per bending angle, K(phi) =
    sum_n c_n (1 + cos(n phi - d_n)) from coeff rows 5, 6 and
    gamma0(phi) = 1.85 - 0.227 cos(phi - 0.785) (Levitt-Warshel); the second
    angle (at C) only for the last backbone term."""
    B = oracle_built
    coeff = np.zeros((6, 5), dtype=np.float32)
    coeff[4] = [30.0, 4.0, -6.0, 2.5, 1.0]
    coeff[5] = [0.0, 0.35, -0.9, 1.4, 0.2]
    c, d = coeff[4].astype(np.float64), coeff[5].astype(np.float64)
    r = _four_beads(g1, g2, p, seed=2)
    phi = -p
    K = float(np.sum(c * (1 + np.cos(np.arange(5) * phi - d))))
    g0 = 1.85 - 0.227 * np.cos(phi - 0.785)
    expect = 0.5 * K * (g1 - g0) ** 2 + (0.5 * K * (g2 - g0) ** 2 if last else 0.0)
    for flag in (0, 1):
        e = _cdf1(B, r, coeff, 1, last, flag)[3]
        assert e == pytest.approx(expect, rel=1e-12)


def _dipole_frame(ra, rb, rc):
    """The frame of dipole_reconstruction.f90:101-160: w, v unit vectors from
    B to A and to C, n = w x v / sin(gamma), m = n x v."""
    w = (ra - rb) / np.linalg.norm(ra - rb)
    v = (rc - rb) / np.linalg.norm(rc - rb)
    n = np.cross(w, v)
    n /= np.linalg.norm(n)
    return v, n, np.cross(n, v), float(np.arccos(np.clip(w @ v, -1, 1)))


@pytest.mark.parametrize("g1, g2", [(1.45, 1.95), (1.70, 1.76), (1.73, 2.30)])
def test_dipole_pair_geometry(oracle_built, g1, g2):
    """Each dipole pair sits at the midpoint of its bond (B-C, and C-D for
    the last term), is delta = 0.3 nm long, makes the angle psi = 1.392947
    with the bond, and its azimuth in the (n, m) frame is theta(gamma) of the
    fit, compared with known values."""
    B = oracle_built
    coeff = np.zeros((6, 5), dtype=np.float32)
    coeff[4, 0] = 10.0
    r = _four_beads(g1, g2, 0.9, seed=3)
    _, dip, _, _ = _cdf1(B, r, coeff, 1, 1, 1)
    for pair, (ia, ib, ic), g in ((0, (0, 1, 2), g1), (1, (1, 2, 3), g2)):
        p1, p2 = dip[0, 2 * pair], dip[0, 2 * pair + 1]
        v, n, m, gam = _dipole_frame(r[ia], r[ib], r[ic])
        assert gam == pytest.approx(g, abs=1e-12)
        np.testing.assert_allclose(0.5 * (p1 + p2), 0.5 * (r[ib] + r[ic]), atol=1e-12)
        dd = p1 - p2
        # cos psi and sin psi are real(4) constants: cos^2 + sin^2 = 1 to 4e-8
        assert np.linalg.norm(dd) == pytest.approx(DELTA, rel=1e-7)
        assert dd @ v / DELTA == pytest.approx(np.cos(PSI), abs=1e-7)
        theta = np.arctan2(dd @ m, dd @ n)
        expect = THETA_KNOWN[g]
        assert np.angle(np.exp(1j * (theta - expect))) == pytest.approx(0.0, abs=1e-7)


def _dipole_linearisation(B, r, coeff, last, pair, bead, h=1e-6):
    """Derivatives of the dipole displacement d = (p1 - p2)/2 of one pair:
    J[alpha, beta] = dd_alpha / dr_{bead, beta} by central differences, the
    gradient of the bending angle gamma of that pair, and dd/dgamma along an
    in-plane rotation of the pair's first bead about the second (v, n, m fixed)."""
    ia, ib, ic = (0, 1, 2) if pair == 0 else (1, 2, 3)

    def disp(rr):
        dip = _cdf1(B, rr, coeff, 1, last, 1)[1]
        return 0.5 * (dip[0, 2 * pair] - dip[0, 2 * pair + 1])

    J = np.zeros((3, 3))
    gg = np.zeros(3)
    for beta in range(3):
        rp, rm = r.copy(), r.copy()
        rp[bead, beta] += h
        rm[bead, beta] -= h
        J[:, beta] = (disp(rp) - disp(rm)) / (2 * h)
        gg[beta] = (_dipole_frame(rp[ia], rp[ib], rp[ic])[3] - _dipole_frame(rm[ia], rm[ib], rm[ic])[3]) / (2 * h)
    v, n, m, gam = _dipole_frame(r[ia], r[ib], r[ic])
    # rotate bead A of the pair about bead B in the (w, v) plane by +-h: gamma -> gamma -+ h
    u = r[ia] - r[ib]
    ax = -n  # rotation about -n moves w towards... sign fixed by measuring gamma below

    def rot(t):
        rr = r.copy()
        c_, s_ = np.cos(t), np.sin(t)
        rr[ia] = r[ib] + u * c_ + np.cross(ax, u) * s_ + ax * (ax @ u) * (1 - c_)
        return rr
    rp, rm = rot(h), rot(-h)
    dg = _dipole_frame(rp[ia], rp[ib], rp[ic])[3] - _dipole_frame(rm[ia], rm[ib], rm[ic])[3]
    dd_dgamma = (disp(rp) - disp(rm)) / dg
    d0 = disp(r)
    d_perp = d0 - (d0 @ v) * v
    return J, gg, dd_dgamma, d_perp, gam


@pytest.mark.parametrize("seed, g1, g2, p", [(4, 1.9, 2.05, 0.8), (5, 1.62, 2.3, -1.7), (6, 2.15, 1.71, 2.9)])
def test_transfer_matrices_against_finite_differences(oracle_built, seed, g1, g2, p):
    """Transfer matrix D_k of bead k (dipole_reconstruction.f90:167-213) in the
    convention of force.py:866-877 (F_k += D_k @ (F_1 - F_2)): the transpose
    of the Jacobian of the dipole displacement d, except that the reference's
    terms proportional to dgamma/dr_k enter with the opposite sign to the
    exact derivative (the 1/sin(gamma) normalisation of n in N_k, and the
    theta(gamma) terms FN_k - FM_k).  So
        D_k = J_k^T + 2 grad_k(gamma) (x) (cot(gamma) d_perp - dd/dgamma),
    every term measured by central differences of the dipole positions and of
    the bending angle.  A transposed matrix, a dropped term or a wrong
    prefactor fails."""
    B = oracle_built
    coeff = np.zeros((6, 5), dtype=np.float32)
    coeff[4, 0] = 10.0
    r = _four_beads(g1, g2, p, seed=seed)
    _, _, tm, _ = _cdf1(B, r, coeff, 1, 1, 1)
    for pair, beads in ((0, (0, 1, 2)), (1, (1, 2, 3))):
        for k, bead in enumerate(beads):
            J, gg, dd_dg, d_perp, gam = _dipole_linearisation(B, r, coeff, 1, pair, bead)
            expect = J.T + 2.0 * np.outer(gg, d_perp / np.tan(gam) - dd_dg)
            got = tm[0, 3 * pair + k]
            np.testing.assert_allclose(got, expect, atol=2e-7 * max(1.0, np.abs(expect).max()))
            assert np.abs(got - J.T).max() > 1e-3  # the reference's sign is not the exact one
            assert np.abs(got.T - expect).max() > 1e-3  # the convention is not transposed


def test_redistribution_virtual_work(oracle_built):
    """force.py:855-877: for any virtual displacement dr of the beads,
    sum_k F_k . dr_k = sum_pairs [(F_1 + F_2) . dr_mid + (F_1 - F_2) . dd],
    with dr_mid the move of the pair's bond midpoint and dd the reference's
    linearised move of the displacement (previous test:
    dd = J dr + 2 (grad gamma . dr)(cot gamma d_perp - dd/dgamma)).  A
    transposed matrix or a missing 1/2 on the midpoint share fails."""
    B = oracle_built
    rng = np.random.default_rng(12)
    coeff = np.zeros((6, 5), dtype=np.float32)
    coeff[4, 0] = 10.0
    r = _four_beads(1.85, 2.1, 1.1, seed=7)
    _, _, tm, _ = _cdf1(B, r, coeff, 1, 1, 1)
    fdip = rng.normal(size=(1, 4, 3))
    a, b, c, d = [0], [1], [2], [3]
    out = B.dipole_forces_redistribution(4, fdip, tm, a, b, c, d, [1], [1])
    for _ in range(3):
        dr = rng.normal(size=(4, 3))
        lhs = float(np.sum(out * dr))
        rhs = 0.0
        for pair, beads in ((0, (0, 1, 2)), (1, (1, 2, 3))):
            f1, f2 = fdip[0, 2 * pair], fdip[0, 2 * pair + 1]
            rhs += (f1 + f2) @ (0.5 * (dr[beads[1]] + dr[beads[2]]))
            dd = np.zeros(3)
            for bead in beads:
                J, gg, dd_dg, d_perp, gam = _dipole_linearisation(B, r, coeff, 1, pair, bead)
                dd += J @ dr[bead] + 2.0 * (gg @ dr[bead]) * (d_perp / np.tan(gam) - dd_dg)
            rhs += (f1 - f2) @ dd
        assert lhs == pytest.approx(rhs, rel=1e-6, abs=1e-8)
    # the transposed convention (D_k^T @ diff) breaks the identity
    out_t = np.zeros_like(out)
    for pair, beads in ((0, (0, 1, 2)), (1, (1, 2, 3))):
        f1, f2 = fdip[0, 2 * pair], fdip[0, 2 * pair + 1]
        for k, bead in enumerate(beads):
            out_t[bead] += tm[0, 3 * pair + k].T @ (f1 - f2)
            if k > 0:
                out_t[bead] += 0.5 * (f1 + f2)
    assert np.abs(out_t - out).max() > 1e-3
