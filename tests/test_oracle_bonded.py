"""Pin the bonded-kernel oracle (oracle/bonded.c) before trusting it.

* reference goldens: test/test_force.py:47-273 (energies and forces);
* central finite differences of the energies (forces = -grad E);
* the reference's own Fortran compiled into oracle/_ref (when built);
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


def _ref_or_skip(B):
    if not B.ref_available():
        pytest.skip("reference Fortran not compiled (oracle/_ref)")


@pytest.mark.parametrize("single", [False, True])
def test_oracle_matches_compiled_reference(oracle_built, single):
    """The C restatement against the reference's own Fortran on random terms,
    including CBT dihedrals with dipole reconstruction and last-angle terms."""
    B = oracle_built
    _ref_or_skip(B)
    rng = np.random.default_rng(1)
    n, m = 200, 150
    box = np.array([4.0, 5.0, 6.0])
    r = rng.uniform(0, 1, (n, 3)) * box
    if single:
        r = r.astype(np.float32).astype(np.float64)
    a = rng.integers(0, n, m)
    b, c, d = (a + 1) % n, (a + 2) % n, (a + 3) % n
    r0, k, t0 = rng.uniform(0.3, 0.6, m), rng.uniform(100, 1500, m), rng.uniform(1.5, 3.1, m)
    atol = 0.0 if single else 1e-10
    f1, e1, p1 = B.cbf(r, box, a, b, r0, k, single)
    f2, e2, p2 = B.ref_cbf(r, box, a, b, r0, k, single)
    np.testing.assert_allclose(f1, f2, rtol=0, atol=atol * 10)
    assert e1 == pytest.approx(e2, rel=1e-14)
    f1, e1, p1 = B.caf(r, box, a, b, c, t0, k, single)
    f2, e2, p2 = B.ref_caf(r, box, a, b, c, t0, k, single)
    np.testing.assert_allclose(f1, f2, rtol=0, atol=atol * 100)
    assert e1 == pytest.approx(e2, rel=1e-14)
    coeff = rng.uniform(-5, 5, (m, 6, 5)).astype(np.float32)
    dt = rng.integers(0, 3, m)
    coeff[dt == 2, 0, 1] = np.abs(coeff[dt == 2, 0, 1])
    last = rng.integers(0, 2, m)
    for flag in (0, 1):
        f1, d1, t1, e1 = B.cdf(r, box, a, b, c, d, coeff, dt, last, flag, single)
        f2, d2, t2, e2 = B.ref_cdf(r, box, a, b, c, d, coeff, dt, last, flag, single)
        np.testing.assert_allclose(f1, f2, rtol=0, atol=atol * 100)
        np.testing.assert_allclose(d1, d2, rtol=0, atol=atol)
        np.testing.assert_allclose(t1, t2, rtol=0, atol=atol)
        assert e1 == pytest.approx(e2, rel=1e-13)


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
