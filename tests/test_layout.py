"""Spatial layout of the MD state (CPU): a permutation that keeps molecules
whole and contiguous, orders them by the mesh cell of their first particle,
and renumbers every bonded term onto the same particles."""
import numpy as np

from hymd_amd import layout, synth


def _check(s, mesh):
    order = layout.spatial_order(s.positions, s.box, mesh, s.bonds, s.angles, s.extra.get("dihedrals"))
    n = s.n
    assert sorted(order.tolist()) == list(range(n))          # a permutation
    p, inv = layout.permute_system(s, order)
    np.testing.assert_array_equal(inv[order], np.arange(n))
    np.testing.assert_array_equal(p.positions, s.positions[order])
    np.testing.assert_array_equal(p.types, s.types[order])
    # every term acts on the same particles (same positions) after renumbering
    for t_old, t_new, k in [(s.bonds, p.bonds, 2), (s.angles, p.angles, 3)] + (
            [(s.extra["dihedrals"], p.extra["dihedrals"], 4)] if "dihedrals" in s.extra else []):
        for j in range(k):
            np.testing.assert_array_equal(p.positions[t_new[j]], s.positions[t_old[j]])
        for j in range(k, len(t_old)):
            np.testing.assert_array_equal(np.asarray(t_new[j]), np.asarray(t_old[j]))
    # molecules contiguous, in order of their first particle's cell
    nmol, lab = layout.molecule_labels(n, p.bonds, p.angles, p.extra.get("dihedrals"))
    starts = np.flatnonzero(np.r_[True, lab[1:] != lab[:-1]])
    assert len(starts) == nmol                                   # each label one contiguous run
    m = np.asarray(mesh)
    cell = np.clip(np.floor(np.mod(p.positions[starts], s.box) / s.box * m).astype(int), 0, m - 1)
    key = (cell[:, 0] * m[1] + cell[:, 1]) * m[2] + cell[:, 2]
    single = np.diff(np.r_[starts, n]) == 1
    assert not np.any(single[:-1] & ~single[1:])   # single particles after every molecule
    assert np.all(np.diff(key[~single]) >= 0) and np.all(np.diff(key[single]) >= 0)


def test_layout_lipid_water():
    _check(synth.lipid_water(n_chains=300, n_water=2000, box=8.0, mesh=24, seed=5), (24, 24, 24))


def test_layout_peptide_with_dihedrals():
    _check(synth.peptide_solvent(n_chains=10, n_res=8, n_solvent=500, box=6.0, mesh=16, seed=9), (16, 16, 16))


def test_layout_joins_a_molecule_with_a_gap():
    s = synth.diblock_melt(n_chains=50, box=5.0, mesh=16, seed=3)
    a, b = s.bonds[0].copy(), s.bonds[1].copy()
    b[0] = 25
    s.bonds = (a, b, s.bonds[2], s.bonds[3])
    _check(s, (16, 16, 16))
