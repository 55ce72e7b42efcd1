"""Device prepare_bonds (hymd_amd.topology) against the oracle
(oracle/topology.py): the 16 arrays element for element, in order (index
work: bit-exact), or the same exception; the reference's fixtures; and at the
lipid/water size the term sets of the synthetic chains."""
import json
import os
import time

import numpy as np
import pytest
import torch

from oracle import topology as OT

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "topology_goldens.json")))


class _Cfg:
    def __init__(self, bonds=(), angles=(), dihedrals=()):
        self.bonds, self.angle_bonds, self.dihedrals = list(bonds), list(angles), list(dihedrals)


def _compare(mol, names, bonds, idx, bt, at, dt):
    from hymd_amd.topology import prepare_bonds
    try:
        want = OT.prepare_bonds(mol, names, bonds, idx, bt, at, dt)
    except (KeyError, ValueError) as exc:
        with pytest.raises(type(exc)):
            prepare_bonds(mol, names, bonds, idx, _Cfg(bt, at, dt))
        return None
    got = prepare_bonds(mol, names, bonds, idx, _Cfg(bt, at, dt))
    assert len(got) == 16
    for k, (g, w) in enumerate(zip(got, want)):
        g = g.cpu().numpy()
        assert g.shape == w.shape, (k, g.shape, w.shape)
        np.testing.assert_array_equal(g, w, err_msg=f"array {k}")
    return want


def test_reference_fixtures(hip):
    g = GOLD["dppc_single"]
    args = (np.array(g["molecules"]), np.array(g["names"]), np.array(g["bonds"]), np.array(g["indices"]))
    w = _compare(*args, g["bond_2"], g["bond_3"], [])
    assert len(w[0]) == 11 and len(w[4]) == 8
    a = GOLD["alanine_octapeptide"]
    args = (np.array(a["molecules"]), np.array(a["names"]), np.array(a["bonds"]), np.array(a["indices"]))
    w = _compare(*args, a["bond_2"], a["bond_3"], a["bond_4"])
    assert len(w[9]) == 5
    # type 1 entries mark the last backbone dihedral of each molecule
    d1 = [["BB", "BB", "BB", "BB", [[1, 1, 1, 1, 1], [0, 0, 0, 0, 0]], 1]]
    w = _compare(*args, [], [], d1)
    assert list(w[15]) == [0, 0, 0, 0, 1]


def _random_molecule(rng, n, extra_edges, max_bonds):
    edges = [(int(rng.integers(0, i)), i) for i in range(1, n)]
    for _ in range(extra_edges if n > 2 else 0):
        a, b = rng.choice(n, 2, replace=False)
        edges.append((int(a), int(b)))
    table = [[] for _ in range(n)]
    for a, b in edges:
        if b not in table[a] and len(table[a]) < max_bonds and len(table[b]) < max_bonds:
            table[a].append(b)
            table[b].append(a)
    for t in table:
        rng.shuffle(t)
    return table


def _random_system(seed, n_mol, alphabet="ABC", max_bonds=4, one_sided=False):
    """Molecules of 1-14 particles (trees with ring closures), rows of the
    molecules interleaved and molecule ids unsorted, global indices a random
    injection; with one_sided each bond is listed from one end only."""
    rng = np.random.default_rng(seed)
    rows, mols, tables = [], [], []
    ids = rng.permutation(10 * n_mol)[:n_mol]
    for m in range(n_mol):
        n = int(rng.integers(1, 15))
        t = _random_molecule(rng, n, int(rng.integers(0, 3)), max_bonds)
        if one_sided:
            t = [[b for b in tb if b > a or rng.random() < 0.3] for a, tb in enumerate(t)]
        tables.append(t)
        mols += [int(ids[m])] * n
    N = len(mols)
    perm = rng.permutation(N)                      # row of each particle
    gids = rng.permutation(20 * N)[:N]             # global index of each row
    names = rng.choice(np.array(list(alphabet)), N)
    bonds = -np.ones((N, max_bonds), dtype=np.int64)
    molecules = np.zeros(N, dtype=np.int64)
    k = 0
    base = []
    for m, t in enumerate(tables):
        base.append(k)
        k += len(t)
    for m, t in enumerate(tables):
        for a, tb in enumerate(t):
            r = perm[base[m] + a]
            molecules[r] = mols[base[m] + a]
            for c, b in enumerate(tb):
                bonds[r, c] = gids[perm[base[m] + b]]
    return molecules, names, bonds, gids, rng


def _random_types(rng, alphabet="ABC", n2=4, n3=5, n4=4, kinds=(0, 1)):
    L = list(alphabet)
    bt = [[rng.choice(L), rng.choice(L), float(rng.uniform(0.3, 0.5)), float(rng.uniform(1e3, 2e3))] for _ in range(n2)]
    at = [[rng.choice(L), rng.choice(L), rng.choice(L), float(rng.uniform(90, 180)), float(rng.uniform(10, 30))]
          for _ in range(n3)]
    dt = [[rng.choice(L), rng.choice(L), rng.choice(L), rng.choice(L), rng.normal(size=(int(rng.integers(1, 5)), 5)),
           int(rng.choice(kinds))] for _ in range(n4)]
    return bt, at, dt


@pytest.mark.parametrize("seed", range(10))
def test_random_topologies_match_oracle(hip, seed):
    print(f"topology seed {seed}", flush=True)
    mol, names, bonds, gids, rng = _random_system(seed, n_mol=60, one_sided=seed % 3 == 0)
    bt, at, dt = _random_types(rng)
    if seed % 2:                                   # palindromic and repeated entries
        dt.append(["A", "B", "B", "A", np.ones((2, 5)), 0])
        bt.append(list(bt[0]))
    _compare(mol, names, bonds, gids, bt, at, dt)


def test_dihedral_reverse_rule_with_two_alphabet(hip):
    """Two names make reversed paths that match the same or another type."""
    for seed in range(6):
        mol, names, bonds, gids, rng = _random_system(100 + seed, n_mol=40, alphabet="AB")
        _, _, dt = _random_types(rng, alphabet="AB", n4=3)
        _compare(mol, names, bonds, gids, [], [], dt)


def test_errors_follow_the_reference(hip):
    from hymd_amd.topology import prepare_bonds
    g = GOLD["dppc_single"]
    mol = np.array(g["molecules"])
    mol[6:] = 1                                    # bond 5-6 crosses two molecules
    with pytest.raises(KeyError):
        prepare_bonds(mol, np.array(g["names"]), np.array(g["bonds"]), np.array(g["indices"]),
                      _Cfg(g["bond_2"], g["bond_3"]))
    a = GOLD["alanine_octapeptide"]
    two = [a["bond_4"][0], ["BB", "BB", "BB", "BB", [[2, 2, 2, 2, 2], [0, 0, 0, 0, 0]], 0]]
    with pytest.raises(ValueError):
        prepare_bonds(np.array(a["molecules"]), np.array(a["names"]), np.array(a["bonds"]),
                      np.array(a["indices"]), _Cfg([], [], two))


def test_lipid_water_terms_at_full_size(hip):
    """The bond table of the 1,048,576-bead lipid/water system (52,428 chains
    of 12, 419,440 waters) gives back the chains' bonds and angles, with the
    chain parameters, and the water rows no terms."""
    from hymd_amd import synth
    from hymd_amd.topology import prepare_bonds
    s = synth.lipid_water(seed=1001)
    n = s.n
    a, b = s.bonds[0], s.bonds[1]
    table = -np.ones((n, 2), dtype=np.int64)
    table[a, 1] = b                                 # chain i -> i+1 in column 1, i-1 in column 0
    table[b, 0] = a
    names = np.array(s.type_names)[s.types]
    molecules = np.r_[np.repeat(np.arange(52428), 12), 52428 + np.arange(n - 52428 * 12)]
    cfg = _Cfg(bonds=[[p, q, 0.47, 1250.0] for p, q in (("H", "H"), ("H", "L"), ("L", "L"), ("L", "T"), ("T", "T"))],
               angles=[["H", "H", "L", 180.0, 25.0], ["H", "L", "L", 120.0, 25.0], ["L", "L", "T", 120.0, 25.0],
                       ["L", "T", "T", 180.0, 25.0], ["T", "T", "T", 180.0, 25.0]])
    prepare_bonds(molecules[:24], names[:24], np.where(table[:24] < 24, table[:24], -1), np.arange(24), cfg)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = prepare_bonds(molecules, names, table, np.arange(n), cfg)
    torch.cuda.synchronize()
    print(f"prepare_bonds at {n} beads: {1e3 * (time.perf_counter() - t0):.1f} ms", flush=True)
    a2, b2, r0, k2 = (t.cpu().numpy() for t in out[:4])
    assert len(a2) == len(a)
    np.testing.assert_array_equal(a2, a)            # chain order = molecule order, bonds along the chain
    np.testing.assert_array_equal(b2, b)
    assert np.all(r0 == 0.47) and np.all(k2 == 1250.0)
    a3, b3, c3, t0_, k3 = (t.cpu().numpy() for t in out[4:9])
    np.testing.assert_array_equal(np.stack([a3, b3, c3]), np.stack(s.angles[:3]))
    np.testing.assert_array_equal(t0_, s.angles[3])
    assert len(out[9]) == 0


def _compare_topol(molecules, topol):
    from hymd_amd.topology import prepare_bonds
    want = OT.prepare_bonds_topol(molecules, topol)
    got = prepare_bonds(molecules, None, None, None, None, topol=topol)
    for k, (g, w) in enumerate(zip(got, want)):
        g = g.cpu().numpy()
        assert g.shape == w.shape, (k, g.shape, w.shape)
        np.testing.assert_array_equal(g, w, err_msg=f"array {k}")
    return want


def test_index_based_topology_matches_oracle(hip):
    from tests.test_oracle_topology import TOPOL
    w = _compare_topol(np.array([0, 0, 0, 0, 2, 1, 1, 1, 1, 3]), TOPOL)
    assert len(w[0]) == 6 and len(w[9]) == 8


@pytest.mark.parametrize("seed", range(4))
def test_random_index_based_topologies(hip, seed):
    """Residue templates of 1-12 beads with random terms, molecule ids
    shuffled over the rows, a tail of molecules beyond the topology's count
    (where the reference stops)."""
    rng = np.random.default_rng(seed)
    topol = {"system": {"molecules": []}}
    for r in range(4):
        nb = int(rng.integers(1, 12))
        res = {}
        if nb > 1:
            res["bonds"] = [[int(i), int(i) + 1, 1, float(rng.uniform(0.3, 0.5)), 1000.0] for i in range(1, nb)]
        if nb > 2:
            res["angles"] = [[int(i), int(i) + 1, int(i) + 2, 1, float(rng.uniform(90, 180)), 25.0]
                             for i in range(1, nb - 1)]
        if nb > 3:
            res["dihedrals"] = [[1, 2, 3, 4, 0, [float(rng.choice([-1.0, -0.4, 0.0, 0.7, 1.0]))]],
                                [2, 3, 4, 1, 2, rng.normal(size=(2, 5)).tolist()],
                                [4, 3, 2, 1, 1, [[0.3], [1, 2, 3, 4, 5], [5, 4, 3, 2, 1]]]]
        topol[f"R{r}"] = res
        topol["system"]["molecules"].append([f"R{r}", int(rng.integers(1, 40))])
    sizes = {f"R{r}": 1 + max([t[1] for t in topol[f"R{r}"].get("bonds", [[0, 0]])]) for r in range(4)}
    mols, rows = [], []
    mid = 0
    for name, count in topol["system"]["molecules"]:
        for _ in range(count):
            mols += [mid] * sizes[name]
            mid += 1
    mols += [mid] * 3 + [mid + 1] * 2                     # beyond the topology
    molecules = np.array(mols)[rng.permutation(len(mols))] if seed % 2 else np.array(mols)
    _compare_topol(molecules, topol)
