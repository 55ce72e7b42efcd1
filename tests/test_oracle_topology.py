"""The topology oracle (oracle/topology.py) against the reference's expected
term lists (test/test_force.py:18-45, 103-127, 189-212 on the dppc_single and
alanine_octapeptide fixtures) and against networkx itself, whose traversal
orders the reference's output inherits."""
import json
import os

import numpy as np
import pytest

from oracle import topology as OT

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "topology_goldens.json")))


def _fixture(name):
    g = GOLD[name]
    return (np.array(g["molecules"]), np.array(g["names"]), np.array(g["bonds"]), np.array(g["indices"]),
            g.get("bond_2", []), g.get("bond_3", []), g.get("bond_4", []))


def test_dppc_bonds_and_angles():
    mol, names, bonds, idx, bt, at, _ = _fixture("dppc_single")
    b2, b3, b4, bb = OT.prepare_bonds_old(mol, names, bonds, idx, bt, at, [])
    g = GOLD["dppc_single"]
    assert len(b2) == 11 and len(b3) == 8 and b4 == [] and bb == []
    got2 = {tuple(t[:2]): t[2:] for t in b2}
    for e in g["expected_bonds"]:
        assert got2[tuple(e[:2])] == pytest.approx(e[2:], abs=1e-13)
    got3 = {tuple(t[:3]): t[3:] for t in b3}
    for e in g["expected_angles"]:
        assert got3[tuple(e[:3])] == pytest.approx([np.radians(e[3]), e[4]], abs=1e-13)


def test_alanine_dihedrals():
    mol, names, bonds, idx, bt, at, dt = _fixture("alanine_octapeptide")
    _, _, b4, bb = OT.prepare_bonds_old(mol, names, bonds, idx, [], [], dt)
    g = GOLD["alanine_octapeptide"]
    assert len(b4) == 5 and bb == []
    assert {tuple(t[:4]) for t in b4} == {tuple(e[:4]) for e in g["expected_dihedrals"]}
    for t in b4:
        np.testing.assert_array_equal(dt[t[4]][4], g["expected_dihedral_coeffs"])
        assert t[5] == 0
    arrays = OT.prepare_bonds(mol, names, bonds, idx, [], [], dt)
    coeff = arrays[13]
    # np.resize repeats the two given rows over the six (force.py:711)
    np.testing.assert_array_equal(coeff[0], [[1] * 5, [0] * 5] * 3)
    assert list(arrays[15]) == [0] * 5


def test_type_1_marks_the_last_backbone_dihedral_per_molecule():
    mol, names, bonds, idx, _, _, dt = _fixture("alanine_octapeptide")
    mol = np.array([0] * 8 + [1] * 8)                   # two molecules (the bond between 6 and 8 crosses them:
    bonds = bonds.copy()                                # cut it)
    bonds[6] = [4, 7, -1]
    bonds[8] = [9, 10, -1]
    dt1 = [["BB", "BB", "BB", "BB", [[1, 1, 1, 1, 1], [0, 0, 0, 0, 0]], 1]]
    arrays = OT.prepare_bonds(mol, names, bonds, idx, [], [], dt1)
    a1, last = arrays[9], arrays[15]
    assert list(a1) == [0, 8]                            # (0,2,4,6) in molecule 0, (8,10,12,14) in 1
    assert list(last) == [1, 1]
    mol, names, bonds, idx, _, _, _ = _fixture("alanine_octapeptide")
    arrays = OT.prepare_bonds(mol, names, bonds, idx, [], [], dt1)
    assert list(arrays[15]) == [0, 0, 0, 0, 1]           # one molecule: its last type-1 term


def test_partner_outside_the_molecule_raises_keyerror():
    mol, names, bonds, idx, bt, at, _ = _fixture("dppc_single")
    mol = mol.copy()
    mol[6:] = 1                                          # bond 5-6 now crosses two molecules
    with pytest.raises(KeyError):
        OT.prepare_bonds_old(mol, names, bonds, idx, bt, at, [])


def test_reversed_duplicate_with_another_type_raises_valueerror():
    mol, names, bonds, idx, _, _, dt = _fixture("alanine_octapeptide")
    two = [dt[0], ["BB", "BB", "BB", "BB", [[2, 2, 2, 2, 2], [0, 0, 0, 0, 0]], 0]]
    with pytest.raises(ValueError):
        OT.prepare_bonds_old(mol, names, bonds, idx, [], [], two)


def _random_molecule(rng, n, extra_edges, max_bonds=4):
    """A random tree plus a few ring-closing edges, as a (n, max_bonds) table
    listing each bond from both ends, rows in a random order."""
    edges = [(int(rng.integers(0, i)), i) for i in range(1, n)]
    for _ in range(extra_edges):
        a, b = rng.choice(n, 2, replace=False)
        edges.append((int(a), int(b)))
    table = [[] for _ in range(n)]
    for a, b in edges:
        if b not in table[a] and len(table[a]) < max_bonds and len(table[b]) < max_bonds:
            table[a].append(b)
            table[b].append(a)
    for t in table:
        rng.shuffle(t)
    return np.array([t + [-1] * (max_bonds - len(t)) for t in table])


@pytest.mark.parametrize("seed", range(12))
def test_traversal_orders_match_networkx(seed):
    nx = pytest.importorskip("networkx")
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 25))
    bonds = _random_molecule(rng, n, int(rng.integers(0, 4)))
    gids = rng.permutation(1000)[:n]                     # global indices need not be rows
    bonds_g = np.where(bonds >= 0, gids[np.maximum(bonds, 0)], -1)
    rows = np.arange(n)
    G = OT.molecule_graph(rows, gids, ["X"] * n, bonds_g)
    H = nx.Graph()                                        # built as force.py:405-417 builds it
    for r in rows:
        H.add_node(int(gids[r]))
        for b in bonds_g[r]:
            if b != -1:
                H.add_edge(int(gids[r]), int(b))
    assert G.order == list(H.nodes)
    for g in G.order:
        assert G.adj[g] == list(H.adj[g])
    for s in G.order:
        sp = nx.single_source_shortest_path(H, s)
        want = [(j, p[-2]) for j, p in sp.items() if len(p) in (2, 3)]
        assert G.bfs2(s) == want
        simple = {tuple(p) for t in H.nodes for p in nx.all_simple_paths(H, s, t, cutoff=3) if len(p) == 4}
        assert set(G.paths4(s)) == simple and len(G.paths4(s)) == len(simple)


TOPOL = {
    "system": {"molecules": [["PEP", 2], ["WAT", 1]]},
    "PEP": {"bonds": [[1, 2, 1, 0.35, 4000.0], [2, 3, 1, 0.36, 4100.0], [3, 4, 1, 0.37, 4200.0]],
            "angles": [[1, 2, 3, 1, 127.0, 20.0], [2, 3, 4, 1, 100.0, 25.0]],
            "dihedrals": [[1, 2, 3, 4, 0, [-1.0]],
                          [4, 3, 2, 1, 1, [[0.5], [1, 2, 3, 4, 5], [6, 7, 8, 9, 10]]],
                          [1, 2, 3, 4, 2, [[1, 1, 1, 1, 1], [2, 2, 2, 2, 2]]],
                          [2, 3, 4, 1, 3, [[1, 0, 0, 0, 0], [0, 1, 0, 0, 0]]]]},
    "WAT": {},
}


def test_index_based_topology_by_hand():
    """Two 4-bead peptides (rows 0-3 and 5-8, molecule ids 0 and 1), a water
    (row 4, id 2) and a particle of molecule 3, which has no residue: the
    expansion stops there."""
    molecules = np.array([0, 0, 0, 0, 2, 1, 1, 1, 1, 3])
    b2, b3, b4 = OT.prepare_index_based_bonds(molecules, TOPOL)
    assert [t[:2] for t in b2] == [[0, 1], [1, 2], [2, 3], [5, 6], [6, 7], [7, 8]]
    assert [t[2:] for t in b2] == [[0.35, 4000.0], [0.36, 4100.0], [0.37, 4200.0]] * 2
    assert [t[:3] for t in b3] == [[0, 1, 2], [1, 2, 3], [5, 6, 7], [6, 7, 8]]
    assert b3[0][3] == np.radians(127.0) and b3[1][4] == 25.0
    assert [t[:4] for t in b4] == [[0, 1, 2, 3], [3, 2, 1, 0], [0, 1, 2, 3], [1, 2, 3, 0],
                                   [5, 6, 7, 8], [8, 7, 6, 5], [5, 6, 7, 8], [6, 7, 8, 5]]
    assert [t[5] for t in b4] == [0, 1, 2, 3] * 2 and all(t[6] == 0 for t in b4)
    np.testing.assert_array_equal(b4[0][4][:2], [[7.406, -5.298, -2.570, 1.336, 0.739],
                                                 [-0.28632126, 1.2099146, 1.18122138, 0.49075168, 0.98495911]])
    np.testing.assert_array_equal(b4[0][4][2:], 0.0)
    assert b4[1][4].shape == (6, 5)
    np.testing.assert_array_equal(b4[1][4][4:], [[1, 2, 3, 4, 5], [6, 7, 8, 9, 10]])
    np.testing.assert_array_equal(b4[2][4], [[1] * 5, [2] * 5])
    np.testing.assert_array_equal(b4[3][4], [[1, 0, 0, 0, 0], [0, 1, 0, 0, 0], [0] * 5, [0] * 5])
    arrays = OT.prepare_bonds_topol(molecules, TOPOL)
    np.testing.assert_array_equal(arrays[13][2], np.resize(np.array([[1] * 5, [2] * 5], float), (6, 5)))
    assert list(arrays[15]) == [0] * 8


def test_propensity_closed_forms():
    a, b, c = np.array(OT._ALPHA), np.array(OT._BETA), np.array(OT._COIL)
    np.testing.assert_array_equal(OT.propensity_potential_coeffs(-1.0)[:2], a)
    np.testing.assert_array_equal(OT.propensity_potential_coeffs(0.0)[:2], c)
    np.testing.assert_array_equal(OT.propensity_potential_coeffs(1.0)[:2], b)
    for x in (-0.3, 0.6):
        p = OT.propensity_potential_coeffs(x)
        main = a if x < 0 else b
        np.testing.assert_allclose(p[0], abs(x) * main[0], rtol=1e-15)
        np.testing.assert_array_equal(p[1], main[1])
        np.testing.assert_allclose(p[2], (1 - abs(x)) * c[0], rtol=1e-15)
        np.testing.assert_array_equal(p[3], c[1])
    assert OT.propensity_potential_coeffs(1.5) is None
