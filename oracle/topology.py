"""Topology -> bonded index arrays (oracle; test infrastructure only).

prepare_bonds: force.py:573-728 over prepare_bonds_old :325-505, restated
without networkx.  The order of the output terms follows networkx's
documented traversal orders, which the reference's results inherit:

* graph: nodes in order of first appearance (add_node of each member, in row
  order, and add_edge of its bond partners, in column order); each node's
  neighbours in order of first edge insertion (networkx.Graph adjacency
  dicts);
* bonds / angles: for every node s in node order, the breadth-first
  discovery order of single_source_shortest_path (level 1 = the neighbours
  of s, level 2 = for each level-1 node v in order, its neighbours not yet
  seen, reached through v); a bond for each level-1 node j, an angle
  (s, v, j) for each level-2 node j, both only when the global index of j
  exceeds that of s, once per matching type (forward or backward names);
* dihedrals: the simple 4-node paths s-a-b-c of find_all_paths (force.py:
  266-276), nested in neighbour order, once per type whose names match
  forward, unless the reversed term with the same type was already stored
  (it was iff c precedes s in node order and the reversed path matches that
  type); the reference compares stored terms as lists, and when such a
  reversed term was stored with a different type object first it raises
  ValueError (numpy array truth value), as this restatement does;
* per molecule (np.unique order) the last stored dihedral of type 1 gets
  last = 1; coefficients are np.resize(coeffs, (6, 5)) (force.py:711).

A bond partner outside the molecule has no node attributes in the
reference (KeyError); here too.

prepare_index_based_bonds (force.py:508-570) with propensity_potential_coeffs
(force.py:280-322): per molecule (np.unique order; resid = mol + 1 looked
up in the cumulative counts of topol["system"]["molecules"], stopping at the
first molecule without a residue) the residue's 1-based term lists offset by
the molecule's first row; angles in degrees -> np.radians; dihedral
coefficients from a propensity, inserted zero rows or as given; last = 0.

Pinning: the term lists of test/test_force.py:18-45, 103-127, 189-212 on the
dppc_single and alanine_octapeptide fixtures (tests/golden/
topology_goldens.json), and networkx itself on random molecular graphs
(adjacency order, BFS discovery order and parents, the set of simple 4-node
paths; tests/test_oracle_topology.py); the index-based path, which no
reference test covers, by hand-written expected arrays of a two-residue
topology, the stop at the first molecule without a residue, and the
propensity coefficients' closed forms (lambda = -1, 0, 1 give the alpha, coil
and beta tables; in between the first row of the helix or sheet table is
scaled by |lambda| and the coil's by 1 - |lambda|).
"""
import numpy as np


def _name(x):
    return x.decode("utf-8") if isinstance(x, bytes) else str(x)


class MolGraph:
    """networkx.Graph insertion order for one molecule: nodes keyed by
    global index, holding their row (local index) and name."""

    def __init__(self):
        self.order = []       # global ids in node order
        self.adj = {}         # gid -> [gid] in insertion order
        self.row = {}         # gid -> local index
        self.name = {}        # gid -> name

    def add_node(self, g):
        if g not in self.adj:
            self.adj[g] = []
            self.order.append(g)

    def add_edge(self, u, v):
        self.add_node(u)
        self.add_node(v)
        if v not in self.adj[u]:
            self.adj[u].append(v)
            if u != v:
                self.adj[v].append(u)

    def bfs2(self, s):
        """[(j, parent)] in discovery order for levels 1 and 2."""
        seen = {s}
        out = []
        for j in self.adj[s]:
            if j not in seen:
                seen.add(j)
                out.append((j, s))
        level1 = [j for j, _ in out]
        for v in level1:
            for w in self.adj[v]:
                if w not in seen:
                    seen.add(w)
                    out.append((w, v))
        return out

    def paths4(self, s):
        """find_all_paths(G, s, 3): [s, a, b, c] nested in neighbour order."""
        out = []
        for a in self.adj[s]:
            for b in self.adj[a]:
                for c in self.adj[b]:
                    if c != b and a != b and a != c and s != a and s != b and s != c:
                        out.append((s, a, b, c))
        return out


def molecule_graph(rows, indices, names, bonds):
    G = MolGraph()
    for r in rows:
        g = int(indices[r])
        G.add_node(g)
        G.row[g] = int(r)
        G.name[g] = _name(names[r])
        for b in bonds[r]:
            if b != -1:
                G.add_edge(g, int(b))
    return G


def prepare_bonds_old(molecules, names, bonds, indices, bond_types, angle_types, dihedral_types):
    """Term lists (force.py:325-505).  bond_types: (a1, a2, r0, k);
    angle_types: (a1, a2, a3, theta0 in degrees, k); dihedral_types:
    (a1, a2, a3, a4, coeffs, dih_type).  Returns bonds_2, bonds_3, bonds_4
    (with the type index instead of the coefficient object) and bb_index."""
    molecules = np.asarray(molecules)
    b2, b3, b4, bb_index = [], [], [], []
    for mol in np.unique(molecules):
        rows = np.nonzero(molecules == mol)[0]
        G = molecule_graph(rows, indices, names, bonds)

        def row(g):
            if g not in G.row:
                raise KeyError("name")   # a partner outside the molecule has no attributes
            return G.row[g]

        def nm(g):
            row(g)
            return G.name[g]

        bb = 0
        start = len(b4)
        for s in G.order:
            for j, v in G.bfs2(s):
                if j <= s:
                    continue
                if v == s:   # level 1: a bond
                    n2 = (nm(s), nm(j))
                    for bt in bond_types:
                        if n2 in ((bt[0], bt[1]), (bt[1], bt[0])):
                            b2.append([row(s), row(j), bt[2], bt[3]])
                else:        # level 2: an angle through v
                    n3 = (nm(s), nm(v), nm(j))
                    for at in angle_types:
                        if n3 in ((at[0], at[1], at[2]), (at[2], at[1], at[0])):
                            b3.append([row(s), row(v), row(j), np.radians(at[3]), at[4]])
            for p in G.paths4(s):
                names4 = tuple(nm(g) for g in p)
                for e, dt in enumerate(dihedral_types):
                    if names4 != tuple(dt[:4]):
                        continue
                    rev = [row(p[3]), row(p[2]), row(p[1]), row(p[0])]
                    stored = [t for t in b4[start:] if t[:4] == rev]
                    skip = False
                    for t in stored:      # the reference's list comparison, in stored order
                        if t[4] == e:
                            skip = True
                            break
                        raise ValueError("The truth value of an array with more than one element is ambiguous")
                    if skip:
                        continue
                    b4.append([row(p[0]), row(p[1]), row(p[2]), row(p[3]), e, dt[5]])
                    if dt[5] == 1:
                        bb = len(b4)
        if bb:
            bb_index.append(bb - 1)
    return b2, b3, b4, bb_index


def resize_coeffs(coeffs):
    """np.resize(coeffs, (6, 5)) (force.py:711): the flattened values repeated."""
    return np.resize(np.asarray(coeffs, dtype=np.float64), (6, 5))


def prepare_bonds(molecules, names, bonds, indices, bond_types, angle_types, dihedral_types):
    """force.py:573-728 (connectivity path): the 16 arrays, dtype int64 /
    float64 as the reference builds them."""
    b2, b3, b4, bb_index = prepare_bonds_old(molecules, names, bonds, indices, bond_types, angle_types,
                                             dihedral_types)
    a2 = np.array([t[:2] for t in b2], dtype=np.int64).reshape(-1, 2)
    p2 = np.array([t[2:] for t in b2], dtype=np.float64).reshape(-1, 2)
    a3 = np.array([t[:3] for t in b3], dtype=np.int64).reshape(-1, 3)
    p3 = np.array([t[3:] for t in b3], dtype=np.float64).reshape(-1, 2)
    a4 = np.array([t[:4] for t in b4], dtype=np.int64).reshape(-1, 4)
    coeff = np.array([resize_coeffs(dihedral_types[t[4]][4]) for t in b4], dtype=np.float64).reshape(-1, 6, 5)
    dtype = np.array([t[5] for t in b4], dtype=np.int64)
    last = np.zeros(len(b4), dtype=np.int64)
    last[bb_index] = 1
    return (a2[:, 0], a2[:, 1], p2[:, 0], p2[:, 1],
            a3[:, 0], a3[:, 1], a3[:, 2], p3[:, 0], p3[:, 1],
            a4[:, 0], a4[:, 1], a4[:, 2], a4[:, 3], coeff, dtype, last)


# ---------------------------------------------------------------------------
# index-based topology input (force.py:280-322, 508-570)
# ---------------------------------------------------------------------------
_ALPHA = [[7.406, -5.298, -2.570, 1.336, 0.739], [-0.28632126, 1.2099146, 1.18122138, 0.49075168, 0.98495911]]
_BETA = [[3.770, 5.929, -4.151, -0.846, 0.190], [-0.2300693, -0.0583289, 0.99342396, 1.03237971, 2.90160988]]
_COIL = [[1.416, -0.739, 0.990, -0.397, 0.136], [1.3495933, 0.45649087, 2.30441057, -0.12274901, -0.26179939]]


def propensity_potential_coeffs(x):
    """force.py:280-322: helical / coil / sheet cosine-series coefficients
    (4, 5) for a propensity lambda in [-1, 1] (None outside, as the reference
    logs an error and returns nothing)."""
    alpha, beta, coil = np.array(_ALPHA), np.array(_BETA), np.array(_COIL)
    zero = np.zeros((2, 5))
    if np.isclose(x, -1):
        return np.concatenate((alpha, zero))
    if np.isclose(x, 0):
        return np.concatenate((coil, zero))
    if np.isclose(x, 1):
        return np.concatenate((beta, zero))
    ax = abs(x)
    if ax > 1:
        return None
    coil[0] *= 1 - ax
    if x < 0:
        alpha[0] *= 0.5 * (ax - x)
        return np.concatenate((alpha, coil))
    beta[0] *= 0.5 * (ax + x)
    return np.concatenate((beta, coil))


def dihedral_coeffs_of_entry(dih_type, coeff):
    """The coefficient array of one topology dihedral entry (force.py:555-567)."""
    if dih_type == 0 and isinstance(coeff[0], (float, int)):
        return propensity_potential_coeffs(coeff[0])
    if dih_type == 1 and len(coeff) == 3:
        return np.array(propensity_potential_coeffs(coeff[0][0]).tolist() + list(coeff[1:]))
    if dih_type == 2:
        return np.array(coeff)
    return np.insert(np.array(coeff), 2, np.zeros((2, 5)), axis=0)


def prepare_index_based_bonds(molecules, topol):
    """force.py:508-570: bonds_2, bonds_3, bonds_4 from per-residue term
    lists with 1-based indices, offset by each molecule's first row."""
    molecules = np.asarray(molecules)
    b2, b3, b4 = [], [], []
    for mol in np.unique(molecules):
        resid = mol + 1
        resname, count = None, 0
        for name, k in topol["system"]["molecules"]:
            count += k
            if resid <= count:
                resname = name
                break
        if resname is None:
            break
        first = int(np.nonzero(molecules == mol)[0][0])
        res = topol[resname]
        for t in res.get("bonds", []):
            b2.append([t[0] - 1 + first, t[1] - 1 + first, t[3], t[4]])
        for t in res.get("angles", []):
            b3.append([t[0] - 1 + first, t[1] - 1 + first, t[2] - 1 + first, np.radians(t[4]), t[5]])
        for t in res.get("dihedrals", []):
            b4.append([t[0] - 1 + first, t[1] - 1 + first, t[2] - 1 + first, t[3] - 1 + first,
                       dihedral_coeffs_of_entry(t[4], t[5]), t[4], 0])
    return b2, b3, b4


def prepare_bonds_topol(molecules, topol):
    """force.py:573-728 with a topology: the 16 arrays."""
    b2, b3, b4 = prepare_index_based_bonds(molecules, topol)
    a2 = np.array([t[:2] for t in b2], dtype=np.int64).reshape(-1, 2)
    p2 = np.array([t[2:] for t in b2], dtype=np.float64).reshape(-1, 2)
    a3 = np.array([t[:3] for t in b3], dtype=np.int64).reshape(-1, 3)
    p3 = np.array([t[3:] for t in b3], dtype=np.float64).reshape(-1, 2)
    a4 = np.array([t[:4] for t in b4], dtype=np.int64).reshape(-1, 4)
    coeff = np.array([resize_coeffs(t[4]) for t in b4], dtype=np.float64).reshape(-1, 6, 5)
    dtype = np.array([t[5] for t in b4], dtype=np.int64)
    last = np.array([t[6] for t in b4], dtype=np.int64)
    return (a2[:, 0], a2[:, 1], p2[:, 0], p2[:, 1],
            a3[:, 0], a3[:, 1], a3[:, 2], p3[:, 0], p3[:, 1],
            a4[:, 0], a4[:, 1], a4[:, 2], a4[:, 3], coeff, dtype, last)
