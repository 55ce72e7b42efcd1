"""Spatially ordered particle layout of the MD state.

The field path bins particles by mesh row every outer step; when the state
arrays are already in spatial order (whole molecules ordered by the mesh
cell of their first particle, x-major rows as the binning uses them, z
within a row), the binning's scatter and the readout's force writes touch
neighbouring addresses, and the row counts of consecutive particles
coincide.  A molecule (particles connected by any bonded term) stays whole,
with its particles in their original relative order, so the molecule-blocked
inner loop applies to any input order.  The reference keeps molecules whole
per rank in the same spirit (file_io.py:781-874).
"""
import dataclasses

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components


def molecule_labels(n, bonds=None, angles=None, dihedrals=None):
    """Connected components of the bonded graph: (n_molecules, label per particle)."""
    rows, cols = [], []
    if bonds is not None and len(bonds[0]):
        rows.append(np.asarray(bonds[0]))
        cols.append(np.asarray(bonds[1]))
    if angles is not None and len(angles[0]):
        rows += [np.asarray(angles[0]), np.asarray(angles[1])]
        cols += [np.asarray(angles[1]), np.asarray(angles[2])]
    if dihedrals is not None and len(dihedrals[0]):
        rows += [np.asarray(dihedrals[0]), np.asarray(dihedrals[1]), np.asarray(dihedrals[2])]
        cols += [np.asarray(dihedrals[1]), np.asarray(dihedrals[2]), np.asarray(dihedrals[3])]
    r = np.concatenate(rows) if rows else np.zeros(0, dtype=np.int64)
    c = np.concatenate(cols) if cols else np.zeros(0, dtype=np.int64)
    g = coo_matrix((np.ones(len(r), dtype=np.int8), (r, c)), shape=(n, n))
    return connected_components(g, directed=False)


def spatial_order(positions, box, mesh, bonds=None, angles=None, dihedrals=None, singles_last=True):
    """New-to-old particle order (see the module docstring).  With
    `singles_last`, molecules of several particles come first and single
    particles (solvent) after them, each group in spatial order: blocks of the
    molecule-blocked inner loop then hold either bonded molecules or solvent,
    not a mix whose wavefronts all pay for the bonded terms."""
    x = np.asarray(positions, dtype=np.float64)
    n = len(x)
    if n == 0:
        return np.zeros(0, dtype=np.int64)
    nmol, lab = molecule_labels(n, bonds, angles, dihedrals)
    first = np.full(nmol, n, dtype=np.int64)
    np.minimum.at(first, lab, np.arange(n))
    box = np.asarray(box, dtype=np.float64)
    mesh = np.asarray(mesh, dtype=np.int64)
    cell = np.floor(np.mod(x[first], box) / box * mesh).astype(np.int64)
    cell = np.clip(cell, 0, mesh - 1)
    key = (cell[:, 0] * mesh[1] + cell[:, 1]) * mesh[2] + cell[:, 2]
    if singles_last:
        single = np.bincount(lab, minlength=nmol) == 1
        key = key + single.astype(np.int64) * int(mesh.prod())
    rank = np.empty(nmol, dtype=np.int64)
    rank[np.argsort(key, kind="stable")] = np.arange(nmol)
    return np.lexsort((np.arange(n), rank[lab]))


def permute_system(system, order):
    """A copy of `system` with its particles in `order` (new-to-old) and every
    bonded term's indices renumbered; term order is unchanged.  Returns the
    copy and the old-to-new map."""
    s = system
    order = np.asarray(order, dtype=np.int64)
    inv = np.empty(len(order), dtype=np.int64)
    inv[order] = np.arange(len(order))

    def renum(t, k):
        if t is None:
            return None
        return tuple(inv[np.asarray(v)].astype(np.int32) for v in t[:k]) + tuple(t[k:])

    extra = dict(s.extra)
    if "dihedrals" in extra and extra["dihedrals"] is not None:
        extra["dihedrals"] = renum(extra["dihedrals"], 4)
    return dataclasses.replace(
        s, positions=np.asarray(s.positions)[order], velocities=np.asarray(s.velocities)[order],
        types=np.asarray(s.types)[order], bonds=renum(s.bonds, 2), angles=renum(s.angles, 3),
        charges=None if s.charges is None else np.asarray(s.charges)[order], extra=extra), inv
