"""Bonded term arrays from the topology, on the device (force.py:573-728,
prepare_bonds over prepare_bonds_old :325-505, via ``hymd_bond_builder_*``).

``prepare_bonds(molecules, names, bonds, indices, config)`` takes the
reference's arguments (host arrays as read from the input file, or device
tensors) and returns its 16 arrays as device tensors, in the reference's
order, ready for ``BondedTopology`` / ``DihedralTopology``:

    bonds_2_atom1, bonds_2_atom2, bonds_2_equilibrium, bonds_2_strength,
    bonds_3_atom1, bonds_3_atom2, bonds_3_atom3, bonds_3_equilibrium,
    bonds_3_strength, bonds_4_atom1, ..._atom4, bonds_4_coeff (D, 6, 5),
    bonds_4_type, bonds_4_last

Index arrays are int32 (the kernels' index type; the reference's are
int64), parameters float64.  The host part is marshalling only: name ids,
the molecule segments (np.unique order, rows in order) and the type tables
(np.radians of the angle equilibria and np.resize of the dihedral
coefficients to (6, 5), force.py:711, computed as the reference computes
them); the graph traversal runs on the device, one thread per molecule.
A bond partner outside its molecule raises KeyError and an ambiguous
reversed dihedral ValueError, as the reference does.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import HymdError, dptr


def _name(x):
    return x.decode("utf-8") if isinstance(x, bytes) else str(x)


def _field(obj, name, i):
    return getattr(obj, name) if hasattr(obj, name) else obj[i]


def _type_tables(config, name_id):
    """(names, params) host arrays of config.bonds / angle_bonds / dihedrals."""
    def nid(x):
        return name_id.get(_name(x), -1)

    bonds = list(getattr(config, "bonds", None) or [])
    angles = list(getattr(config, "angle_bonds", None) or [])
    dihedrals = list(getattr(config, "dihedrals", None) or [])
    n2 = np.array([[nid(_field(b, "atom_1", 0)), nid(_field(b, "atom_2", 1))] for b in bonds],
                  dtype=np.int32).reshape(-1, 2)
    p2 = np.array([[_field(b, "equilibrium", 2), _field(b, "strength", 3)] for b in bonds],
                  dtype=np.float64).reshape(-1, 2)
    n3 = np.array([[nid(_field(a, "atom_1", 0)), nid(_field(a, "atom_2", 1)), nid(_field(a, "atom_3", 2))]
                   for a in angles], dtype=np.int32).reshape(-1, 3)
    p3 = np.array([[np.radians(_field(a, "equilibrium", 3)), _field(a, "strength", 4)] for a in angles],
                  dtype=np.float64).reshape(-1, 2)
    n4 = np.array([[nid(_field(d, "atom_%d" % (k + 1), k)) for k in range(4)] for d in dihedrals],
                  dtype=np.int32).reshape(-1, 4)
    k4 = np.array([int(_field(d, "dih_type", 5)) for d in dihedrals], dtype=np.int32)
    c4 = np.array([np.resize(np.asarray(_field(d, "coeffs", 4), dtype=np.float64), (6, 5)) for d in dihedrals],
                  dtype=np.float64).reshape(-1, 30)
    return n2, p2, n3, p3, n4, k4, c4


def _host(a):
    return a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def _c(a, ctype):
    a = np.ascontiguousarray(a)
    return a.ctypes.data_as(ctypes.POINTER(ctype)) if a.size else None, a


def prepare_bonds(molecules, names, bonds, indices, config, topol=None, device=None, stream=None,
                  return_types=False):
    """force.py:573-728 on the device (see the module docstring).  With
    ``return_types`` a 17th item holds the config entry index of every term,
    (type2, type3, type4)."""
    if topol is not None:
        raise NotImplementedError("the index-based topology input (prepare_index_based_bonds, force.py:508-570) "
                                  "is not built here; pass the connectivity table")
    _lib.require_gpu()
    dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
    molecules = _host(molecules).astype(np.int64, copy=False)
    indices = _host(indices).astype(np.int64, copy=False)
    bonds = _host(bonds).astype(np.int64, copy=False)
    if bonds.ndim == 1:
        bonds = bonds[:, None]
    names = _host(names)
    if names.dtype.kind == "S":
        names = np.char.decode(names, "utf-8")
    n = len(molecules)
    if not (len(indices) == len(names) == n and bonds.shape[0] == n):
        raise ValueError("molecules, names, bonds and indices must have one row per particle")
    # molecule segments: np.unique order, rows in order within a molecule
    order = np.argsort(molecules, kind="stable")
    mol_sorted = molecules[order]
    starts = np.flatnonzero(np.r_[True, mol_sorted[1:] != mol_sorted[:-1]]) if n else np.zeros(0, np.int64)
    mol_start = np.r_[starts, n].astype(np.int64)
    n_mol = len(starts)
    uniq, name_ids = np.unique(names.astype(str), return_inverse=True)
    name_id = {str(u): i for i, u in enumerate(uniq)}
    name_ids = name_ids.astype(np.int32).reshape(-1)
    n2, p2, n3, p3, n4, k4, c4 = _type_tables(config, name_id)

    def up(a, dtype):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype), device=dev)

    d_start = up(mol_start, np.int64)
    d_gid = up(indices[order], np.int64)
    d_row = up(order, np.int64)
    d_name = up(name_ids[order], np.int32)
    d_bonds = up(bonds[order], np.int64)
    args = []
    for a, ct in ((n2, ctypes.c_int32), (p2, ctypes.c_double), (n3, ctypes.c_int32), (p3, ctypes.c_double),
                  (n4, ctypes.c_int32), (k4, ctypes.c_int32), (c4, ctypes.c_double)):
        args.append(_c(a, ct))
    counts = (ctypes.c_int64 * 3)()
    h = ctypes.c_void_p()
    L = _lib.lib()
    st = L.hymd_bond_builder_create(ctypes.byref(h), n, n_mol, dptr(d_start), dptr(d_gid), dptr(d_row),
                                    dptr(d_name), dptr(d_bonds), int(bonds.shape[1]),
                                    args[0][0], args[1][0], len(n2), args[2][0], args[3][0], len(n3),
                                    args[4][0], args[5][0], args[6][0], len(n4), counts, _lib.stream_ptr(stream))
    if st != _lib.HYMD_OK:
        msg = L.hymd_last_error().decode()
        if msg.startswith("KeyError"):
            raise KeyError(msg)
        if msg.startswith("ValueError"):
            raise ValueError(msg)
        raise HymdError(f"status {st}: {msg}")
    try:
        nb, na, nd = (int(c) for c in counts)

        def i32(m):
            return torch.empty(m, dtype=torch.int32, device=dev)

        def f64(*shape):
            return torch.empty(shape, dtype=torch.float64, device=dev)

        out = [i32(nb), i32(nb), f64(nb), f64(nb),
               i32(na), i32(na), i32(na), f64(na), f64(na),
               i32(nd), i32(nd), i32(nd), i32(nd), f64(nd, 6, 5), i32(nd), i32(nd)]
        types = [i32(nb), i32(na), i32(nd)]
        _lib.call("hymd_bond_builder_fill", h, *[dptr(t) for t in out], *[dptr(t) for t in types])
        torch.cuda.current_stream().synchronize() if stream is None else stream.synchronize()
    finally:
        L.hymd_bond_builder_destroy(h)
    return tuple(out) + ((tuple(types),) if return_types else ())


__all__ = ["prepare_bonds"]
