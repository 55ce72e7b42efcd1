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
reversed dihedral ValueError, as the reference does.  With ``topol`` the
index-based input of force.py:508-570 is expanded on the device instead.
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


# Cosine-series coefficient tables of the helix, sheet and coil propensity
# dihedrals (force.py:280-302): the same numbers as the reference's table.
_ALPHA = [[7.406, -5.298, -2.570, 1.336, 0.739], [-0.28632126, 1.2099146, 1.18122138, 0.49075168, 0.98495911]]
_BETA = [[3.770, 5.929, -4.151, -0.846, 0.190], [-0.2300693, -0.0583289, 0.99342396, 1.03237971, 2.90160988]]
_COIL = [[1.416, -0.739, 0.990, -0.397, 0.136], [1.3495933, 0.45649087, 2.30441057, -0.12274901, -0.26179939]]


def propensity_potential_coeffs(x):
    """force.py:280-322: (4, 5) coefficients for a propensity lambda in
    [-1, 1]; outside it the reference logs an error and returns None."""
    alpha, beta, coil = np.array(_ALPHA), np.array(_BETA), np.array(_COIL)
    zero = np.zeros((2, 5))
    if np.isclose(x, -1):
        return np.concatenate((alpha, zero))
    if np.isclose(x, 0):
        return np.concatenate((coil, zero))
    if np.isclose(x, 1):
        return np.concatenate((beta, zero))
    if abs(x) > 1:
        return None
    coil[0] *= 1 - abs(x)
    if x < 0:
        alpha[0] *= 0.5 * (abs(x) - x)
        return np.concatenate((alpha, coil))
    beta[0] *= 0.5 * (abs(x) + x)
    return np.concatenate((beta, coil))


def _topol_coeffs(dih_type, coeff):
    """Coefficients of one topology dihedral entry (force.py:555-567)."""
    if dih_type == 0 and isinstance(coeff[0], (float, int)):
        return propensity_potential_coeffs(coeff[0])
    if dih_type == 1 and len(coeff) == 3:
        return np.array(propensity_potential_coeffs(coeff[0][0]).tolist() + list(coeff[1:]))
    if dih_type == 2:
        return np.array(coeff)
    return np.insert(np.array(coeff), 2, np.zeros((2, 5)), axis=0)


def _prepare_index_based(molecules, topol, dev, stream):
    """prepare_index_based_bonds (force.py:508-570) + prepare_bonds' arrays on
    the device: residue templates marshalled on the host, one thread per
    molecule expands them."""
    molecules = _host(molecules).astype(np.int64, copy=False).reshape(-1)
    uniq, first = np.unique(molecules, return_index=True)
    entries = topol["system"]["molecules"]
    names = [str(e[0]) for e in entries]
    cum = np.cumsum([int(e[1]) for e in entries]) if entries else np.zeros(0, np.int64)
    k = np.searchsorted(cum, uniq + 1, side="left")       # first entry with resid <= cumulative count
    bad = np.flatnonzero(k >= len(names))
    stop = int(bad[0]) if len(bad) else len(uniq)        # the reference stops at the first miss
    used = sorted({names[int(i)] for i in k[:stop]}, key=names.index)
    tmpl_id = {nm: t for t, nm in enumerate(used)}
    template_of = np.full(len(uniq), -1, dtype=np.int32)
    template_of[:stop] = [tmpl_id[names[int(i)]] for i in k[:stop]]
    off = np.zeros((3, len(used) + 1), dtype=np.int32)
    i2, p2, i3, p3, i4, c4, k4 = [], [], [], [], [], [], []
    for t, nm in enumerate(used):
        res = topol[nm]
        for b in res.get("bonds", []):
            i2.append([b[0] - 1, b[1] - 1])
            p2.append([b[3], b[4]])
        for a in res.get("angles", []):
            i3.append([a[0] - 1, a[1] - 1, a[2] - 1])
            p3.append([np.radians(a[4]), a[5]])
        for d in res.get("dihedrals", []):
            i4.append([d[0] - 1, d[1] - 1, d[2] - 1, d[3] - 1])
            c4.append(np.resize(np.asarray(_topol_coeffs(d[4], d[5]), dtype=np.float64), (6, 5)).ravel())
            k4.append(int(d[4]))
        off[:, t + 1] = [len(i2), len(i3), len(i4)]
    arr = [_c(np.array(i2, np.int32).reshape(-1, 2), ctypes.c_int32), _c(np.array(p2, np.float64).reshape(-1, 2),
           ctypes.c_double), _c(np.array(i3, np.int32).reshape(-1, 3), ctypes.c_int32),
           _c(np.array(p3, np.float64).reshape(-1, 2), ctypes.c_double), _c(np.array(i4, np.int32).reshape(-1, 4),
           ctypes.c_int32), _c(np.array(c4, np.float64).reshape(-1, 30), ctypes.c_double),
           _c(np.array(k4, np.int32), ctypes.c_int32)]
    offp, offa = _c(off, ctypes.c_int32)
    d_first = torch.as_tensor(first.astype(np.int64), device=dev)
    d_tmpl = torch.as_tensor(template_of, device=dev)
    counts = (ctypes.c_int64 * 3)()
    h = ctypes.c_void_p()
    _lib.call("hymd_index_bonds_create", ctypes.byref(h), len(uniq), dptr(d_first), dptr(d_tmpl), len(used), offp,
              *[a[0] for a in arr], counts, _lib.stream_ptr(stream))
    try:
        nb, na, nd = (int(c) for c in counts)
        out = [torch.empty(nb, dtype=torch.int32, device=dev), torch.empty(nb, dtype=torch.int32, device=dev),
               torch.empty(nb, dtype=torch.float64, device=dev), torch.empty(nb, dtype=torch.float64, device=dev),
               *[torch.empty(na, dtype=torch.int32, device=dev) for _ in range(3)],
               torch.empty(na, dtype=torch.float64, device=dev), torch.empty(na, dtype=torch.float64, device=dev),
               *[torch.empty(nd, dtype=torch.int32, device=dev) for _ in range(4)],
               torch.empty((nd, 6, 5), dtype=torch.float64, device=dev),
               torch.empty(nd, dtype=torch.int32, device=dev), torch.empty(nd, dtype=torch.int32, device=dev)]
        _lib.call("hymd_index_bonds_fill", h, *[dptr(t) for t in out])
        torch.cuda.current_stream().synchronize() if stream is None else stream.synchronize()
    finally:
        _lib.lib().hymd_index_bonds_destroy(h)
    return tuple(out)


def prepare_bonds(molecules, names, bonds, indices, config, topol=None, device=None, stream=None,
                  return_types=False):
    """force.py:573-728 on the device (see the module docstring).  With
    ``topol`` (a parsed TOML topology) the terms come from its residue
    templates (prepare_index_based_bonds, force.py:508-570) and ``names``,
    ``bonds``, ``indices`` are not used.  With ``return_types`` a 17th item
    holds the config entry index of every term, (type2, type3, type4)
    (connectivity path only)."""
    if topol is not None:
        _lib.require_gpu()
        dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        return _prepare_index_based(molecules, topol, dev, stream)
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
