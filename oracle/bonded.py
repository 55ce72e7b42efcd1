"""ctypes wrappers of the bonded-kernel checkers (oracle; test infrastructure only).

``cbf``/``caf``/``cdf`` call ``bonded.c`` (the C restatement of
compute_bond_forces*.f90, compute_angle_forces*.f90,
compute_dihedral_forces*.f90 and dipole_reconstruction*.f90); ``single=True``
gives the real(4) variants, ``single=False`` the __double ones.
All return values follow the f2py call surface of force.py:8-23:
``(f, energy, pr)`` for bonds/angles, ``(f, dipoles, tmat, energy)`` for
dihedrals, with C-order float64 arrays.

Pinning (tests/test_oracle_bonded.py): bonds/angles/dihedral energies and
forces against the reference's goldens (test/test_force.py:47-273; dihedral
forces there carry the opposite sign), central finite differences of the
energies, the cosine-series and CBT energies on geometries of known gamma and
phi, the dipole pair's geometry (delta, psi, theta(gamma) of the fit), the
transfer matrices against finite differences of the dipole displacement and
the redistribution by virtual work.  The reference's Fortran is reached only
through f2py's generated wrapper (f2py needs meson, absent here), so it is not
built.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboracle_bonded.so")
_lib = None

_d = ctypes.c_double
_i64 = ctypes.c_int64
_p = ctypes.c_void_p


def build():
    """Compile the C restatement."""
    subprocess.run(["make", "-s", "-C", _HERE, "oracle"], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        lib = ctypes.CDLL(_LIB)
        lib.oracle_cbf.restype = _d
        lib.oracle_cbf.argtypes = [_i64, _p, _p, _p, _i64, _p, _p, _p, _p, _p, ctypes.c_int]
        lib.oracle_caf.restype = _d
        lib.oracle_caf.argtypes = [_i64, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, ctypes.c_int]
        lib.oracle_cdf.restype = _d
        lib.oracle_cdf.argtypes = [_i64, _p, _p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, _p,
                                   ctypes.c_int, ctypes.c_int]
        lib.oracle_dipole_redistribution.restype = None
        lib.oracle_dipole_redistribution.argtypes = [_i64, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p]
        lib.oracle_clamp_count.restype = ctypes.c_long
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def cbf(r, box, a, b, r0, k, single=False):
    lib = _load()
    r = _f64(r)
    n = len(r)
    f = np.zeros((n, 3))
    pr = np.zeros(3)
    a, b = _i32(a), _i32(b)
    r0, k, box = _f64(r0), _f64(k), _f64(box)
    e = lib.oracle_cbf(n, _ptr(f), _ptr(r), _ptr(box), len(a), _ptr(a), _ptr(b), _ptr(r0),
                       _ptr(k), _ptr(pr), int(single))
    return f, e, pr


def caf(r, box, a, b, c, t0, k, single=False):
    lib = _load()
    r = _f64(r)
    n = len(r)
    f = np.zeros((n, 3))
    pr = np.zeros(3)
    a, b, c = _i32(a), _i32(b), _i32(c)
    t0, k, box = _f64(t0), _f64(k), _f64(box)
    e = lib.oracle_caf(n, _ptr(f), _ptr(r), _ptr(box), len(a), _ptr(a), _ptr(b), _ptr(c),
                       _ptr(t0), _ptr(k), _ptr(pr), int(single))
    return f, e, pr


def cdf(r, box, a, b, c, d, coeff, dtype, bb_index, dipole_flag, single=False):
    lib = _load()
    r = _f64(r)
    n = len(r)
    m = len(a)
    f = np.zeros((n, 3))
    dip = np.zeros((m, 4, 3))
    tm = np.zeros((m, 6, 3, 3))
    a, b, c, d = _i32(a), _i32(b), _i32(c), _i32(d)
    dtype, bb_index = _i32(dtype), _i32(bb_index)
    coeff = np.ascontiguousarray(coeff, dtype=np.float32).reshape(m, 6, 5)
    box = _f64(box)
    e = lib.oracle_cdf(n, _ptr(f), _ptr(r), _ptr(dip), _ptr(tm), _ptr(box), m, _ptr(a), _ptr(b),
                       _ptr(c), _ptr(d), _ptr(coeff), _ptr(dtype), _ptr(bb_index),
                       int(dipole_flag), int(single))
    return f, dip, tm, e


def clamp_count():
    """Number of sin(gamma) < 0.1 clamps seen (dipole_reconstruction.f90:112-117)."""
    return int(_load().oracle_clamp_count())


def dipole_forces_redistribution(n, f_dipoles, tmat, a, b, c, d, dtype, last):
    """force.py:855-877."""
    lib = _load()
    m = len(a)
    out = np.zeros((n, 3))
    fd = _f64(f_dipoles, (m, 4, 3))
    tm = _f64(tmat, (m, 6, 3, 3))
    a, b, c, d = _i32(a), _i32(b), _i32(c), _i32(d)
    dtype, last = _i32(dtype), _i32(last)
    lib.oracle_dipole_redistribution(n, _ptr(out), _ptr(fd), _ptr(tm), m, _ptr(a), _ptr(b),
                                     _ptr(c), _ptr(d), _ptr(dtype), _ptr(last))
    return out
