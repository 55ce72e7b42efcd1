"""ctypes binding of ``libhymd_amd.so`` (the C-ABI declared in include/hymd_amd.h).

torch is imported first on purpose: its bundled HIP runtime has the same
SONAME as the system one, so the library binds to the runtime torch already
loaded and both share one HIP context, one set of streams and pointers.
There is no CPU fallback: a missing library raises ``HymdError``.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede loading the HIP library)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhymd_amd.so")

HYMD_OK = 0
HYMD_F32 = 0
HYMD_F64 = 1
HYMD_MAX_TYPES = 16
HYMD_FIXED_SHIFT = 32

_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_d = ctypes.c_double
_pd = ctypes.POINTER(ctypes.c_double)
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pi64 = ctypes.POINTER(ctypes.c_int64)

# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "hymd_last_error": [],
    "hymd_abi_version": [],
    "hymd_pm_create": [ctypes.POINTER(_p), _pi32, _pd, _i32, _p],
    "hymd_pm_destroy": [_p],
    "hymd_pm_set_stream": [_p, _p],
    "hymd_pm_set_box": [_p, _pd],
    "hymd_pm_set_tuning": [_p, _i32, _i32],
    "hymd_pm_get_tuning": [_p, _i32, _pi32],
    "hymd_update_field": [_p, _p, _p, _i64, _i32, _pd, _pd, _d, _pd, _d, _p, _p, _p, _p, _pd],
    "hymd_compute_field_force": [_p, _p, _p, _i64, _i32, _p, _p],
    "hymd_update_field_force_q": [_p, _p, _p, _i64, _d, _d, _d, _p, _p, _p, _p, _pd],
    "hymd_update_field_force_q_gpe": [_p, _p, _p, _p, _i64, _i32, _p, _pd, _pd, _d, _d, _d, _d, _i32, _i32, _p, _p,
                                      _p, _p, _p, _p, _p, _p, _pd, _pi32, _pd],
    "hymd_field_energy": [_p, _p, _i32, _pd, _pd, _d, _pd],
    "hymd_field_q_energy": [_p, _p, _p, _d, _pd],
    "hymd_self_energy_q": [_p, _i64, _d, _d, _pd, _p],
    "hymd_field_moments": [_p, _p, _p, _i64, _i32, _pd, _d, _pd, _pd, _pd, _p],
    "hymd_field_forces_held": [_p, _i32, _pd, _pd, _d, _p, _p],
    "hymd_field_adjoint": [_p, _p, _p, _i64, _i32, _pd, _d, _pd, _p, _i32, _pd, _pd],
    "hymd_field_pressure": [_p, _p, _i32, _pd, _pd, _d, _d, _p, _p, _pd, _d, _pd],
    "hymd_md_set_peptide": [_p, _p, _d, _d, _p, _i64, _p, _p],
    "hymd_md_peptide_state": [_p, _p, _p, _pd],
    "hymd_dihedrals_create": [ctypes.POINTER(_p), _i64, _i64, _p, _p, _p, _p, _p, _p, _p, _p],
    "hymd_dihedrals_destroy": [_p],
    "hymd_dihedrals_forces": [_p, _p, _p, _p, _p, _pd, _i32, _pd, _p],
    "hymd_dihedrals_redistribute": [_p, _p, _p, _p, _p],
    "hymd_dihedrals_dipoles": [_p, _p, _p, _p, _pd, _p],
    "hymd_batch_create": [ctypes.POINTER(_p), _i32, _pi32, _pd, _pi64, _i32, _pd],
    "hymd_batch_destroy": [_p],
    "hymd_batch_moments": [_p, _p, _p, _pd, _pd, _i32, _p],
    "hymd_batch_forces": [_p, _p, _p, _pd, _pd, _pd, _p, _i32, _p],
    "hymd_paint_fixed": [_p, _p, _p, _p, _i64, _i32, _pd, _p, _i32],
    "hymd_fft_r2c": [_p, _p, _p, _i32],
    "hymd_fft_c2r": [_p, _p, _p, _i32],
    "hymd_readout": [_p, _p, _p, _i64, _p],
    "hymd_cbf": [_p, _p, _i64, _pd, _p, _p, _p, _p, _i64, _i32, _pd, _pd, _p],
    "hymd_cbf_noalloc": [_p, _p, _i64, _pd, _p, _p, _p, _p, _i64, _i32, _pd, _pd, _p],
    "hymd_caf": [_p, _p, _i64, _pd, _p, _p, _p, _p, _p, _i64, _i32, _pd, _pd, _p],
    "hymd_cdf": [_p, _p, _p, _p, _i64, _pd, _p, _p, _p, _p, _p, _p, _p, _i64, _i32, _i32, _pd, _pi64, _p],
    "hymd_dipole_redistribution": [_p, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _i32, _p],
    "hymd_topology_create": [ctypes.POINTER(_p), _i64, _i64, _p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p],
    "hymd_topology_destroy": [_p],
    "hymd_topology_forces": [_p, _p, _p, _pd, _pd, _pd, _p],
    "hymd_respa_kick_drift": [_p, _p, _p, _i64, _p, _d, _p, _d, _d, _pd, _p],
    "hymd_respa_kick": [_p, _i64, _p, _d, _p, _d, _p],
    "hymd_kinetic_energy": [_p, _i64, _i32, _d, _pd, _p],
    "hymd_scale_positions": [_p, _i32, _p, _i64, _i64, _i64, _pd, _p],
    "hymd_md_create": [ctypes.POINTER(_p), _p, _p, _i64, _i32, _p, _p, _p, _p, _p, _p, _p, _pd, _pd, _d, _pd, _d,
                       _d, _d, _i32, _pd],
    "hymd_md_step": [_p, _i32, _i32, _pd, _p],
    "hymd_md_step_half": [_p, _i32, _pd, _pd, _p],
    "hymd_md_set_spectra": [_p, _p, _p, _p],
    "hymd_md_destroy": [_p],
    "hymd_md_time_stage": [_p, _i32, _i32],
    "hymd_md_set_fused": [_p, _i32],
    "hymd_md_set_term_precision": [_p, _i32],
    "hymd_md_set_thermostat": [_p, _i32, _pi32, _i32, _pi64, _d, _d, _d, _i32, _i32, _p, _p, _i64],
    "hymd_md_push_draws": [_p, _pd, _pd, _i32],
    "hymd_md_set_com_removal": [_p, _i32, _i64, _p, _i64],
    "hymd_md_set_step": [_p, _i64],
    "hymd_fused_set_precision": [_p, _i32],
    "hymd_md_is_fused": [_p],
    "hymd_md_set_reorder": [_p, _i32, _p, _p, _p],
    "hymd_md_single_tail": [_p],
    "hymd_fused_create": [ctypes.POINTER(_p), _i64, _i64, _p, _p, _p, _p, _i64, _p, _p, _p, _p, _p],
    "hymd_fused_create_dihedrals": [ctypes.POINTER(_p), _i64, _i64, _p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _i64,
                                    _p, _p, _p, _p, _p, _p, _p],
    "hymd_fused_destroy": [_p],
    "hymd_fused_inner": [_p, _p, _p, _p, _p, _p, _d, _d, _d, _i32, _i32, _pd, _pd, _pd, _p],
    "hymd_fused_inner_dihedrals": [_p, _p, _p, _p, _p, _p, _d, _d, _d, _i32, _i32, _pd, _pd, _pd, _p],
    "hymd_md_stage_time": [_p, _pd, _pi32],
    "hymd_thermostat_workspace_bytes": [],
    "hymd_csvr_thermostat": [_p, _p, _i64, _i32, _pi32, _i32, _pi64, _pd, _pd, _d, _d, _d, _d, _i32, _p, _p, _i64,
                             _p],
    "hymd_cancel_com_momentum": [_p, _i64, _i64, _p, _i64, _p],
    "hymd_bond_builder_create": [ctypes.POINTER(_p), _i64, _i64, _p, _p, _p, _p, _p, _i32, _pi32, _pd, _i32, _pi32,
                                 _pd, _i32, _pi32, _pi32, _pd, _i32, _pi64, _p],
    "hymd_bond_builder_fill": [_p] + [_p] * 19,
    "hymd_bond_builder_destroy": [_p],
    "hymd_index_bonds_create": [ctypes.POINTER(_p), _i64, _p, _p, _i32, _pi32, _pi32, _pd, _pi32, _pd, _pi32, _pd,
                                _pi32, _pi64, _p],
    "hymd_index_bonds_fill": [_p] + [_p] * 16,
    "hymd_index_bonds_destroy": [_p],
}
_RESTYPES = {"hymd_last_error": ctypes.c_char_p, "hymd_md_single_tail": ctypes.c_int64,
             "hymd_thermostat_workspace_bytes": ctypes.c_int64}


class HymdError(RuntimeError):
    """Raised for a non-zero status from the HIP library (or a missing library)."""


_lib = None


def lib():
    """Load the library once; raise HymdError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HymdError(f"{LIB_PATH} is missing: build it with `python -m hymd_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(status):
    if status != HYMD_OK:
        msg = lib().hymd_last_error()
        raise HymdError(f"status {status}: {msg.decode() if msg else ''}")


def call(name, *args):
    check(getattr(lib(), name)(*args))


def require_gpu():
    if not torch.cuda.is_available():
        raise HymdError("no HIP device visible: the MI355X path needs a GPU (there is no CPU fallback)")


# ---------------------------------------------------------------------------
# argument helpers
# ---------------------------------------------------------------------------
class DevPtr(ctypes.c_void_p):
    """A device pointer that keeps its tensor alive: a temporary tensor passed
    as dptr(f(x)) must not return to torch's allocator (and be handed to the
    next argument) before the library call that uses it."""


def dptr(t, dtype=None, name="tensor"):
    """Device pointer of a contiguous CUDA/HIP tensor (None -> NULL)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a device tensor")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    p = DevPtr(t.data_ptr())
    p.tensor = t
    return p


def darr(values, n=None):
    """Host double array (ctypes) from a sequence."""
    vals = [float(v) for v in values]
    if n is not None and len(vals) != n:
        raise ValueError(f"expected {n} values, got {len(vals)}")
    return (ctypes.c_double * len(vals))(*vals)


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return _p(s.cuda_stream)
