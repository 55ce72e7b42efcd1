"""hymd_cbf_noalloc (the allocation-free bond entry point) against hymd_cbf,
called through the C-ABI: two calls in a row, the second with fewer
particles and bonds, in both precisions.  The two entry points run the same
kernels over the same term lists, so forces, energy and virial must agree bit
for bit; the scratch array keeps the first call's contents, so the second call
also shows that everything that must start at zero is cleared.
"""
import ctypes

import numpy as np
import pytest
import torch

from hymd_amd import _lib
from hymd_amd._lib import HYMD_F32, HYMD_F64, darr, dptr

pytestmark = pytest.mark.gpu

MAX_N = 1 << 20


def _system(seed, n, m, box):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.0, 1.0, (n, 3)) * np.asarray(box)
    a = rng.integers(0, n, m).astype(np.int32)
    b = ((a + rng.integers(1, 9, m)) % n).astype(np.int32)
    r0 = rng.uniform(0.3, 0.6, m)
    k = rng.uniform(500.0, 1500.0, m)
    return r, a, b, r0, k


def _call(name, r, a, b, r0, k, box, precision):
    dt = torch.float32 if precision == HYMD_F32 else torch.float64
    dev = torch.device("cuda")
    rd = torch.as_tensor(r, dtype=dt, device=dev).contiguous()
    # a poisoned output: the callee must overwrite every element
    f = torch.full_like(rd, float("nan"))
    ad = torch.as_tensor(a, device=dev)
    bd = torch.as_tensor(b, device=dev)
    r0d = torch.as_tensor(r0, dtype=torch.float64, device=dev)
    kd = torch.as_tensor(k, dtype=torch.float64, device=dev)
    e = ctypes.c_double(float("nan"))
    pr = (ctypes.c_double * 3)(float("nan"), float("nan"), float("nan"))
    _lib.call(name, dptr(f), dptr(rd), rd.shape[0], darr(box), dptr(ad), dptr(bd), dptr(r0d), dptr(kd), ad.numel(),
              precision, ctypes.byref(e), pr, _lib.stream_ptr())
    torch.cuda.synchronize()
    return f.cpu().numpy(), e.value, np.array(pr[:])


@pytest.mark.parametrize("precision", [HYMD_F32, HYMD_F64])
def test_noalloc_matches_cbf_twice(hip, precision):
    box = (9.0, 10.0, 11.0)
    for seed, n, m in ((11, 200_000, 190_000), (12, 30_000, 20_000)):
        r, a, b, r0, k = _system(seed + precision, n, m, box)
        f0, e0, pr0 = _call("hymd_cbf", r, a, b, r0, k, box, precision)
        f1, e1, pr1 = _call("hymd_cbf_noalloc", r, a, b, r0, k, box, precision)
        assert np.isfinite(f0).all() and np.abs(f0).max() > 0
        np.testing.assert_array_equal(f1, f0)
        assert e1 == e0
        np.testing.assert_array_equal(pr1, pr0)


def test_noalloc_at_its_limits(hip):
    """The largest accepted sizes run; one particle or bond more is rejected
    with a message before any device work."""
    box = (40.0, 40.0, 40.0)
    r, a, b, r0, k = _system(7, MAX_N, MAX_N, box)
    f0, e0, pr0 = _call("hymd_cbf", r, a, b, r0, k, box, HYMD_F64)
    f1, e1, pr1 = _call("hymd_cbf_noalloc", r, a, b, r0, k, box, HYMD_F64)
    np.testing.assert_array_equal(f1, f0)
    assert e1 == e0
    np.testing.assert_array_equal(pr1, pr0)
    L = _lib.lib()
    e = ctypes.c_double()
    st = L.hymd_cbf_noalloc(ctypes.c_void_p(1), ctypes.c_void_p(1), MAX_N + 1, darr(box), None, None, None, None, 0,
                            HYMD_F64, ctypes.byref(e), None, None)
    assert st == -1 and b"allocation-free limit" in L.hymd_last_error()
    st = L.hymd_cbf_noalloc(ctypes.c_void_p(1), ctypes.c_void_p(1), 10, darr(box), ctypes.c_void_p(1),
                            ctypes.c_void_p(1), ctypes.c_void_p(1), ctypes.c_void_p(1), MAX_N + 1, HYMD_F64,
                            ctypes.byref(e), None, None)
    assert st == -1 and b"allocation-free limit" in L.hymd_last_error()
