import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def pytest_collection_modifyitems(config, items):
    try:
        import torch
        have_gpu = torch.cuda.is_available()
    except Exception:  # pragma: no cover
        have_gpu = False
    if have_gpu:
        return
    skip = pytest.mark.skip(reason="no HIP device in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def goldens():
    with open(os.path.join(GOLDEN, "reference_goldens.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def oracle_built():
    """Compile oracle/bonded.c."""
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)
    from oracle import bonded
    return bonded


@pytest.fixture(scope="session")
def hip():
    """The HIP library on a GPU (fails loudly if it is not built)."""
    import torch
    from hymd_amd import _lib
    assert torch.cuda.is_available()
    _lib.lib()
    return _lib


def _pair(a, b):
    """Promote to float64 or complex128 (complex data keeps its imaginary part)."""
    a = np.asarray(a)
    b = np.asarray(b)
    kind = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    return a.astype(kind), b.astype(kind)


def rel_err(a, b):
    """max |a - b| / max |b| (moduli for complex data)."""
    a, b = _pair(a, b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def rms_rel(a, b):
    """sqrt(mean |a - b|^2) / sqrt(mean |b|^2) (moduli for complex data)."""
    a, b = _pair(a, b)
    return float(np.sqrt(np.mean(np.abs(a - b) ** 2)) / max(np.sqrt(np.mean(np.abs(b) ** 2)), 1e-300))
