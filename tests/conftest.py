import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def case_from_golden(d):
    """Rebuilds reference-layout inputs stored by tests/golden/make_golden.py."""
    from mps3d_amd.workloads import Case
    n = tuple(int(v) for v in d["n"])
    codes = tuple(str(s).strip() for s in d["bc_codes"])
    faces = [np.asfortranarray(d["bc%d" % i]) for i in range(6)]
    c = Case(n=n, length=tuple(float(v) for v in d["length"]), bc_codes=codes, bc_vals=faces,
             timederiv=bool(int(d["timederiv"])), dt=float(d["dt"]))
    for k in ("dcoeff", "velx", "vely", "velz", "reac", "source"):
        setattr(c, k, np.asfortranarray(d[k]))
    return c


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a GPU and the built library are present; the library must load."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from mps3d_amd import _lib
    _lib.load()  # raises if the HIP library is missing: no fallback
    torch.cuda.set_device(0)
    return torch.device("cuda:0")
