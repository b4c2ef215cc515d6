"""The C ABI library loads and exports every entry point include/mps3d.h declares.

No compute call is made here (no GPU in this environment); only host-side
paths: option defaults, error strings and the missing-device status.
"""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT
from mps3d_amd import _lib as M


def declared_symbols():
    with open(os.path.join(ROOT, "include", "mps3d.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(mps_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert sorted(M.EXPORTS) == declared_symbols()


def test_library_exports_every_declared_symbol():
    lib = M.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_default_options_are_the_reference_constants():
    lib = M.load()
    o = M.Opts()
    lib.mps_default_opts(C.byref(o))
    # nsmoothsteps = 2 (mgridsteps.f90:473), 20 V-cycles, ERRTOLR = 1e-11
    assert (o.nu1, o.nu2, o.max_cycles, o.tol) == (2, 2, 20, 1e-11)
    assert (o.hierarchy, o.precision, o.flux, o.tol_kind) == (M.HIER_REF, M.PREC_FP64, M.FLUX_UPWIND, M.TOL_RMS)
    assert (o.arith, o.start) == (0, 0)  # exact rows, V-cycles from the caller's guess
    assert o.dc_inner == 0                # FP32 cycles per FP64 defect step: chosen per step
    assert o.omega == 0.0                 # fast solves' sweep relaxation: the library's default


def test_error_strings():
    lib = M.load()
    for code in (M.OK, M.NOT_CONVERGED, M.ERR_NAN, M.ERR_ARG, M.ERR_HIP, M.ERR_STATE, M.ERR_NOLIB):
        assert lib.mps_error_string(code).decode()


def test_create_rejects_bad_grid_before_touching_a_device():
    lib = M.load()
    h = C.c_void_p()
    g = M.Grid(0, 4, 4, 1.0, 1.0, 1.0, 1, 1, 1)
    assert lib.mps_create(C.byref(h), C.byref(g), None) == M.ERR_ARG
    assert not h.value


def test_null_context_is_an_argument_error():
    lib = M.load()
    assert lib.mps_solve(None, None, None, None, None) == M.ERR_ARG
    assert lib.mps_destroy(None) == M.OK


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU device node is present")
def test_create_without_device_reports_no_gpu():
    lib = M.load()
    h = C.c_void_p()
    g = M.Grid(4, 4, 4, 1.0, 1.0, 1.0, 1, 1, 1)
    assert lib.mps_create(C.byref(h), C.byref(g), None) == M.ERR_NOLIB


def test_wrong_sized_device_face_is_refused_before_any_library_call():
    """set_bc copies ny*nz / nx*nz / nx*ny values from a device face; a face of
    another size (here a left face shaped (nz, nx) on a non-cubic grid) is
    refused on the host, before the library (or the GPU) is touched."""
    torch = pytest.importorskip("torch")
    from mps3d_amd import FieldSolver

    class NoLib:
        def __getattr__(self, name):
            raise AssertionError("library called: %s" % name)

    s = FieldSolver.__new__(FieldSolver)
    s.n = (6, 4, 5)
    s.lib = NoLib()
    s._keep, s._ext, s._torch_call, s._h = [], None, False, None
    bad = torch.zeros(5, 6, dtype=torch.float64)        # (nz, nx): 30 values, left face needs ny*nz = 20
    with pytest.raises(ValueError, match="face 0 has 30 elements, expected 20"):
        s.set_bc(("DIRC",) * 6, (bad, 0.0, 0.0, 0.0, 0.0, 0.0))
    ok_left = torch.zeros(4, 5, dtype=torch.float64)
    bad_front = torch.zeros(6 * 4 + 1, dtype=torch.float64)
    with pytest.raises(ValueError, match="face 5 has 25 elements, expected 24"):
        s.set_bc(("DIRC",) * 6, (ok_left, 0.0, 0.0, 0.0, 0.0, bad_front))
