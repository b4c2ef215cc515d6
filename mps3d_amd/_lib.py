"""ctypes binding of the C ABI in include/mps3d.h.

Loading fails loudly (ImportError) when the built library is missing: there
is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MPS3D_AMD_LIB") or os.path.join(HERE, "lib", "libmps3d_hip.so")

# status codes
OK, NOT_CONVERGED, ERR_NAN, ERR_ARG, ERR_HIP, ERR_STATE, ERR_NOLIB = 0, 1, -1, -2, -3, -4, -5
# BC codes
BC_NONE, BC_DIRC, BC_ZGRD, BC_FLUX = 0, 1, 2, 3
BC_BY_NAME = {"DIRC": BC_DIRC, "ZGRD": BC_ZGRD, "FLUX": BC_FLUX}
# solver forms
HIER_REF, HIER_ALIGNED = 0, 1
PREC_FP64, PREC_FP32_DC = 0, 1
FLUX_UPWIND, FLUX_SG = 0, 1
TOL_RMS, TOL_REL_L2 = 0, 1
ARITH_EXACT, ARITH_FAST = 0, 1

# every entry point declared in include/mps3d.h
EXPORTS = (
    "mps_default_opts", "mps_create", "mps_destroy", "mps_set_bc", "mps_set_bc_str", "mps_set_coeffs",
    "mps_vcycle", "mps_solve", "mps_apply", "mps_apply_transpose", "mps_bvec", "mps_smooth",
    "mps_fill_ghosts", "mps_gmres", "mps_solve_transpose", "mps_last_timing", "mps_last_vcycles", "mps_stream", "mps_solve_rhs",
    "mps_efield", "mps_set_chemistry", "mps_species_transport", "mps_eenergy_transport",
    "mps_poisson_transport", "mps_electron_temp", "mps_ion_drift_faces",
    "mps_error_string", "mps_last_error", "mps_time_kernel", "mps_grad_velocity", "mps_time_stream",
    "mps_time_in_solve",
    "mps_get_grid", "mps_gmres_noalloc", "mps_restrict", "mps_prolong1",
    "mps_eden_faces", "mps_neutral_faces", "mps_eenergy_faces", "mps_floor", "mps_set_source",
)


_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lenx", C.c_double), ("leny", C.c_double), ("lenz", C.c_double),
                ("px", C.c_int), ("py", C.c_int), ("pz", C.c_int)]


class Chem(C.Structure):
    """mps_chem: chemistry tables, Fortran-order arrays, 1-based species."""
    _fields_ = [("nspecies", C.c_int), ("nreac", C.c_int), ("especnum", C.c_int), ("bgspecnum", C.c_int),
                ("ionspecmin", C.c_int), ("ionspecmax", C.c_int),
                ("reactants", _vp), ("products", _vp), ("k_arrh", _vp), ("elecenergy", _vp),
                ("mobility", _vp), ("dcoeff", _vp), ("molmass", _vp), ("charge", _vp),
                ("mob_kind", _vp), ("mob_terms", _vp), ("mob_coef", _vp), ("dcoeff_kind", _vp),
                ("rate_kind", _vp), ("rate_coef", _vp)]


class Opts(C.Structure):
    _fields_ = [("hierarchy", C.c_int), ("precision", C.c_int), ("flux", C.c_int),
                ("nu1", C.c_int), ("nu2", C.c_int), ("max_cycles", C.c_int),
                ("tol", C.c_double), ("tol_kind", C.c_int), ("arith", C.c_int), ("start", C.c_int),
                ("dc_inner", C.c_int), ("omega", C.c_double)]


_lib = None


class MpsError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (status %d)" % (msg, status))
        self.status = status


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("mps3d_amd HIP library not built (%s); run python -m mps3d_amd.build" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    L.mps_default_opts.argtypes = [C.POINTER(Opts)]
    L.mps_default_opts.restype = None
    L.mps_create.argtypes = [C.POINTER(_vp), C.POINTER(Grid), C.POINTER(Opts)]
    L.mps_destroy.argtypes = [_vp]
    L.mps_set_bc.argtypes = [_vp, C.POINTER(C.c_int), C.POINTER(_vp)]
    L.mps_set_bc_str.argtypes = [_vp, C.c_char_p, C.POINTER(_vp)]
    L.mps_set_coeffs.argtypes = [_vp] + [_vp] * 6 + [C.c_int, C.c_double]
    L.mps_vcycle.argtypes = [_vp, _vp, _vp, _dp]
    L.mps_solve.argtypes = [_vp, _vp, C.POINTER(C.c_int), _dp, _vp]
    L.mps_solve_rhs.argtypes = [_vp, _vp, _vp, C.POINTER(C.c_int), _dp, _vp]
    L.mps_apply.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.mps_apply_transpose.argtypes = [_vp, _vp, _vp]
    L.mps_bvec.argtypes = [_vp, _vp, C.c_int, _vp]
    L.mps_smooth.argtypes = [_vp, _vp, _vp, C.c_int, _vp]
    L.mps_fill_ghosts.argtypes = [_vp, _vp]
    L.mps_efield.argtypes = [_vp, _vp, _vp, C.c_double, C.c_int]
    L.mps_set_chemistry.argtypes = [_vp, _vp]
    L.mps_species_transport.argtypes = [_vp, C.c_int, C.c_double, C.c_int, _vp, _vp, _vp, C.c_double, C.c_double,
                                        _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.mps_eenergy_transport.argtypes = [_vp, C.c_double, C.c_int, C.c_double, _vp, _vp, _vp, C.c_double,
                                        C.c_double, _vp, _vp, _vp, _vp, _vp, _vp]
    L.mps_poisson_transport.argtypes = [_vp, C.c_double, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.mps_electron_temp.argtypes = [_vp, _vp, _vp, C.c_double, C.c_double, _vp]
    L.mps_ion_drift_faces.argtypes = [_vp, C.c_int, _vp, _vp, _vp, C.c_double, _vp]
    L.mps_gmres.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int), _dp]
    L.mps_solve_transpose.argtypes = [_vp, _vp, _vp, C.POINTER(C.c_int), _dp]
    L.mps_last_timing.argtypes = [_vp, _dp, C.POINTER(C.c_int)]
    L.mps_last_vcycles.argtypes = [_vp, C.POINTER(C.c_int)]
    L.mps_stream.argtypes = [_vp]
    L.mps_stream.restype = _vp
    L.mps_grad_velocity.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
    L.mps_time_kernel.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_char_p, C.c_int]
    L.mps_gmres_noalloc.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_double]
    L.mps_eden_faces.argtypes = [_vp, C.POINTER(C.c_int), _vp, _vp, C.c_double, _vp]
    L.mps_neutral_faces.argtypes = [_vp, C.POINTER(C.c_int), _vp, C.c_double, C.c_double, C.c_double, _vp]
    L.mps_eenergy_faces.argtypes = [_vp, C.POINTER(C.c_int), _vp, _vp, C.c_double, _vp, C.c_double, _vp]
    L.mps_floor.argtypes = [_vp, _vp, C.c_double]
    L.mps_set_source.argtypes = [_vp, _vp]
    L.mps_restrict.argtypes = [_vp, _vp, _vp]
    L.mps_prolong1.argtypes = [_vp, _vp, _vp]
    L.mps_get_grid.argtypes = [_vp, C.POINTER(Grid), C.POINTER(C.c_int)]
    L.mps_time_stream.argtypes = [C.c_int, C.c_longlong, C.c_int, _dp, _dp]
    L.mps_time_in_solve.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp]
    L.mps_error_string.argtypes = [C.c_int]
    L.mps_error_string.restype = C.c_char_p
    L.mps_last_error.argtypes = []
    L.mps_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(status, allow=(OK,)):
    if status in allow:
        return status
    L = load()
    raise MpsError(status, "%s: %s" % (L.mps_error_string(status).decode(), L.mps_last_error().decode()))
