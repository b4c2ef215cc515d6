"""mps_gmres_noalloc: the reference GMRES (precond 0) with its work space in
the library's static device array, called through the C ABI, against
mps_gmres with precond = 0 on the same inputs (bit for bit), two calls in a
row on one GPU, the second on a smaller grid; and the size limit refused on
the host."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import case_from_golden, golden
from mps3d_amd import FieldSolver, _lib as M, ghosted, to_device
from mps3d_amd import workloads as W

pytestmark = pytest.mark.gpu


def _interior_dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).transpose(2, 1, 0))).cuda()


def _both(case, b, x0, m, restarts, tol):
    """x from mps_gmres(precond=0) (host arrays) and from mps_gmres_noalloc
    (device arrays), same context, same inputs."""
    with FieldSolver(case.n, case.length) as s:
        s.set_case(case)
        xa = ghosted(x0)
        s.gmres(b, xa, m=m, restarts=restarts, tol=tol, precond=0, raise_on_failure=True)
        bd = _interior_dev(b)
        xd = to_device(ghosted(x0))
        torch.cuda.synchronize()
        st = s.lib.mps_gmres_noalloc(s._h, C.c_void_p(bd.data_ptr()), C.c_void_p(xd.data_ptr()), int(m),
                                     int(restarts), float(tol))
        M.check(st)
        xb = xd.cpu().numpy().transpose(2, 1, 0)
    return xa, xb


def test_noalloc_gmres_equals_the_allocating_entry_point_twice_in_a_row(gpu):
    # first call: 8^3 random convection-diffusion-reaction rows, GMRES(64)
    big = W.random_cdr((8, 8, 8), seed=29)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(big.n)
    xa, xb = _both(big, b, np.zeros(big.n), 64, 1, 1e-15)
    print("noalloc: 8^3 GMRES(64)", flush=True)
    assert np.array_equal(xa, xb)
    assert np.all(np.isfinite(xb))
    # second call, smaller: the reference fixture's 4^3 system, GMRES(6) x 8
    d = golden("gmres_cdr")
    small = case_from_golden(d)
    m, nr, tol = d["cfg_restarted"]
    xa, xb = _both(small, d["b"], d["x0"], int(m), int(nr), float(tol))
    print("noalloc: 4^3 GMRES(6) x 8", flush=True)
    assert np.array_equal(xa, xb)
    got = xb[1:-1, 1:-1, 1:-1]
    ref = d["X_restarted"]
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-12


def test_noalloc_gmres_refuses_work_beyond_its_static_space(gpu):
    case = W.random_cdr((8, 8, 8), seed=3)
    with FieldSolver(case.n, case.length) as s:
        s.set_case(case)
        bd = torch.zeros((8, 8, 8), dtype=torch.float64, device="cuda")
        xd = torch.zeros((10, 10, 10), dtype=torch.float64, device="cuda")
        # GMRES(600) on 10^3 ghosted cells needs ~610k doubles > 2^19
        st = s.lib.mps_gmres_noalloc(s._h, C.c_void_p(bd.data_ptr()), C.c_void_p(xd.data_ptr()), 600, 1, 1e-12)
        assert st == M.ERR_ARG
        assert b"static work space" in s.lib.mps_last_error()
