"""The alternative fast sweeps, both off by default (slower than
k_sweep_fast on MI355X) and switched on per context by environment:

* k_sweep2_fast, two red-black sweeps in one pass (temporal blocking;
  MPS3D_SWEEP2=3 for both precisions): mps_smooth with two sweeps in fast
  arithmetic equals two reference sweeps (the oracle's
  perform_gseidel_smoothing restatement) to rounding;
* k_sweep_ref, one sweep of the reference operator with the boundary rows
  carried by staged ghost coefficients (MPS3D_SWEEP_REF=1): three sweeps
  equal three reference sweeps to rounding;

both on grids with partial tiles in x and y, many z-chunks, and Dirichlet
and no-flux faces in every direction.
* In FP32 (the default), the FP32 + FP64 defect-correction solve with the
  one-pass pairs follows the solve with single sweeps: the same defect
  history to FP32 rounding and the same converged solution.
"""
import os

import numpy as np
import pytest

from mps3d_amd import FieldSolver, ghosted
from mps3d_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _solver(monkeypatch, mask, *args, sweep_ref="0", **kw):
    # the library reads MPS3D_SWEEP2 and MPS3D_SWEEP_REF when the context is
    # created
    monkeypatch.setenv("MPS3D_SWEEP2", str(mask))
    monkeypatch.setenv("MPS3D_SWEEP_REF", sweep_ref)
    s = FieldSolver(*args, **kw)
    monkeypatch.delenv("MPS3D_SWEEP2")
    monkeypatch.delenv("MPS3D_SWEEP_REF")
    return s


@pytest.mark.parametrize("kernel,mask,sweep_ref,nsweeps", [("k_sweep2_fast", 3, "0", 2), ("k_sweep_ref", 0, "1", 3)])
@pytest.mark.parametrize("n,codes", [
    ((24, 20, 70), None),
    ((65, 33, 17), ("DIRC", "ZGRD", "FLUX", "DIRC", "DIRC", "ZGRD")),
    ((67, 9, 33), ("ZGRD", "DIRC", "DIRC", "ZGRD", "ZGRD", "DIRC")),
    ((40, 41, 96), ("DIRC", "DIRC", "DIRC", "DIRC", "DIRC", "DIRC")),
])
def test_fast_sweep_variant_matches_reference_sweeps(gpu, monkeypatch, n, codes, kernel, mask, sweep_ref, nsweeps):
    c = W.varcoef_poisson(n)
    rng = np.random.default_rng(9)
    if codes is not None:
        c.bc_codes = codes
        c.bc_vals = [rng.standard_normal(f.shape) for f in c.bc_vals]
    u0 = ghosted(rng.standard_normal(c.n))
    u0[0, :, :] = 7.0             # the caller's ghost values stay as they were
    P = O.Problem(c.n, c.length, (1, 1, 1), c.bc_codes, c.bc_vals, c.dcoeff, source=c.source)
    ref_u, ref_res = O.smooth(P, u0.copy(order="F"), nsweeps)
    print("%s: n=%s codes=%s" % (kernel, n, codes), flush=True)
    s = _solver(monkeypatch, mask, c.n, c.length, arith="fast", sweep_ref=sweep_ref)
    try:
        s.set_case(c)
        u = u0.copy(order="F")
        res = s.smooth(u, nsweeps)
    finally:
        s.close()
    scale = np.abs(ref_u[1:-1, 1:-1, 1:-1]).max()
    err = np.abs(u[1:-1, 1:-1, 1:-1] - ref_u[1:-1, 1:-1, 1:-1]).max() / scale
    assert err <= 1e-12, err
    assert np.all(u[0, :, :] == 7.0)
    assert np.abs(res - ref_res).max() <= 1e-10 * np.abs(ref_res).max()


@pytest.mark.parametrize("n", [65, 129])
def test_fp32dc_with_one_pass_pairs_follows_single_sweeps(gpu, monkeypatch, n):
    c = W.varcoef_poisson(n)
    out = {}
    for mask in (0, 1):
        s = _solver(monkeypatch, mask, c.n, c.length, hierarchy="aligned", precision="fp32dc", max_cycles=40,
                    tol=1e-10, tol_kind="rel", arith="fast", start="fmg", dc_inner=2, omega=1.0)
        try:
            s.set_case(c)
            u = np.zeros(s.shape, order="F")
            k, r, h, st = s.solve(u, raise_on_failure=True)
            out[mask] = (u[1:-1, 1:-1, 1:-1].copy(), h)
        finally:
            s.close()
    (u0, h0), (u1, h1) = out[0], out[1]
    assert len(h0) == len(h1)
    assert np.max(np.abs(h1 - h0) / h0) < 1e-3
    assert np.linalg.norm(u1 - u0) / np.linalg.norm(u0) < 1e-9
