"""The oracle (oracle/mps_oracle.c) against the reference's own outputs.

Fixtures in tests/golden/ were written by the reference Fortran/MPI library
itself (tests/golden/make_golden.py). The oracle is a one-rank restatement
built without FMA contraction, so single-rank results must agree bit for
bit; multi-rank residual histories differ only by the MPI_Allreduce sum
order (a few ulps).
"""
import numpy as np
import pytest

from conftest import case_from_golden, golden
from oracle import oracle as O


def problem(c, p=(1, 1, 1)):
    return O.Problem(c.n, c.length, p, c.bc_codes, c.bc_vals, c.dcoeff, c.velx, c.vely, c.velz, c.reac,
                     c.source, c.timederiv, c.dt)


@pytest.mark.parametrize("name", ["operator_mixed", "operator_poisson"])
def test_find_ax_and_bvec_bitwise(name):
    d = golden(name)
    c = case_from_golden(d)
    P = problem(c)
    ax, dg, of = O.apply(P, O.ghosted(d["u0"]))
    assert np.array_equal(ax, d["AX"])
    assert np.array_equal(dg, d["DIAG"])
    assert np.array_equal(of, d["OFFDIAG"])
    assert np.array_equal(O.bvec(P, d["sterm"], True), d["BVEC"])
    if "BVECNOBC" in d:
        assert np.array_equal(O.bvec(P, d["sterm"], False), d["BVECNOBC"])


def test_gseidel_smoothing_bitwise():
    d = golden("operator_mixed")
    P = problem(case_from_golden(d))
    u, res = O.smooth(P, O.ghosted(d["u0"]), 2, d["sterm"])
    assert np.array_equal(u[1:-1, 1:-1, 1:-1], d["SMOOTH_U"])
    assert np.array_equal(res, d["SMOOTH_RES"])


def test_transfers_and_ghosts_bitwise():
    d = golden("transfer")
    assert np.array_equal(O.restrict(d["fine"]), d["RESTR"])
    assert np.array_equal(O.prolong1(d["coarse_g"]), d["PROLONG"])
    names = ["BCLEFT", "BCRIGHT", "BCBOTTOM", "BCTOP", "BCBACK", "BCFRONT"]
    for i, nm in enumerate(names):
        assert np.array_equal(O.restrict_face(d["bc%d" % i]).ravel(order="F"), d[nm])
    P = problem(case_from_golden(d))
    assert np.array_equal(O.ghosts(P, d["ghost_in"]), d["GHOSTED"])


def test_gmres_bitwise():
    d = golden("gmres_cdr")
    P = problem(case_from_golden(d))
    for tag in ("full", "restarted"):
        m, nr, tol = d["cfg_" + tag]
        x = O.gmres(P, d["b"], O.ghosted(d["x0"]), int(m), int(nr), float(tol))
        assert np.array_equal(x[1:-1, 1:-1, 1:-1], d["X_" + tag]), tag


SOLVES = ["solve_laplace16_p1", "solve_laplace16_p8", "solve_varcoef16_p1", "solve_varcoef16_p8",
          "solve_mms20x12x8_p512", "solve_mms16x8x8_p1"]


@pytest.mark.parametrize("name", SOLVES)
def test_vcycle_trajectory(name):
    d = golden(name)
    c = case_from_golden(d)
    p = tuple(int(v) for v in d["p"])
    P = problem(c, p)
    u, hist = O.solve(P, P.zeros(), int(d["ncyc"]), float(d["tol"]))
    ref = d["hist"]
    assert len(hist) == len(ref)
    if p == (1, 1, 1):
        assert np.array_equal(hist, ref)
        assert np.array_equal(u[1:-1, 1:-1, 1:-1], d["U"])
    else:
        # one rank vs an MPI_Allreduce of partial sums: ulp-level history
        assert np.max(np.abs(hist - ref) / ref) < 1e-14
        assert np.array_equal(u[1:-1, 1:-1, 1:-1], d["U"])


def test_sg_flux_matches_reference():
    d = golden("sg1d")
    got = np.array([O.sg_flux(*row) for row in d["rows"]])
    assert np.array_equal(got, d["out"])
