"""The oracle against what the mathematics fixes, independently of the
reference build: the Laplace deck's exact discrete solution
(tests/Laplace_test/README:1-4 of the reference), a dense brute-force
assembly of the operator from its face rules (mgridlib/discretepde.f90:620-670
and the boundary rows :798-865), symmetry of the diffusion operator,
transfer invariants (mgridlib/mgridsteps.f90:11-53, :170-216), the RMS
definition (mgridlib/globalvars.f90:154-156) and GMRES with a full Krylov
space against a dense solve (mgridlib/gmres.f90:197-434).

Each check is chosen so that a plausible slip in the oracle -- a sign, a
transposed neighbour, a missing boundary term, a swapped child -- fails it.
"""
import numpy as np
import pytest

from oracle import oracle as O

rng = np.random.default_rng(20261202)


def dense_from_columns(P):
    """A by applying the oracle's find_AX to unit vectors (zero ghosts)."""
    nx, ny, nz = P.n
    N = nx * ny * nz
    A = np.zeros((N, N))
    for col in range(N):
        e = np.zeros(N)
        e[col] = 1.0
        ax, _, _ = O.apply(P, O.ghosted(e.reshape(P.n, order="F")))
        A[:, col] = ax.ravel(order="F")
    return A


def dense_from_faces(n, h, D, v, reac, codes, dtinv):
    """Independent assembly from the face rules: face mean k = (kL+kR)/2
    gives k/h^2 to both diagonals and -k/h^2 couplings; the upwind face
    velocity v = (vL+vR)/2 moves v/h onto the upstream cell's diagonal
    (v > 0: left) and -v/h onto the downstream row's coupling; DIRC rows add
    2 D/h^2 plus outflow v/h, ZGRD rows add the outflow term only, FLUX
    rows nothing."""
    nx, ny, nz = n
    idx = np.arange(nx * ny * nz).reshape(n, order="F")
    A = np.zeros((idx.size, idx.size))
    for ax in range(3):
        hh = h[ax]
        vel = v[ax]
        sl = [slice(None)] * 3
        sr = [slice(None)] * 3
        sl[ax], sr[ax] = slice(0, -1), slice(1, None)
        L, R = idx[tuple(sl)].ravel(), idx[tuple(sr)].ravel()
        kL, kR = D[tuple(sl)].ravel(), D[tuple(sr)].ravel()
        vL, vR = vel[tuple(sl)].ravel(), vel[tuple(sr)].ravel()
        k = 0.5 * (kL + kR) / hh ** 2
        vb = 0.5 * (vL + vR) / hh
        for a, b, kk, vv in zip(L, R, k, vb):
            A[a, a] += kk
            A[b, b] += kk
            A[a, b] -= kk
            A[b, a] -= kk
            if vv > 0:
                A[a, a] += vv
                A[b, a] -= vv
            else:
                A[b, b] -= vv
                A[a, b] += vv
        for side, code in ((0, codes[2 * ax]), (1, codes[2 * ax + 1])):
            s = [slice(None)] * 3
            s[ax] = slice(0, 1) if side == 0 else slice(-1, None)
            cells = idx[tuple(s)].ravel()
            Db, vbd = D[tuple(s)].ravel(), vel[tuple(s)].ravel()
            out = vbd > 0 if side == 1 else vbd < 0
            sign = 1.0 if side == 1 else -1.0
            for c, d_, w, o in zip(cells, Db, vbd, out):
                if code == "DIRC":
                    A[c, c] += 2.0 * d_ / hh ** 2 + (sign * w / hh if o else 0.0)
                elif code == "ZGRD":
                    A[c, c] += sign * w / hh
    A[np.diag_indices_from(A)] += -reac.ravel(order="F") + dtinv
    return A


def test_operator_matches_brute_force_face_assembly():
    n = (5, 4, 3)
    h = (0.3 / 5, 0.2 / 4, 0.5 / 3)
    D = 1.0 + rng.random(n)
    v = [rng.standard_normal(n) for _ in range(3)]
    reac = 0.1 * rng.random(n)
    codes = ("DIRC", "ZGRD", "FLUX", "DIRC", "ZGRD", "DIRC")
    P = O.Problem(n, (0.3, 0.2, 0.5), (1, 1, 1), codes, (0.0,) * 6, D, *v, reac, None, True, 0.01)
    A = dense_from_columns(P)
    B = dense_from_faces(n, h, D, v, reac, codes, 100.0)
    assert np.allclose(A, B, rtol=1e-13, atol=1e-12 * np.abs(B).max())


def test_laplace_deck_linear_solution_is_exact():
    """phi = x at cell centres solves the deck's discrete system: x faces
    DIRC 0 / 1, zero gradient elsewhere, D = 1 (the DIRC ghost mirror makes
    the boundary rows exact too)."""
    n = (8, 8, 4)
    codes = ("DIRC", "DIRC", "ZGRD", "ZGRD", "ZGRD", "ZGRD")
    P = O.Problem(n, (1.0, 1.0, 1.0), (1, 1, 1), codes, (0.0, 1.0, 0.0, 0.0, 0.0, 0.0), 1.0)
    x = (np.arange(n[0]) + 0.5) / n[0]
    u = np.broadcast_to(x[:, None, None], n)
    ax, _, _ = O.apply(P, O.ghosted(u))
    b = O.bvec(P)
    assert np.abs(ax - b).max() <= 1e-12 * np.abs(b).max()
    # and the V-cycle solve converges to it
    sol, hist = O.solve(P, P.zeros(), 30, 1e-13)
    assert np.abs(sol[1:-1, 1:-1, 1:-1] - u).max() < 1e-10


def test_diffusion_operator_is_symmetric_with_zero_interior_row_sums():
    n = (6, 5, 4)
    D = 0.5 + rng.random(n)
    codes = ("DIRC", "FLUX", "ZGRD", "DIRC", "FLUX", "ZGRD")
    P = O.Problem(n, (1.0, 0.7, 0.4), (1, 1, 1), codes, (0.0,) * 6, D)
    A = dense_from_columns(P)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    rows = A.sum(axis=1).reshape(n, order="F")
    # only rows touching a DIRC face carry a non-zero sum (2 D / h^2)
    assert np.abs(rows[1:, 1:-1, :]).max() <= 1e-9
    assert rows[0].min() > 0 and rows[:, -1, :].min() > 0


def test_restriction_invariants():
    f = rng.standard_normal((8, 6, 4))
    c = O.restrict(f)
    assert np.allclose(O.restrict(np.full((8, 6, 4), 3.25)), 3.25, rtol=0, atol=0)
    assert np.isclose(8.0 * c.sum(), f.sum(), rtol=1e-13)
    # child (2,3,1) (1-based) belongs to coarse (1,2,1) only
    g = np.zeros((8, 6, 4))
    g[1, 2, 0] = 8.0
    cg = O.restrict(g)
    assert cg[0, 1, 0] == 1.0 and np.count_nonzero(cg) == 1


def test_prolongation_reproduces_linear_fields():
    """prolongation1 with linear ghost values reproduces a linear field at the
    fine centres (catches a swapped child or slope sign)."""
    cn = (3, 4, 2)
    a, bx, by, bz = 0.7, 1.3, -0.4, 2.1

    def f(x, y, z):
        return a + bx * x + by * y + bz * z

    H = 1.0
    ci = [(np.arange(-1, m + 1) + 0.5) * H for m in cn]  # coarse centres incl. ghosts
    E = f(ci[0][:, None, None], ci[1][None, :, None], ci[2][None, None, :])
    fine = O.prolong1(np.asfortranarray(E))
    fi = [(np.arange(2 * m) + 0.5) * H / 2 for m in cn]
    want = f(fi[0][:, None, None], fi[1][None, :, None], fi[2][None, None, :])
    assert np.allclose(fine, want, rtol=0, atol=1e-13)


def test_rms_definition():
    r = rng.standard_normal((4, 3, 5))
    assert np.isclose(O.rms(r), np.sqrt((r ** 2).sum() / r.size), rtol=1e-15)


def test_gmres_full_krylov_space_solves_the_dense_system():
    n = (4, 3, 2)
    D = 1.0 + rng.random(n)
    v = [3.0 * rng.standard_normal(n) for _ in range(3)]
    codes = ("DIRC", "ZGRD", "DIRC", "ZGRD", "FLUX", "DIRC")
    P = O.Problem(n, (1.0, 1.0, 1.0), (1, 1, 1), codes, (0.0,) * 6, D, *v, None, None, True, 0.05)
    A = dense_from_columns(P)
    b = rng.standard_normal(A.shape[0])
    x = O.gmres(P, b.reshape(n, order="F"), P.zeros(), maxkspdim=A.shape[0], nrestarts=2, tol=1e-15)
    want = np.linalg.solve(A, b)
    got = x[1:-1, 1:-1, 1:-1].ravel(order="F")
    assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want)


@pytest.mark.parametrize("mutation", ["flip_velocity_rule", "drop_dirc_term"])
def test_brute_force_assembly_detects_planted_mistakes(mutation):
    """The face-rule assembly is sharp enough to catch the slips it is meant
    for: a mistake planted in the reference rules (not in the oracle) makes
    the comparison fail."""
    n = (4, 3, 3)
    h = (1 / 4, 1 / 3, 1 / 3)
    D = 1.0 + rng.random(n)
    v = [rng.standard_normal(n) for _ in range(3)]
    codes = ("DIRC", "ZGRD", "ZGRD", "DIRC", "DIRC", "ZGRD")
    P = O.Problem(n, (1.0, 1.0, 1.0), (1, 1, 1), codes, (0.0,) * 6, D, *v)
    A = dense_from_columns(P)
    if mutation == "flip_velocity_rule":
        B = dense_from_faces(n, h, D, [-w for w in v], np.zeros(n), codes, 0.0)
    else:
        B = dense_from_faces(n, h, D, v, np.zeros(n), ("FLUX",) + codes[1:], 0.0)
    assert not np.allclose(A, B, rtol=1e-10, atol=1e-10)


def _efield_problem(n, length, codes, faces=(0.0,) * 6):
    return O.Problem(n, length, (1, 1, 1), codes, faces)


def test_efield_of_linear_potentials_is_exact():
    """update_efield (plasma_pde/poissoneq.f90:565-680) on phi = a + bx + cy
    + dz: with DIRC faces carrying phi's wall values the ghost 2g - phi_1 is
    the linear extrapolation, so E = -scale (b, c, d) in every cell; with
    zero-gradient faces the boundary cells see half the slope."""
    n, L = (7, 5, 4), (0.7, 0.4, 0.6)
    h = [Li / ni for Li, ni in zip(L, n)]
    a, b, c, d = 0.3, 2.0, -1.5, 0.75
    x, y, z = [(np.arange(ni) + 0.5) * hi for ni, hi in zip(n, h)]
    phi = a + b * x[:, None, None] + c * y[None, :, None] + d * z[None, None, :]

    def wall(xv, yv, zv):
        return a + b * xv + c * yv + d * zv

    faces = [wall(0.0, y[:, None], z[None, :]), wall(L[0], y[:, None], z[None, :]),
             wall(x[:, None], 0.0, z[None, :]), wall(x[:, None], L[1], z[None, :]),
             wall(x[:, None], y[None, :], 0.0), wall(x[:, None], y[None, :], L[2])]
    scale = 1.7
    P = _efield_problem(n, L, ("DIRC",) * 6, faces)
    _, E = O.efield(P, O.ghosted(phi), scale)
    for q, slope in enumerate((b, c, d)):
        assert np.allclose(E[1:-1, 1:-1, 1:-1, q], -scale * slope, rtol=1e-12, atol=0)
    P = _efield_problem(n, L, ("ZGRD",) * 6)
    _, E = O.efield(P, O.ghosted(phi), scale)
    Ex = E[1:-1, 1:-1, 1:-1, 0]
    assert np.allclose(Ex[1:-1], -scale * b, rtol=1e-12)
    assert np.allclose(Ex[0], -scale * b / 2, rtol=1e-12) and np.allclose(Ex[-1], -scale * b / 2, rtol=1e-12)
    # ghost faces of E copy the first interior cell
    assert np.array_equal(E[0, 1:-1, 1:-1, 0], E[1, 1:-1, 1:-1, 0])
    assert np.array_equal(E[1:-1, -1, 1:-1, 1], E[1:-1, -2, 1:-1, 1])
    assert np.array_equal(E[1:-1, 1:-1, 0, 2], E[1:-1, 1:-1, 1, 2])


def test_efield_central_difference_is_exact_for_quadratics_and_limited_at_kinks():
    n, L = (9, 3, 3), (0.9, 0.3, 0.3)
    h = L[0] / n[0]
    x = (np.arange(n[0]) + 0.5) * h
    P = _efield_problem(n, L, ("ZGRD",) * 6)
    phi = np.broadcast_to((x ** 2)[:, None, None], n).copy()
    _, E = O.efield(P, O.ghosted(phi), 1.0)
    assert np.allclose(E[2:-2, 1, 1, 0], -2.0 * x[1:-1], rtol=1e-12)
    # minmod: r = (x_i^2 - x_{i-1}^2) / (x_{i+1}^2 - x_i^2) = (2x_i - h) / (2x_i + h)
    _, El = O.efield(P, O.ghosted(phi), 1.0, limiter=True)
    r = (2 * x[1:-1] - h) / (2 * x[1:-1] + h)
    assert np.allclose(El[2:-2, 1, 1, 0], -2.0 * x[1:-1] * r, rtol=1e-12)
    # an asymmetric kink at cell 4: slopes -1 and 2, central difference 0.5,
    # ratio of the one-sided differences -1/2 < 0, so the limiter zeroes E
    x0 = x[4]
    phi = np.broadcast_to(np.maximum(-(x - x0), 2 * (x - x0))[:, None, None], n).copy()
    _, E = O.efield(P, O.ghosted(phi), 1.0)
    _, El = O.efield(P, O.ghosted(phi), 1.0, limiter=True)
    assert np.isclose(E[5, 1, 1, 0], -0.5, rtol=1e-12) and El[5, 1, 1, 0] == 0.0
