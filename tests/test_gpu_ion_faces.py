"""mps_ion_drift_faces: the FLUX faces of ions_update_bcs
(plasma_pde/ionseq.f90:247-476) in one launch, against the face values the
plasma loop formed before it with one torch operation per step of the
formula: on the boundary cells of each face, mu = getspecmobility (the
constant, or N_atm mu0 / sum_j n_j / K_j in species order), drift = mu n
E_normal vscale with n the ion's scaled solution, kept where mu E_normal
points out of the domain (-drift on the low faces, +drift on the high
ones), else 0. Each operation is one correctly rounded float64 step, so the
kernel must agree bit for bit. The divisions are tensor by tensor here:
torch divides a tensor by a Python number (and a number by a tensor)
through the reciprocal, two roundings where the reference's n_j / K_j and
mu0 N_atm / sum are one; the loop's former torch faces differed from the
reference in the last bits for that reason. Grids are not cubic, so a swapped axis or
face layout fails; E takes both signs, so both branches are taken.
"""
import numpy as np
import pytest
import torch

from mps3d_amd import FieldSolver
from mps3d_amd.chemistry import K_BMAN, h2_o2, ion_elec

pytestmark = pytest.mark.gpu

FACES = ("left", "right", "bottom", "top", "back", "front")
N = (12, 10, 7)


def boundary(G, f, n):
    nx, ny, nz = n
    sl = {"left": (slice(1, -1), slice(1, -1), 1), "right": (slice(1, -1), slice(1, -1), nx),
          "bottom": (slice(1, -1), 1, slice(1, -1)), "top": (slice(1, -1), ny, slice(1, -1)),
          "back": (1, slice(1, -1), slice(1, -1)), "front": (nz, slice(1, -1), slice(1, -1))}[f]
    return G[sl]


def torch_faces(ch, sp, E, numden, ni, vscale, n):
    """The face values as the torch formula gives them."""
    out = []
    for k, f in enumerate(FACES):
        kind = 0 if ch.mob_kind is None else int(ch.mob_kind[sp])
        if kind == 0:
            mu = float(ch.mobility[sp])
        else:
            total = None
            for j in range(ch.nspecies):
                if ch.mob_terms[j, sp]:
                    nj = boundary(numden[j], f, n)
                    inv = nj / torch.full_like(nj, float(ch.mob_coef[0, j, sp]))
                    total = inv if total is None else total + inv
            mu = torch.full_like(total, float(ch.mobility[sp]) * (101325.0 / K_BMAN / 300.0)) / total
        En = boundary(E[k // 2], f, n)
        nb = boundary(ni, f, n)
        low = k % 2 == 0
        drift = mu * nb * En * vscale
        leaves = (mu * En < 0) if low else (mu * En > 0)
        val = -drift if low else drift
        out.append(torch.where(leaves, val, torch.zeros_like(val)).contiguous())
    return out


def fields(ch, seed):
    nx, ny, nz = N
    g = (nz + 2, ny + 2, nx + 2)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    E = (torch.rand((3,) + g, dtype=torch.float64, device="cuda", generator=gen) - 0.5) * 2e4
    numden = 10.0 ** (torch.rand((ch.nspecies,) + g, dtype=torch.float64, device="cuda", generator=gen) * 8 + 14)
    ni = torch.rand(g, dtype=torch.float64, device="cuda", generator=gen) * 1e3
    return E, numden, ni


@pytest.mark.parametrize("factory", [h2_o2, ion_elec])
def test_faces_match_the_torch_formula_bit_for_bit(factory):
    ch = factory()
    ions = range(ch.ionspecmin - 1, ch.ionspecmax)
    assert len(ions) > 0
    with FieldSolver(N, (0.012, 0.010, 0.007)) as s:
        s.set_chemistry(ch)
        for q, sp in enumerate(ions):
            E, numden, ni = fields(ch, 100 + q)
            vscale = 0.37 + q
            expect = torch_faces(ch, sp, E, numden, ni, vscale, N)
            got = [torch.full_like(e, float("nan")) for e in expect]
            s.ion_drift_faces(sp + 1, E.reshape(-1), numden.reshape(-1), ni.reshape(-1), vscale, got)
            torch.cuda.synchronize()
            for k in range(6):
                assert torch.equal(got[k], expect[k]), (ch.names[sp], FACES[k])
                assert (got[k] != 0).any() and (got[k] == 0).any(), (ch.names[sp], FACES[k])


def test_null_faces_are_skipped_and_host_arrays_work():
    ch = h2_o2()
    sp = ch.ionspecmin - 1
    with FieldSolver(N, (0.012, 0.010, 0.007)) as s:
        s.set_chemistry(ch)
        E, numden, ni = fields(ch, 7)
        expect = torch_faces(ch, sp, E, numden, ni, 1.0, N)
        got = [None if k in (1, 4) else torch.full_like(expect[k], -1.0) for k in range(6)]
        s.ion_drift_faces(sp + 1, E.reshape(-1), numden.reshape(-1), ni.reshape(-1), 1.0, got)
        torch.cuda.synchronize()
        for k in (0, 2, 3, 5):
            assert torch.equal(got[k], expect[k])
        # host inputs and outputs (Fortran-order face arrays), staged by the library
        nx, ny, nz = N
        g = (nx + 2, ny + 2, nz + 2)
        Eh = np.asfortranarray(np.stack([c.cpu().numpy().T for c in E], axis=-1))
        nh = np.asfortranarray(np.stack([c.cpu().numpy().T for c in numden], axis=-1))
        nih = np.asfortranarray(ni.cpu().numpy().T)
        assert Eh.shape == g + (3,)
        host = [np.zeros(e.shape[::-1], order="F") for e in expect]
        s.ion_drift_faces(sp + 1, Eh, nh, nih, 1.0, host)
        for k in range(6):
            assert np.array_equal(host[k], expect[k].cpu().numpy().T)


def test_rejects_electron_fit_mobility_and_bad_faces():
    from mps3d_amd import _lib as M
    ch = h2_o2()
    with FieldSolver(N) as s:
        s.set_chemistry(ch)
        E, numden, ni = fields(ch, 1)
        faces = [torch.zeros(ny * nz, dtype=torch.float64, device="cuda") for ny, nz in
                 ((10, 7), (10, 7), (12, 7), (12, 7), (12, 10), (12, 10))]
        with pytest.raises(M.MpsError):   # the electron's mobility is the log-quadratic fit
            s.ion_drift_faces(ch.especnum, E.reshape(-1), numden.reshape(-1), ni.reshape(-1), 1.0, faces)
        faces[2] = torch.zeros(5, dtype=torch.float64, device="cuda")
        with pytest.raises(ValueError):
            s.ion_drift_faces(ch.ionspecmin, E.reshape(-1), numden.reshape(-1), ni.reshape(-1), 1.0, faces)
