"""Transport assembly on the device (mps_species_transport,
mps_eenergy_transport, mps_poisson_transport, mps_electron_temp: the
*_update_transport routines of plasma_pde/{edeneq,ionseq,eenergyeq,
poissoneq}.f90 and update_electrontemp) against the oracle's restatement, on
random fields with random ghost values and ragged grid sizes, from host
arrays and from device tensors.

Bar: bit for bit wherever no exp/pow enters (velocities, diffusion,
reaction, the Poisson source, the electron temperature, Joule heating and
the elastic loss, and the production of a table whose rates are T^0 e^0);
with Arrhenius rates exp and pow come from the device math library on one
side and the C library on the other, and the sources agree to a relative
1e-14 (a few ulps), written in each test.
"""
import numpy as np
import pytest

from mps3d_amd import FieldSolver
from mps3d_amd.chemistry import Chemistry, etempmin_default, ion_elec
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SCALING = (0.7, 1.3, 2.9, 0.25)
RTOL_EXP = 1e-14
RTOL_CHEM = 1e-13


def case(n, seed):
    rng = np.random.default_rng(seed)
    g = tuple(v + 2 for v in n)
    ef = np.asfortranarray(rng.uniform(-3e3, 3e3, g + (3,)))
    te = np.asfortranarray(rng.uniform(5e3, 6e4, g))
    nd = np.asfortranarray(np.stack([rng.uniform(1e21, 3e22, g), rng.uniform(1e14, 1e16, g),
                                     rng.uniform(1e14, 1e16, g)], axis=-1))
    soln = np.asfortranarray(rng.uniform(-1, 1, g))
    return ef, te, nd, soln


def flat_rates(ch):
    """The same chemistry with alpha = Ea = 0: every rate is A exactly."""
    d = dict(ch.__dict__)
    k = ch.k_arrh.copy()
    k[:, 1:] = 0.0
    d["k_arrh"] = k
    return Chemistry(**d)


def no_reaction_energy(ch):
    d = dict(ch.__dict__)
    d["elecenergy"] = np.zeros_like(ch.elecenergy)
    return Chemistry(**d)


def to_dev(a, torch):
    """A Fortran-order (nx+2, ny+2, nz+2[, m]) array as a device tensor with
    the same memory layout."""
    return torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda")


def from_dev(t, shape):
    return t.cpu().numpy().reshape(shape[::-1]).T


N_RAGGED = [(13, 9, 7), (31, 5, 3)]


@pytest.mark.parametrize("n", N_RAGGED)
@pytest.mark.parametrize("specnum", [1, 2, 3])
def test_species_transport_matches_oracle(gpu, n, specnum):
    ef, te, nd, soln = case(n, 11 + specnum)
    P = O.Problem(n, (0.013, 0.009, 0.007))
    with FieldSolver(n, (0.013, 0.009, 0.007)) as s:
        for ch, exact in ((flat_rates(ion_elec()), True), (ion_elec(), False)):
            s.set_chemistry(ch)
            ref = O.species_transport(P, ch, specnum, ef, te, nd, SCALING, soln=soln, dt=0.3)
            got = s.species_transport(specnum, ef, te, nd, SCALING, soln=soln, dt=0.3)
            for k in ("vel", "dcoeff", "reac"):
                assert np.array_equal(got[k], ref[k]), k
            for k in ("source", "prod"):
                if exact:
                    assert np.array_equal(got[k], ref[k]), k
                else:
                    np.testing.assert_allclose(got[k], ref[k], rtol=RTOL_EXP, atol=0)


def test_species_transport_device_tensors(gpu):
    import torch
    n = (13, 9, 7)
    g = tuple(v + 2 for v in n)
    ef, te, nd, soln = case(n, 5)
    ch = flat_rates(ion_elec())
    P = O.Problem(n, (1.0, 1.0, 1.0))
    ref = O.species_transport(P, ch, 2, ef, te, nd, SCALING, soln=soln, dt=0.3)
    with FieldSolver(n) as s:
        s.set_chemistry(ch)
        got = s.species_transport(2, to_dev(ef, torch), to_dev(te, torch), to_dev(nd, torch), SCALING,
                                  soln=to_dev(soln, torch), dt=0.3)
        assert np.array_equal(from_dev(got["vel"], g + (3,)), ref["vel"])
        for k in ("dcoeff", "reac", "source"):
            assert np.array_equal(from_dev(got[k], g), ref[k]), k
        # prod: interior only; the fresh output tensor is zero elsewhere
        assert np.array_equal(from_dev(got["prod"], g), ref["prod"])


@pytest.mark.parametrize("n", N_RAGGED)
def test_eenergy_transport_matches_oracle(gpu, n):
    """Joule heating (upwinded fluxes per axis, ghost densities and fields
    included), elastic loss and vel/dcoeff bit for bit with the reaction
    energy removed; with it, the inelastic term's exp within RTOL_EXP of the
    largest source term."""
    ef, te, nd, soln = case(n, 7)
    L = (0.013, 0.009, 0.007)
    P = O.Problem(n, L)
    lscale = 0.02
    with FieldSolver(n, L) as s:
        for ch, exact in ((no_reaction_energy(ion_elec()), True), (ion_elec(), False)):
            s.set_chemistry(ch)
            ref = O.eenergy_transport(P, ch, ef, te, nd, SCALING, lscale=lscale, soln=soln, dt=0.3, Tg=310.0)
            got = s.eenergy_transport(ef, te, nd, SCALING, lscale=lscale, soln=soln, dt=0.3, Tg=310.0)
            for k in ("vel", "dcoeff", "reac"):
                assert np.array_equal(got[k], ref[k]), k
            if exact:
                assert np.array_equal(got["source"], ref["source"])
                assert np.count_nonzero(ref["source"][1:-1, 1:-1, 1:-1]) > 0
            else:
                scale = np.abs(ref["source"]).max()
                np.testing.assert_allclose(got["source"], ref["source"], rtol=0, atol=RTOL_EXP * scale)


@pytest.mark.parametrize("device", [False, True])
def test_poisson_transport_matches_oracle(gpu, device):
    import torch
    n = (13, 9, 7)
    g = tuple(v + 2 for v in n)
    _, _, nd, soln = case(n, 9)
    ne = np.asfortranarray(nd[..., 1])
    ni = np.asfortranarray(nd[..., 2])
    ch = ion_elec()
    P = O.Problem(n)
    ref = O.poisson_transport(P, ch, ne, ni, SCALING, soln=soln, dt=0.3)
    with FieldSolver(n) as s:
        s.set_chemistry(ch)
        if device:
            got = s.poisson_transport(to_dev(ne, torch), to_dev(ni, torch), SCALING, soln=to_dev(soln, torch), dt=0.3)
            got = {k: from_dev(v, g + ((3,) if k == "vel" else ())) for k, v in got.items()}
        else:
            got = s.poisson_transport(ne, ni, SCALING, soln=soln, dt=0.3)
        for k in ("vel", "dcoeff", "reac", "source"):
            assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("device", [False, True])
def test_electron_temp_matches_oracle(gpu, device):
    import torch
    n = (13, 9, 7)
    g = tuple(v + 2 for v in n)
    rng = np.random.default_rng(13)
    ne = np.asfortranarray(rng.uniform(1e14, 1e16, g))
    escale = 1e15 * 1.380650e-23 * 11604.5
    tmin = etempmin_default()
    # a third of the cells below etempmin
    Te = np.asfortranarray(rng.uniform(0.3 * tmin, 3.0 * tmin, g))
    eps = np.asfortranarray(1.5 * ne * 1.380650e-23 * Te / escale)
    te0 = np.asfortranarray(rng.uniform(-1, 1, g))
    P = O.Problem(n)
    eps_ref, te_ref = O.electron_temp(P, eps, ne, escale, tmin, te0)
    assert 0 < np.count_nonzero(te_ref[1:-1, 1:-1, 1:-1] == tmin) < np.prod(n)
    with FieldSolver(n) as s:
        if device:
            e_d, t_d = to_dev(eps, torch), to_dev(te0, torch)
            s.electron_temp(e_d, to_dev(ne, torch), escale, tmin, t_d)
            eps_got, te_got = from_dev(e_d, g), from_dev(t_d, g)
        else:
            eps_got, te_got = eps.copy(order="F"), te0.copy(order="F")
            s.electron_temp(eps_got, ne, escale, tmin, te_got)
    assert np.array_equal(eps_got, eps_ref)
    assert np.array_equal(te_got, te_ref)


def test_transport_argument_errors(gpu):
    from mps3d_amd import _lib as M
    n = (5, 4, 3)
    ef, te, nd, soln = case(n, 1)
    with FieldSolver(n) as s:
        with pytest.raises(M.MpsError):
            s.species_transport(2, ef, te, nd, SCALING, soln=soln)      # no chemistry yet
        st = s.lib.mps_species_transport(s._h, 2, 1.0, 0, None, None, None, 300.0, 0.0, None, None, None, None,
                                         None, None, None)
        assert st == M.ERR_STATE
        s.set_chemistry(ion_elec())
        with pytest.raises(M.MpsError):
            s.species_transport(4, ef, te, nd, SCALING, soln=soln)      # specnum out of range
        with pytest.raises(M.MpsError):
            s.species_transport(2, ef, te, nd, SCALING, soln=soln, dt=0.0)  # time term with dt = 0


def test_transport_and_field_at_257_cubed(gpu):
    """The full 257^3 grid on device tensors, through properties that do not
    depend on size: a linear potential between DIRC plates gives the exact
    uniform field (ghost faces included); uniform densities and temperature
    in that field give uniform drift velocities mu E, the Arrhenius
    production k n_N n_E at every interior cell, scaled zeros plus the time
    term in the ghost cells, and the Poisson source of the space charge."""
    import math

    import torch
    n = 257
    L = 0.02
    with FieldSolver(n, L) as s:
        ch = ion_elec()
        s.set_chemistry(ch)
        V0 = 1000.0
        # phi = V0 z / L at cell centres; DIRC walls 0 and V0
        h = L / n
        z = (torch.arange(n + 2, dtype=torch.float64, device="cuda") - 0.5) * h
        phi = (V0 * z / L).view(-1, 1, 1).expand(n + 2, n + 2, n + 2).contiguous()
        nx = n
        faces = [0.0, 0.0, 0.0, 0.0, np.zeros((nx, nx)), np.full((nx, nx), V0)]
        s.set_bc(("ZGRD", "ZGRD", "ZGRD", "ZGRD", "DIRC", "DIRC"), faces)
        E = torch.zeros((3, n + 2, n + 2, n + 2), dtype=torch.float64, device="cuda")
        s.efield(phi, 1.0, False, out=E)
        Ez = E[2, 1:-1, 1:-1, 1:-1]
        assert torch.allclose(Ez, torch.full_like(Ez, -V0 / L), rtol=1e-12, atol=0)
        assert float(E[0].abs().max()) == 0.0 and float(E[1].abs().max()) == 0.0
        g = (n + 2) ** 3
        Ea = ch.k_arrh[0, 2]
        te = torch.full((g,), Ea, dtype=torch.float64, device="cuda")
        nN, nE, nI = 3.0e22, 2.0e16, 1.0e16
        nd = torch.cat([torch.full((g,), v, dtype=torch.float64, device="cuda") for v in (nN, nE, nI)])
        soln = torch.full((g,), 0.5, dtype=torch.float64, device="cuda")
        o = s.species_transport(2, E.reshape(-1), te, nd, SCALING, soln=soln, dt=0.25)
        vel = o["vel"].view(3, n + 2, n + 2, n + 2)
        vz = vel[2, 1:-1, 1:-1, 1:-1]
        assert torch.equal(vz, ch.mobility[1] * Ez * SCALING[0])
        src = o["source"].view(n + 2, n + 2, n + 2)
        prod = (ch.k_arrh[0, 0] / math.e) * nN * nE
        inner = src[1:-1, 1:-1, 1:-1]
        assert torch.allclose(inner, torch.full_like(inner, prod * SCALING[3] + 0.5 / 0.25), rtol=1e-14, atol=0)
        assert float(src[0].sub(2.0).abs().max()) == 0.0 and float(src[:, 0].sub(2.0).abs().max()) == 0.0
        p = s.poisson_transport(nd[g:2 * g], nd[2 * g:], SCALING, soln=soln, dt=0.25)
        ps = p["source"].view(n + 2, n + 2, n + 2)[1:-1, 1:-1, 1:-1]
        rho = (nI - nE) * 1.602176e-19 / 8.854188e-12 * SCALING[3] + 0.5 / 0.25
        assert torch.allclose(ps, torch.full_like(ps, rho), rtol=1e-14, atol=0)


@pytest.mark.parametrize("name", ["H2_O2", "ion_elec", "O2", "H2"])
def test_reference_chemistry_on_the_device(gpu, name):
    """The reference's chem_modules (tests/golden/chem_reference.npz: what
    H2_O2_chem.f90, ion_elec_chem.f90, O2_chem.f90 (custom rate fits) and
    H2_chem.f90 return on 64 states) against the
    device transport on a grid whose cells carry those states: mobility
    (vel / E), diffusion and production of every species, and the electron
    energy source at E = 0 (minus the elastic loss plus the inelastic term);
    and the device against the oracle on the same grid with a random field.
    Tolerance 1e-13 relative: the exp and log of the device math library
    differ from the host's by an ulp, and the production sums of O2_chem
    cancel by up to a factor of about 170 (oracle/tests use the same bar)."""
    from conftest import golden
    from mps3d_amd.chemistry import ECHARGE, K_BMAN, h2, h2_o2, o2
    d = golden("chem_reference")
    ch = {"H2_O2": h2_o2, "ion_elec": ion_elec, "O2": o2, "H2": h2}[name]()
    grid = (16, 2, 2)
    g = tuple(v + 2 for v in grid)
    Tg, _ = d[name + "_TgPg"]
    te = np.full(g, 1e4, order="F")
    nd = np.full(g + (ch.nspecies,), 1e15, order="F")
    te[1:-1, 1:-1, 1:-1] = d[name + "_Te"].reshape(grid, order="F")
    for s in range(ch.nspecies):
        nd[1:-1, 1:-1, 1:-1, s] = d[name + "_sd"][:, s].reshape(grid, order="F")

    def inner(a):
        return a[1:-1, 1:-1, 1:-1].reshape(-1, order="F")

    ef = np.zeros(g + (3,), order="F")
    ef[..., 0] = 1.0
    rng = np.random.default_rng(17)
    efr = np.asfortranarray(rng.uniform(-1e4, 1e4, g + (3,)))
    P = O.Problem(grid)
    with FieldSolver(grid) as s:
        s.set_chemistry(ch)
        for sp in range(ch.nspecies):
            o = s.species_transport(sp + 1, ef, te, nd, (1, 1, 1, 1), timederiv=False, Tg=Tg)
            np.testing.assert_allclose(inner(o["prod"]), d[name + "_prod"][:, sp], rtol=RTOL_CHEM, atol=0)
            if ch.charge[sp] != 0:
                np.testing.assert_allclose(inner(o["vel"][..., 0]), d[name + "_mob"][:, sp], rtol=RTOL_CHEM, atol=0)
                np.testing.assert_allclose(inner(o["dcoeff"]), d[name + "_dco"][:, sp], rtol=RTOL_CHEM, atol=0)
            got = s.species_transport(sp + 1, efr, te, nd, SCALING, soln=te, dt=0.5, Tg=Tg)
            ref = O.species_transport(P, ch, sp + 1, efr, te, nd, SCALING, soln=te, dt=0.5, Tg=Tg)
            for k in ("vel", "dcoeff", "source", "prod"):
                np.testing.assert_allclose(got[k], ref[k], rtol=RTOL_CHEM, atol=0, err_msg=k)
        e = ch.especnum - 1
        o = s.eenergy_transport(np.zeros(g + (3,), order="F"), te, nd, (1, 1, 1, 1), timederiv=False, Tg=Tg)
        me, mbg = ch.molmass[e], ch.molmass[ch.bgspecnum - 1]
        elastic = 3.0 * K_BMAN * d[name + "_sd"][:, e] * (d[name + "_Te"] - Tg) * (me / mbg) * d[name + "_collfreq"]
        np.testing.assert_allclose(inner(o["source"]), -elastic + d[name + "_inel"] * ECHARGE, rtol=1e-12, atol=0)
        got = s.eenergy_transport(efr, te, nd, SCALING, lscale=0.01, soln=te, dt=0.5, Tg=Tg)
        ref = O.eenergy_transport(P, ch, efr, te, nd, SCALING, lscale=0.01, soln=te, dt=0.5, Tg=Tg)
        for k in ("vel", "dcoeff"):
            np.testing.assert_allclose(got[k], ref[k], rtol=RTOL_CHEM, atol=0, err_msg=k)
        scale = np.abs(ref["source"]).max()
        np.testing.assert_allclose(got["source"], ref["source"], rtol=0, atol=RTOL_CHEM * scale)


def test_field_dependent_rate_on_the_device(gpu):
    """Rate kind 2, alpha |E| exp(-E0/|E|) (getcustomrate of
    H2_chem.f90:423-437), in the device's species and energy transport
    against the oracle on a random field of both signs."""
    ch = ion_elec()
    ch.rate_kind = np.array([2], dtype=np.int32)
    ch.rate_coef = np.zeros((5, 1))
    ch.rate_coef[:2, 0] = 3.0e-21, 2.5e4
    grid = (12, 10, 9)
    g = tuple(v + 2 for v in grid)
    rng = np.random.default_rng(5)
    ef = np.asfortranarray(rng.uniform(-4e4, 4e4, g + (3,)))
    te = np.asfortranarray(rng.uniform(5e3, 5e4, g))
    nd = np.asfortranarray(10.0 ** rng.uniform(14, 18, g + (3,)))
    P = O.Problem(grid)
    with FieldSolver(grid) as s:
        s.set_chemistry(ch)
        for sp in range(ch.nspecies):
            got = s.species_transport(sp + 1, ef, te, nd, SCALING, soln=te, dt=0.5)
            ref = O.species_transport(P, ch, sp + 1, ef, te, nd, SCALING, soln=te, dt=0.5)
            for k in ("prod", "source"):
                np.testing.assert_allclose(got[k], ref[k], rtol=RTOL_EXP, atol=0, err_msg=k)
        got = s.eenergy_transport(ef, te, nd, SCALING, lscale=0.01, soln=te, dt=0.5)
        ref = O.eenergy_transport(P, ch, ef, te, nd, SCALING, lscale=0.01, soln=te, dt=0.5)
        scale = np.abs(ref["source"]).max()
        np.testing.assert_allclose(got["source"], ref["source"], rtol=0, atol=RTOL_EXP * scale)
