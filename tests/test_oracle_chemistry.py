"""The chemistry tables and the oracle's transport models against the
reference's own chem_modules.

tests/golden/chem_reference.npz holds what plasma_pde/chemistry/H2_O2_chem.f90
(the reference's default chemistry, build/makefile:1) and ion_elec_chem.f90,
compiled and driven by oracle/ref_driver/ref_chem.f90, and O2_chem.f90 and
H2_chem.f90, driven by oracle/ref_driver/ref_chem4.f90, return for 64 seeded
states each: their tables (assignreactions, setspecparams) and per state the
rate of every reaction, the production of every species (getspecproduction),
mobilities and diffusion coefficients (getspecmobility, getspecdcoeff), the
inelastic term (getelectroninelasticterm) and the electron collision
frequency (getelectroncollisionfrequency).

The factories in mps3d_amd/chemistry.py must hold the reference's tables bit
for bit; the oracle's species and energy transport, evaluated on a grid whose
cells carry the 64 states, must give the reference's mobility (as vel / E),
diffusion, production and energy source (E = 0: minus the elastic loss plus
the inelastic term) to 1e-13 relative (exp and log of two math libraries).
"""
import numpy as np
import pytest

from conftest import golden
from mps3d_amd.chemistry import ECHARGE, K_BMAN, h2, h2_o2, ion_elec, o2
from oracle import oracle as O

GRID = (16, 2, 2)     # 64 interior cells, one state each
RTOL = 1e-13


def tables_match(ch, d, name):
    assert np.array_equal(np.asarray(ch.reactants), d[name + "_reactants"])
    assert np.array_equal(np.asarray(ch.products), d[name + "_products"])
    assert np.array_equal(ch.k_arrh, d[name + "_k_arrh"])
    assert np.array_equal(ch.elecenergy, d[name + "_elecenergy"])
    assert np.array_equal(ch.molmass, d[name + "_molmass"])
    assert np.array_equal(ch.charge, d[name + "_charge"])
    assert (ch.especnum, ch.bgspecnum, ch.ionspecmin, ch.ionspecmax) == tuple(
        int(d[name + "_" + k]) for k in ("especnum", "bgspecnum", "ionspecmin", "ionspecmax"))
    # a reaction takes getarrhrate exactly where isratearrh is .true.
    kind = np.zeros(ch.nreac, dtype=int) if ch.rate_kind is None else np.asarray(ch.rate_kind)
    assert np.array_equal(kind == 0, d[name + "_isratearrh"] != 0)


FACTORIES = [("H2_O2", h2_o2), ("ion_elec", ion_elec), ("O2", o2), ("H2", h2)]


@pytest.mark.parametrize("name,factory", FACTORIES)
def test_factory_tables_are_the_reference_tables(name, factory):
    tables_match(factory(), golden("chem_reference"), name)


def grid_fields(d, name):
    """Ghosted fields whose interior cells (x fastest) carry the states."""
    ns = int(d[name + "_nspecies"])
    g = tuple(v + 2 for v in GRID)
    te = np.full(g, 1e4, order="F")
    nd = np.full(g + (ns,), 1e15, order="F")
    te[1:-1, 1:-1, 1:-1] = d[name + "_Te"].reshape(GRID, order="F")
    for s in range(ns):
        nd[1:-1, 1:-1, 1:-1, s] = d[name + "_sd"][:, s].reshape(GRID, order="F")
    return te, nd


def interior(a):
    return a[1:-1, 1:-1, 1:-1].reshape(-1, order="F")


@pytest.mark.parametrize("name,factory", FACTORIES)
def test_oracle_transport_follows_the_reference_chemistry(name, factory):
    d = golden("chem_reference")
    ch = factory()
    Tg, Pg = d[name + "_TgPg"]
    te, nd = grid_fields(d, name)
    P = O.Problem(GRID)
    g = tuple(v + 2 for v in GRID)
    ef = np.zeros(g + (3,), order="F")
    ef[..., 0] = 1.0                     # vel_x = mu
    charged = np.nonzero(ch.charge)[0]
    for s in range(ch.nspecies):
        o = O.species_transport(P, ch, s + 1, ef, te, nd, (1, 1, 1, 1), timederiv=False, Tg=Tg)
        np.testing.assert_allclose(interior(o["prod"]), d[name + "_prod"][:, s], rtol=RTOL, atol=0, err_msg=str(s))
        if s in charged:
            np.testing.assert_allclose(interior(o["vel"][..., 0]), d[name + "_mob"][:, s], rtol=RTOL, atol=0)
        if s in charged or d[name + "_dco"][:, s].any():
            np.testing.assert_allclose(interior(o["dcoeff"]), d[name + "_dco"][:, s], rtol=RTOL, atol=0)
    # electron energy with E = 0: the source is -elastic + inelastic, from
    # the reference's own collision frequency and inelastic term
    e = ch.especnum - 1
    o = O.eenergy_transport(P, ch, np.zeros(g + (3,), order="F"), te, nd, (1, 1, 1, 1), timederiv=False, Tg=Tg)
    me, mbg = ch.molmass[e], ch.molmass[ch.bgspecnum - 1]
    ne = d[name + "_sd"][:, e]
    elastic = 3.0 * K_BMAN * ne * (d[name + "_Te"] - Tg) * (me / mbg) * d[name + "_collfreq"]
    expect = -elastic + d[name + "_inel"] * ECHARGE
    np.testing.assert_allclose(interior(o["source"]), expect, rtol=1e-12, atol=0)
    np.testing.assert_allclose(interior(o["dcoeff"]), 5.0 / 3.0 * d[name + "_dco"][:, e], rtol=RTOL, atol=0)


def test_field_dependent_rate_follows_its_closed_form():
    """Rate kind 2 (getcustomrate of H2_chem.f90:423-437, which that module's
    tables never select): k = alpha |E| exp(-E0/|E|) at the cell's field
    magnitude sqrt(Ex^2 + Ey^2 + Ez^2). On ion_elec's one reaction
    N + E -> I + 2E the electron production is k n_N n_E; the field's sign
    must not matter and a dropped component would change |E|."""
    ch = ion_elec()
    alpha, E0 = 3.0e-21, 2.5e4
    ch.rate_kind = np.array([2], dtype=np.int32)
    ch.rate_coef = np.zeros((5, 1))
    ch.rate_coef[:2, 0] = alpha, E0
    g = tuple(v + 2 for v in GRID)
    rng = np.random.default_rng(3)
    ef = np.asfortranarray(rng.uniform(-4e4, 4e4, g + (3,)))
    te = np.full(g, 2e4, order="F")
    nd = np.asfortranarray(10.0 ** rng.uniform(14, 18, g + (3,)))
    o = O.species_transport(O.Problem(GRID), ch, 2, ef, te, nd, (1, 1, 1, 1), timederiv=False)
    mag = np.sqrt(ef[..., 0] ** 2 + ef[..., 1] ** 2 + ef[..., 2] ** 2)
    k = alpha * mag * np.exp(-E0 / mag)
    expect = k * nd[..., 0] * nd[..., 1]
    np.testing.assert_allclose(interior(o["prod"]), interior(expect), rtol=1e-14, atol=0)
