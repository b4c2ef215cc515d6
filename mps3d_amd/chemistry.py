"""Chemistry tables for the device-side transport assembly.

A :class:`Chemistry` holds what a reference ``chem_module`` defines in its
``assignreactions`` and ``setspecparams`` (species numbers 1-based,
``reactants``/``products`` as (nreac, nspecies), ``k_arrh`` as (nreac, 3)
rows of A, alpha, Ea): Arrhenius rates in the electron temperature, and per
species a constant mobility and diffusion coefficient (the form of
``plasma_pde/chemistry/ion_elec_chem.f90``, :func:`ion_elec`) or the
field-dependent models of ``plasma_pde/chemistry/H2_O2_chem.f90``
(:func:`h2_o2`); reaction rates Arrhenius or the custom fits of
``O2_chem.f90`` (:func:`o2`). ``H2_chem.f90`` is :func:`h2`. ``FieldSolver.set_chemistry`` copies it to the device.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from fractions import Fraction

import numpy as np

from . import _lib as M

# plasma_pde/fundconstants.f90:3-9
ECHARGE = 1.602176e-19
EPS0 = 8.854188e-12
M_ELEC = 9.109382e-31
M_PROT = 1.672621e-27
K_BMAN = 1.380650e-23
EV_TO_K = 11604.5


def real4(x):
    """A Fortran default-real literal (``0.01``, ``-15.578``) as the double it
    becomes when assigned to a real*8 variable."""
    return float(np.float32(x))


@dataclass
class Chemistry:
    names: tuple
    reactants: np.ndarray   # (nreac, nspecies) int
    products: np.ndarray    # (nreac, nspecies) int
    k_arrh: np.ndarray      # (nreac, 3): A, alpha, Ea (K)
    elecenergy: np.ndarray  # (nreac,) eV
    mobility: np.ndarray    # (nspecies,)
    dcoeff: np.ndarray      # (nspecies,)
    molmass: np.ndarray     # (nspecies,) kg
    charge: np.ndarray      # (nspecies,)
    especnum: int
    bgspecnum: int
    ionspecmin: int
    ionspecmax: int
    neutralspecmin: int = 0     # solved neutral species (neutralspecmax < min: none)
    neutralspecmax: int = -1
    # field-dependent transport (None: the constants above); see mps_chem
    mob_kind: np.ndarray | None = None     # (nspecies,) 0 constant, 1 electron fit, 2 reduced ion mobility
    mob_terms: np.ndarray | None = None    # (nspecies partner, nspecies)
    mob_coef: np.ndarray | None = None     # (3, nspecies partner, nspecies)
    dcoeff_kind: np.ndarray | None = None  # (nspecies,) 0 constant, 1 k_B T_e mu/(Z e), 2 k_B T_g mu/(Z e)
    # reaction rates (None: every reaction Arrhenius); see mps_chem
    rate_kind: np.ndarray | None = None    # (nreac,) 0 Arrhenius, 1 exponential fit in 1/T, 2 alpha |E| exp(-E0/|E|)
    rate_coef: np.ndarray | None = None    # (5, nreac)

    @property
    def nspecies(self):
        return len(self.names)

    @property
    def nreac(self):
        return int(self.reactants.shape[0])

    def arrays(self):
        """The tables as the C ABI takes them (Fortran order, int32/float64)."""
        out = {
            "reactants": np.asfortranarray(self.reactants, dtype=np.int32),
            "products": np.asfortranarray(self.products, dtype=np.int32),
            "k_arrh": np.ascontiguousarray(self.k_arrh, dtype=np.float64),  # k_arrh(3, nreac) in memory
            "elecenergy": np.ascontiguousarray(self.elecenergy, dtype=np.float64),
            "mobility": np.ascontiguousarray(self.mobility, dtype=np.float64),
            "dcoeff": np.ascontiguousarray(self.dcoeff, dtype=np.float64),
            "molmass": np.ascontiguousarray(self.molmass, dtype=np.float64),
            "charge": np.ascontiguousarray(self.charge, dtype=np.float64),
        }
        if self.mob_kind is not None:
            out["mob_kind"] = np.ascontiguousarray(self.mob_kind, dtype=np.int32)
            out["mob_terms"] = np.asfortranarray(self.mob_terms, dtype=np.int32)
            out["mob_coef"] = np.asfortranarray(self.mob_coef, dtype=np.float64)
            out["dcoeff_kind"] = np.ascontiguousarray(self.dcoeff_kind, dtype=np.int32)
        if self.rate_kind is not None:
            out["rate_kind"] = np.ascontiguousarray(self.rate_kind, dtype=np.int32)
            out["rate_coef"] = np.asfortranarray(self.rate_coef, dtype=np.float64)
        return out

    def struct(self):
        """(mps_chem structure, arrays it points into: keep both alive)."""
        a = self.arrays()
        s = M.Chem()
        s.nspecies, s.nreac = self.nspecies, self.nreac
        s.especnum, s.bgspecnum = int(self.especnum), int(self.bgspecnum)
        s.ionspecmin, s.ionspecmax = int(self.ionspecmin), int(self.ionspecmax)
        for k, v in a.items():
            setattr(s, k, C.c_void_p(v.ctypes.data))
        return s, a


def ion_elec():
    """The tables of plasma_pde/chemistry/ion_elec_chem.f90:30-102: species
    N (background), E, I+; one reaction N + E -> I + 2E with
    k = 2.5e-6 cm^3/s, Ea = 278508 K, electron energy -15.578 eV; electron
    mobility -20 and diffusion 100, ion mobility 0.2 and diffusion 0.01.

    The values are the doubles the reference holds: its default-real literals
    (0.01, 0.2, -15.578 and cm_to_m = 0.01) are single precision before they
    are stored, and 2.5d-6*(cm_to_m**3) is a constant expression the compiler
    folds with correct rounding."""
    cm_to_m = real4(0.01)
    A = 2.5e-6 * float(Fraction(cm_to_m) ** 3)
    N, E, I = 0, 1, 2
    reactants = np.zeros((1, 3), dtype=np.int32)
    products = np.zeros((1, 3), dtype=np.int32)
    reactants[0, N] = 1
    reactants[0, E] = 1
    products[0, I] = 1
    products[0, E] = 2
    return Chemistry(
        names=("N", "E", "I+"),
        reactants=reactants, products=products,
        k_arrh=np.array([[A, 0.0, real4(278508.0)]]),
        elecenergy=np.array([real4(-15.578)]),
        # getspecmobility / getspecdcoeff (ion_elec_chem.f90:201-242); the
        # background gas has none (the reference prints an error for it)
        mobility=np.array([0.0, -20.0, real4(0.2)]),
        dcoeff=np.array([0.0, 100.0, real4(0.01)]),
        molmass=np.array([4.0 * M_PROT, M_ELEC, 4.0 * M_PROT]),
        charge=np.array([0.0, -1.0, 1.0]),
        especnum=2, bgspecnum=1, ionspecmin=3, ionspecmax=3)


def h2_o2():
    """The reference's default chemistry, plasma_pde/chemistry/H2_O2_chem.f90
    (build/makefile:1): H2, O2 (background), E, H+, H2+, O+, O2+, O-, H, O;
    15 Arrhenius reactions in T_e (every reaction of assignreactions has
    isratearrh = .true.), and field-dependent transport.

    The reaction tables are the doubles that assignreactions leaves in
    reactants, products, k_arrh and elecenergy (:30-384) when compiled: its
    constant expressions such as 6.5023d-9*(cm_to_m**3)/(eVtoK**0.48931) are
    folded by the compiler from single-precision literals, so they are given
    here as the values themselves, taken from the reference build
    (oracle/ref_driver/ref_chem.f90); a test checks them against it.
    Transport (getspecmobility, getspecdcoeff, :532-672): the electron
    mobility is the fit -1 / sum_j n_j / exp(a_j L^2 + b_j L + c_j) over H2
    and O2 (L = ln T_e[eV], single-precision coefficients), ion mobilities are
    N_atm / sum_j n_j / (K_j cm^2) (O-: negative), charged species diffuse by
    the Einstein relation (electrons at T_e, ions at T_g), H and O with
    constants."""
    reactants = np.array([[0, 0, 1, 0, 0, 0, 0, 0, 1, 0], [1, 0, 1, 0, 0, 0, 0, 0, 0, 0], [0, 0, 1, 0, 1, 0, 0, 0, 0, 0], [0, 0, 1, 0, 1, 0, 0, 0, 0, 0], [0, 0, 0, 0, 1, 0, 0, 0, 1, 0], [1, 0, 0, 1, 0, 0, 0, 0, 0, 0], [1, 0, 1, 0, 0, 0, 0, 0, 0, 0], [1, 0, 1, 0, 0, 0, 0, 0, 0, 0], [0, 1, 1, 0, 0, 0, 0, 0, 0, 0], [0, 1, 1, 0, 0, 0, 0, 0, 0, 0], [0, 1, 1, 0, 0, 0, 0, 0, 0, 0], [0, 0, 1, 0, 0, 0, 0, 0, 0, 1], [0, 0, 0, 0, 0, 0, 1, 1, 0, 0], [0, 0, 0, 0, 0, 1, 0, 1, 0, 0], [0, 0, 1, 0, 0, 0, 0, 1, 0, 0]], dtype=np.int32)
    products = np.array([[0, 0, 2, 1, 0, 0, 0, 0, 0, 0], [0, 0, 2, 1, 0, 0, 0, 0, 1, 0], [0, 0, 1, 1, 0, 0, 0, 0, 1, 0], [0, 0, 2, 2, 0, 0, 0, 0, 0, 0], [1, 0, 0, 1, 0, 0, 0, 0, 0, 0], [0, 0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 0, 2, 0, 1, 0, 0, 0, 0, 0], [0, 0, 1, 0, 0, 0, 0, 0, 2, 0], [0, 0, 2, 0, 0, 0, 1, 0, 0, 0], [0, 0, 1, 0, 0, 0, 0, 0, 0, 2], [0, 0, 0, 0, 0, 0, 0, 1, 0, 1], [0, 0, 2, 0, 0, 1, 0, 0, 0, 0], [0, 1, 0, 0, 0, 0, 0, 0, 0, 1], [0, 0, 0, 0, 0, 0, 0, 0, 0, 2], [0, 0, 2, 0, 0, 0, 0, 0, 0, 1]], dtype=np.int32)
    k_arrh = np.array([
        [6.671210258802429e-17, 0.48930999636650085, 149624.359375],
        [4.673050547002412e-16, 0.4445599913597107, 437818.75],
        [6.780687508393681e-14, 0.04876000061631203, 112450.8515625],
        [1.122866241117621e-16, 0.31393998861312866, 270371.5],
        [8.999999396502984e-16, 0.0, 0.0],
        [1.1899999202042835e-28, 0.0, 0.0],
        [6.269267675269048e-15, 0.17156000435352325, 232987.484375],
        [1.8635879765620025e-08, -1.236680030822754, 146128.84375],
        [8.354669297014652e-18, 0.5, 146216.703125],
        [4.229999716356402e-15, 0.0, 64985.19921875],
        [8.79999940991403e-17, 0.0, 51059.80078125],
        [1.2852985741823762e-17, 0.699999988079071, 150858.5],
        [1.9899998665601042e-13, 0.0, 0.0],
        [2.6599998216331046e-13, 0.0, 0.0],
        [1.999999865889552e-13, 0.0, 63824.75],
    ])
    elecenergy = np.array([-12.89365005493164, -37.72835922241211, -9.690279960632324, -23.29884910583496, 0.0, 0.0, -20.07733917236328, -12.592430114746094, -12.600000381469727, -6.400000095367432, -3.5999999046325684, -13.600000381469727, 0.0, 0.0, -5.5])
    cm2 = real4(0.01) * real4(0.01)
    ns = 10
    H2, O2, E, Hp, H2p, Op, O2p, Om, H, O = range(ns)
    mob = np.zeros(ns)
    dco = np.zeros(ns)
    mob_kind = np.zeros(ns, dtype=np.int32)
    dk = np.zeros(ns, dtype=np.int32)
    terms = np.zeros((ns, ns), dtype=np.int32)
    coef = np.zeros((3, ns, ns))
    # electrons: n_H2 / exp(0.061 L^2 - 0.342 L + 55.89) + n_O2 / exp(0.031 L^2 - 0.658 L + 56.71)
    mob_kind[E], mob[E], dk[E] = 1, -1.0, 1
    for j, (a, b, c) in ((H2, (0.061, 0.342, 55.89)), (O2, (0.031, 0.658, 56.71))):
        terms[j, E] = 1
        coef[:, j, E] = (real4(a), -real4(b), real4(c))
    # ions: reduced mobilities (cm^2/V/s) in H2 and O2
    for sp, sign, kh2, ko2 in ((Hp, 1.0, 15.0, 15.0), (H2p, 1.0, 13.0, 13.0), (Op, 1.0, 11.0, 3.5),
                               (O2p, 1.0, 8.0, 2.1), (Om, -1.0, 24.0, 3.8)):
        mob_kind[sp], mob[sp], dk[sp] = 2, sign, 2
        terms[H2, sp] = terms[O2, sp] = 1
        coef[0, H2, sp] = kh2 * cm2
        coef[0, O2, sp] = ko2 * cm2
    dco[H], dco[O] = real4(0.03), real4(0.0075)
    mp = M_PROT
    return Chemistry(
        names=("H2", "O2", "E", "H+", "H2+", "O+", "O2+", "O-", "H", "O"),
        reactants=reactants, products=products, k_arrh=k_arrh, elecenergy=elecenergy,
        mobility=mob, dcoeff=dco,
        molmass=np.array([2.0 * mp, 32.0 * mp, M_ELEC, 1.0 * mp, 2.0 * mp, 16.0 * mp, 32.0 * mp, 16.0 * mp, 1.0 * mp,
                          16.0 * mp]),
        charge=np.array([0.0, 0.0, -1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 0.0, 0.0]),
        especnum=3, bgspecnum=1, ionspecmin=4, ionspecmax=8, neutralspecmin=9, neutralspecmax=10,
        mob_kind=mob_kind, mob_terms=terms, mob_coef=coef, dcoeff_kind=dk)


def o2():
    """plasma_pde/chemistry/O2_chem.f90: O2 (background), E, O+, O2+, O-, O
    (the solved neutral); 7 reactions, of which 1-3 (ionisation to O2+,
    dissociative attachment, dissociative ionisation) take getcustomrate's
    fits in 1/T_e (:413-448, rate_kind 1: A exp(a/T + b/T^2 + c/T^3 + d/T^4)
    with the double literals of the fit) and 4-7 the Arrhenius form;
    constant mobilities and diffusion coefficients (:336-381).

    The tables are the doubles assignreactions (:30-202) leaves when
    compiled, taken from the reference build (oracle/ref_driver/ref_chem4.f90);
    a test checks them against it. The module's masses come from
    mgridlib/fundconstants.f90, whose mass_prot equals plasma_pde's."""
    reactants = np.array([[1, 1, 0, 0, 0, 0], [1, 1, 0, 0, 0, 0], [1, 1, 0, 0, 0, 0], [0, 1, 0, 0, 0, 1],
                          [0, 0, 0, 1, 1, 0], [0, 0, 1, 0, 1, 0], [0, 1, 0, 0, 1, 0]], dtype=np.int32)
    products = np.array([[0, 2, 0, 1, 0, 0], [0, 1, 0, 0, 0, 2], [0, 0, 0, 0, 1, 1], [0, 2, 1, 0, 0, 0],
                         [1, 0, 0, 0, 0, 1], [0, 0, 0, 0, 0, 2], [0, 2, 0, 0, 0, 1]], dtype=np.int32)
    k_arrh = np.array([
        [6.683279354557637e-24, 2.0, 146216.703125],
        [4.229999716356402e-15, 0.0, 64521.01953125],
        [4.59999969154597e-17, 0.0, 33769.09375],
        [1.2852985741823762e-17, 0.699999988079071, 157821.203125],
        [1.9899998665601042e-13, 0.0, 0.0],
        [2.6599998216331046e-13, 0.0, 0.0],
        [2.0999998591840297e-16, 0.5, 39400.0],
    ])
    elecenergy = np.array([-12.600000381469727, -5.559999942779541, -3.5999999046325684, -13.600000381469727,
                           0.0, 0.0, -0.9200000166893005])
    rate_kind = np.array([1, 1, 1, 0, 0, 0, 0], dtype=np.int32)
    rate_coef = np.zeros((5, 7))
    rate_coef[:, 0] = (1.043e-13, -2.74e5, -2.0015e8, 5.33e13, -4.001e17)
    rate_coef[:, 1] = (9.577e-16, 1.246e5, -8.647e9, 1.381e14, -6.953e17)
    rate_coef[:, 2] = (1.46e-14, 6.21e4, -7.27e9, 1.25e14, -6.57e17)
    mp = M_PROT
    return Chemistry(
        names=("O2", "E", "O+", "O2+", "O-", "O"),
        reactants=reactants, products=products, k_arrh=k_arrh, elecenergy=elecenergy,
        mobility=np.array([0.0, -20.0, real4(0.4), real4(0.2), -real4(0.4), 0.0]),
        dcoeff=np.array([0.0, 100.0, real4(0.02), real4(0.01), real4(0.02), real4(0.02)]),
        molmass=np.array([32.0 * mp, M_ELEC, 16.0 * mp, 32.0 * mp, 16.0 * mp, 16.0 * mp]),
        charge=np.array([0.0, -1.0, 1.0, 1.0, -1.0, 0.0]),
        especnum=2, bgspecnum=1, ionspecmin=3, ionspecmax=5, neutralspecmin=6, neutralspecmax=6,
        rate_kind=rate_kind, rate_coef=rate_coef)


def h2():
    """plasma_pde/chemistry/H2_chem.f90: H2 (background), E, H+, H2+, H (the
    solved neutral); 8 Arrhenius reactions (every isratearrh of
    assignreactions, :30-220, is .true., so its field-dependent getcustomrate
    at :423-437 is never reached) and constant mobilities and diffusion
    coefficients (:350-391). Tables as the reference build holds them
    (oracle/ref_driver/ref_chem4.f90), checked by a test."""
    reactants = np.array([[0, 1, 0, 0, 1], [1, 1, 0, 0, 0], [0, 1, 0, 1, 0], [0, 1, 0, 1, 0], [0, 0, 0, 1, 1],
                          [1, 0, 1, 0, 0], [1, 1, 0, 0, 0], [1, 1, 0, 0, 0]], dtype=np.int32)
    products = np.array([[0, 2, 1, 0, 0], [0, 2, 1, 0, 1], [0, 1, 1, 0, 1], [0, 2, 2, 0, 0], [1, 0, 1, 0, 0],
                         [0, 0, 0, 1, 1], [0, 2, 0, 1, 0], [0, 1, 0, 0, 2]], dtype=np.int32)
    k_arrh = np.array([
        [6.671210258802429e-17, 0.48930999636650085, 149624.359375],
        [4.673050547002412e-16, 0.4445599913597107, 437818.75],
        [6.780687508393681e-14, 0.04876000061631203, 112450.8515625],
        [1.122866241117621e-16, 0.31393998861312866, 270371.5],
        [8.999999396502984e-16, 0.0, 0.0],
        [1.1899999202042835e-28, 0.0, 0.0],
        [6.269267675269048e-15, 0.17156000435352325, 232987.484375],
        [1.8635879765620025e-08, -1.236680030822754, 146128.84375],
    ])
    elecenergy = np.array([-12.89365005493164, -37.72835922241211, -9.690279960632324, -23.29884910583496,
                           0.0, 0.0, -20.07733917236328, -12.592430114746094])
    mp = M_PROT
    return Chemistry(
        names=("H2", "E", "H+", "H2+", "H"),
        reactants=reactants, products=products, k_arrh=k_arrh, elecenergy=elecenergy,
        mobility=np.array([0.0, -20.0, real4(0.4), real4(0.2), 0.0]),
        dcoeff=np.array([0.0, 100.0, real4(0.02), real4(0.01), real4(0.01)]),
        molmass=np.array([2.0 * mp, M_ELEC, 1.0 * mp, 2.0 * mp, 1.0 * mp]),
        charge=np.array([0.0, -1.0, 1.0, 1.0, 0.0]),
        especnum=2, bgspecnum=1, ionspecmin=3, ionspecmax=4, neutralspecmin=5, neutralspecmax=5)


def etempmin_default():
    """etempmin of eenergy_initialize (eenergyeq.f90:70): 0.1 eV in K, the
    0.1 a default-real literal."""
    return real4(0.1) * EV_TO_K
