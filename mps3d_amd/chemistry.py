"""Chemistry tables for the device-side transport assembly.

A :class:`Chemistry` holds what a reference ``chem_module`` defines in its
``assignreactions`` and ``setspecparams`` (species numbers 1-based,
``reactants``/``products`` as (nreac, nspecies), ``k_arrh`` as (nreac, 3)
rows of A, alpha, Ea): Arrhenius rates in the electron temperature and one
constant mobility and diffusion coefficient per species, the form of
``plasma_pde/chemistry/ion_elec_chem.f90``. ``FieldSolver.set_chemistry``
copies it to the device.
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

    @property
    def nspecies(self):
        return len(self.names)

    @property
    def nreac(self):
        return int(self.reactants.shape[0])

    def arrays(self):
        """The tables as the C ABI takes them (Fortran order, int32/float64)."""
        return {
            "reactants": np.asfortranarray(self.reactants, dtype=np.int32),
            "products": np.asfortranarray(self.products, dtype=np.int32),
            "k_arrh": np.ascontiguousarray(self.k_arrh, dtype=np.float64),  # k_arrh(3, nreac) in memory
            "elecenergy": np.ascontiguousarray(self.elecenergy, dtype=np.float64),
            "mobility": np.ascontiguousarray(self.mobility, dtype=np.float64),
            "dcoeff": np.ascontiguousarray(self.dcoeff, dtype=np.float64),
            "molmass": np.ascontiguousarray(self.molmass, dtype=np.float64),
            "charge": np.ascontiguousarray(self.charge, dtype=np.float64),
        }

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


def etempmin_default():
    """etempmin of eenergy_initialize (eenergyeq.f90:70): 0.1 eV in K, the
    0.1 a default-real literal."""
    return real4(0.1) * EV_TO_K
