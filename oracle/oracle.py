"""ctypes wrapper of the C restatement in oracle/mps_oracle.c (TEST INFRASTRUCTURE).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this module, as the checker. The product path never does.

Arrays follow the reference layout: ghosted Fortran-order float64 arrays of
shape (nx+2, ny+2, nz+2) and interior arrays of shape (nx, ny, nz).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmps_oracle.so")

BC_NONE, BC_DIRC, BC_ZGRD, BC_FLUX = 0, 1, 2, 3
_BC = {"DIRC": BC_DIRC, "ZGRD": BC_ZGRD, "FLUX": BC_FLUX}

_dp = C.POINTER(C.c_double)


class _Problem(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("px", C.c_int), ("py", C.c_int), ("pz", C.c_int),
                ("lenx", C.c_double), ("leny", C.c_double), ("lenz", C.c_double),
                ("bc", C.c_int * 6), ("bcv", _dp * 6),
                ("D", _dp), ("vx", _dp), ("vy", _dp), ("vz", _dp), ("reac", _dp), ("source", _dp),
                ("timederiv", C.c_int), ("dt", C.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(_Problem)
        L.orc_apply.argtypes = [P, _dp, _dp, _dp, _dp]
        L.orc_bvec.argtypes = [P, _dp, C.c_int, _dp]
        L.orc_smooth.argtypes = [P, _dp, _dp, C.c_int, _dp]
        L.orc_restrict.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_prolong1.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_ghosts.argtypes = [P, _dp]
        L.orc_efield.argtypes = [P, _dp, _dp, C.c_double, C.c_int]
        CP, TO = C.POINTER(_Chem), _TOut
        L.orc_species_transport.argtypes = [P, CP, C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, C.c_double, _dp,
                                            _dp, TO]
        L.orc_eenergy_transport.argtypes = [P, CP, C.c_double, C.c_int, C.c_double, _dp, _dp, _dp, C.c_double,
                                            _dp, _dp, TO]
        L.orc_poisson_transport.argtypes = [P, CP, C.c_double, C.c_int, _dp, _dp, _dp, _dp, TO]
        L.orc_electron_temp.argtypes = [P, _dp, _dp, C.c_double, C.c_double, _dp]
        L.orc_restrict_face.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_vcycle.argtypes = [P, _dp, _dp]
        L.orc_vcycle.restype = C.c_double
        L.orc_solve.argtypes = [P, _dp, C.c_int, C.c_double, _dp]
        L.orc_solve.restype = C.c_int
        L.orc_gmres.argtypes = [P, _dp, _dp, C.c_int, C.c_int, C.c_double]
        L.orc_gmres.restype = C.c_int
        L.orc_rms.argtypes = [C.c_int, _dp]
        L.orc_rms.restype = C.c_double
        L.orc_sg_flux.argtypes = [C.c_double] * 7 + [_dp, _dp, _dp]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def ghosted(interior):
    """Embed an interior (nx,ny,nz) array into a zero-ghost Fortran array."""
    a = np.asarray(interior, dtype=np.float64)
    g = np.zeros((a.shape[0] + 2, a.shape[1] + 2, a.shape[2] + 2), order="F")
    g[1:-1, 1:-1, 1:-1] = a
    return g


class Problem:
    """Reference-mode problem description (see mps_oracle.h)."""

    def __init__(self, n, length=(1.0, 1.0, 1.0), p=(1, 1, 1), bc_codes=("ZGRD",) * 6,
                 bc_vals=(0.0,) * 6, dcoeff=None, velx=None, vely=None, velz=None, reac=None,
                 source=None, timederiv=False, dt=1.0):
        self.n = tuple(int(v) for v in n)
        nx, ny, nz = self.n
        self.p = tuple(int(v) for v in p)
        self.length = tuple(float(v) for v in length)
        self.bc_codes = tuple(bc_codes)
        shapes = [(ny, nz), (ny, nz), (nx, nz), (nx, nz), (nx, ny), (nx, ny)]
        self.faces = [np.asfortranarray(np.broadcast_to(np.asarray(v, dtype=np.float64), s))
                      for v, s in zip(bc_vals, shapes)]

        def g(a):
            return None if a is None else ghosted(np.broadcast_to(np.asarray(a, dtype=np.float64), self.n))

        self.D, self.vx, self.vy, self.vz = g(dcoeff), g(velx), g(vely), g(velz)
        self.reac, self.source = g(reac), g(source)
        self.timederiv, self.dt = bool(timederiv), float(dt)
        s = _Problem()
        s.nx, s.ny, s.nz = self.n
        s.px, s.py, s.pz = self.p
        s.lenx, s.leny, s.lenz = self.length
        for f in range(6):
            s.bc[f] = _BC.get(self.bc_codes[f], BC_NONE)
            s.bcv[f] = _ptr(self.faces[f])
        s.D, s.vx, s.vy, s.vz = _ptr(self.D), _ptr(self.vx), _ptr(self.vy), _ptr(self.vz)
        s.reac, s.source = _ptr(self.reac), _ptr(self.source)
        s.timederiv, s.dt = int(self.timederiv), self.dt
        self._s = s

    @property
    def ref(self):
        return C.byref(self._s)

    def zeros(self, ghost=True):
        nx, ny, nz = self.n
        return np.zeros((nx + 2, ny + 2, nz + 2) if ghost else (nx, ny, nz), order="F")


def apply(p: Problem, u):
    u = np.asfortranarray(u, dtype=np.float64)
    ax, d, o = p.zeros(False), p.zeros(False), p.zeros(False)
    lib().orc_apply(p.ref, _ptr(u), _ptr(ax), _ptr(d), _ptr(o))
    return ax, d, o


def bvec(p: Problem, sterm=None, applybc=True):
    b = p.zeros(False)
    st = None if sterm is None else np.asfortranarray(sterm, dtype=np.float64)
    lib().orc_bvec(p.ref, _ptr(st), int(applybc), _ptr(b))
    return b


def smooth(p: Problem, u, niter=2, sterm=None):
    u = np.array(u, dtype=np.float64, order="F", copy=True)
    res = p.zeros(False)
    st = None if sterm is None else np.asfortranarray(sterm, dtype=np.float64)
    lib().orc_smooth(p.ref, _ptr(u), _ptr(st), int(niter), _ptr(res))
    return u, res


def ghosts(p: Problem, u):
    u = np.array(u, dtype=np.float64, order="F", copy=True)
    lib().orc_ghosts(p.ref, _ptr(u))
    return u


def efield(p: Problem, phi, scale=1.0, limiter=False):
    """update_efield (plasma_pde/poissoneq.f90:565-680): returns (phi with its
    ghost values set, E as a (nx+2, ny+2, nz+2, 3) Fortran array; ghost edges
    and corners of E are zero here, undefined in the reference)."""
    phi = np.array(phi, dtype=np.float64, order="F", copy=True)
    nx, ny, nz = p.n
    E = np.zeros((nx + 2, ny + 2, nz + 2, 3), order="F")
    lib().orc_efield(p.ref, _ptr(phi), _ptr(E), float(scale), int(bool(limiter)))
    return phi, E


# ---------------------------------------------------------------- transport
class _Chem(C.Structure):
    _fields_ = [("nspecies", C.c_int), ("nreac", C.c_int), ("especnum", C.c_int), ("bgspecnum", C.c_int),
                ("ionspecmin", C.c_int), ("ionspecmax", C.c_int),
                ("reactants", C.c_void_p), ("products", C.c_void_p), ("k_arrh", C.c_void_p),
                ("elecenergy", C.c_void_p), ("mobility", C.c_void_p), ("dcoeff", C.c_void_p),
                ("molmass", C.c_void_p), ("charge", C.c_void_p),
                ("mob_kind", C.c_void_p), ("mob_terms", C.c_void_p), ("mob_coef", C.c_void_p),
                ("dcoeff_kind", C.c_void_p), ("rate_kind", C.c_void_p), ("rate_coef", C.c_void_p)]


class _TOut(C.Structure):
    _fields_ = [("vel", _dp), ("dcoeff", _dp), ("reac", _dp), ("source", _dp), ("prod", _dp)]


def _chem(ch):
    """orc_chem for an object with the tables of a chem_module (names,
    reactants/products (nreac, nspecies), k_arrh (nreac, 3), elecenergy,
    mobility, dcoeff, molmass, charge, especnum, bgspecnum, ionspecmin,
    ionspecmax)."""
    keep = [np.asfortranarray(ch.reactants, dtype=np.int32), np.asfortranarray(ch.products, dtype=np.int32),
            np.ascontiguousarray(ch.k_arrh, dtype=np.float64), np.ascontiguousarray(ch.elecenergy, dtype=np.float64),
            np.ascontiguousarray(ch.mobility, dtype=np.float64), np.ascontiguousarray(ch.dcoeff, dtype=np.float64),
            np.ascontiguousarray(ch.molmass, dtype=np.float64), np.ascontiguousarray(ch.charge, dtype=np.float64)]
    models = [None] * 4
    if getattr(ch, "mob_kind", None) is not None:
        models = [np.ascontiguousarray(ch.mob_kind, dtype=np.int32), np.asfortranarray(ch.mob_terms, dtype=np.int32),
                  np.asfortranarray(ch.mob_coef, dtype=np.float64), np.ascontiguousarray(ch.dcoeff_kind, dtype=np.int32)]
        keep += models
    rates = [None] * 2
    if getattr(ch, "rate_kind", None) is not None:
        rates = [np.ascontiguousarray(ch.rate_kind, dtype=np.int32), np.asfortranarray(ch.rate_coef, dtype=np.float64)]
        keep += rates
    s = _Chem(len(ch.mobility), int(np.asarray(ch.reactants).shape[0]), int(ch.especnum), int(ch.bgspecnum),
              int(ch.ionspecmin), int(ch.ionspecmax), *[a.ctypes.data for a in keep[:8]],
              *[None if a is None else a.ctypes.data for a in models + rates])
    return s, keep


def _touts(p: Problem, prod=True):
    nx, ny, nz = p.n
    g = (nx + 2, ny + 2, nz + 2)
    o = {"vel": np.zeros(g + (3,), order="F"), "dcoeff": np.zeros(g, order="F"), "reac": np.zeros(g, order="F"),
         "source": np.zeros(g, order="F")}
    if prod:
        o["prod"] = np.zeros(g, order="F")
    s = _TOut(_ptr(o["vel"]), _ptr(o["dcoeff"]), _ptr(o["reac"]), _ptr(o["source"]), _ptr(o.get("prod")))
    return o, s


def _f(a, ncomp=None):
    return None if a is None else np.asfortranarray(a, dtype=np.float64)


def species_transport(p: Problem, ch, specnum, efield, electemp, numden, scaling, soln=None, dt=1.0,
                      timederiv=True, Tg=300.0):
    """eden_update_transport / ions_update_transport (edeneq.f90:360-444,
    ionseq.f90:478-558): {"vel", "dcoeff", "reac", "source", "prod"}."""
    cs, keep = _chem(ch)
    o, os_ = _touts(p)
    E, T, N, S = _f(efield), _f(electemp), _f(numden), _f(soln)
    sc = np.asarray(scaling, dtype=np.float64)
    lib().orc_species_transport(p.ref, C.byref(cs), int(specnum), float(dt), int(bool(timederiv)), _ptr(E), _ptr(T),
                                _ptr(N), float(Tg), _ptr(sc), _ptr(S), os_)
    return o


def eenergy_transport(p: Problem, ch, efield, electemp, numden, scaling, lscale=1.0, soln=None, dt=1.0,
                      timederiv=True, Tg=300.0):
    """eenergy_update_transport (eenergyeq.f90:394-774)."""
    cs, keep = _chem(ch)
    o, os_ = _touts(p, prod=False)
    E, T, N, S = _f(efield), _f(electemp), _f(numden), _f(soln)
    sc = np.asarray(scaling, dtype=np.float64)
    lib().orc_eenergy_transport(p.ref, C.byref(cs), float(dt), int(bool(timederiv)), float(lscale), _ptr(E), _ptr(T),
                                _ptr(N), float(Tg), _ptr(sc), _ptr(S), os_)
    return o


def poisson_transport(p: Problem, ch, elecden, iondens, scaling, soln=None, dt=1.0, timederiv=True):
    """poisson_update_transport (poissoneq.f90:335-385)."""
    cs, keep = _chem(ch)
    o, os_ = _touts(p, prod=False)
    ne, ni, S = _f(elecden), _f(iondens), _f(soln)
    sc = np.asarray(scaling, dtype=np.float64)
    lib().orc_poisson_transport(p.ref, C.byref(cs), float(dt), int(bool(timederiv)), _ptr(ne), _ptr(ni), _ptr(sc),
                                _ptr(S), os_)
    return o


def electron_temp(p: Problem, eenergysoln, elecden, eenergyscale, etempmin, electemp=None):
    """update_electrontemp (eenergyeq.f90:842-877): (eenergysoln, electemp)."""
    eps = np.array(eenergysoln, dtype=np.float64, order="F", copy=True)
    te = p.zeros() if electemp is None else np.array(electemp, dtype=np.float64, order="F", copy=True)
    ne = _f(elecden)
    lib().orc_electron_temp(p.ref, _ptr(eps), _ptr(ne), float(eenergyscale), float(etempmin), _ptr(te))
    return eps, te


def restrict(fine):
    fine = np.asfortranarray(fine, dtype=np.float64)
    nx, ny, nz = fine.shape
    c = np.zeros((nx // 2, ny // 2, nz // 2), order="F")
    lib().orc_restrict(nx, ny, nz, _ptr(fine), _ptr(c))
    return c


def restrict_face(f):
    f = np.asfortranarray(f, dtype=np.float64)
    c = np.zeros((f.shape[0] // 2, f.shape[1] // 2), order="F")
    lib().orc_restrict_face(f.shape[0], f.shape[1], _ptr(f), _ptr(c))
    return c


def prolong1(coarse_ghosted):
    e = np.asfortranarray(coarse_ghosted, dtype=np.float64)
    cx, cy, cz = (s - 2 for s in e.shape)
    fine = np.zeros((2 * cx, 2 * cy, 2 * cz), order="F")
    lib().orc_prolong1(cx, cy, cz, _ptr(e), _ptr(fine))
    return fine


def vcycle(p: Problem, u, sterm=None):
    u = np.array(u, dtype=np.float64, order="F", copy=True)
    st = None if sterm is None else np.asfortranarray(sterm, dtype=np.float64)
    r = lib().orc_vcycle(p.ref, _ptr(u), _ptr(st))
    return u, r


def solve(p: Problem, u, max_cycles=20, tol=0.0):
    u = np.array(u, dtype=np.float64, order="F", copy=True)
    hist = np.zeros(max_cycles)
    n = lib().orc_solve(p.ref, _ptr(u), int(max_cycles), float(tol), _ptr(hist))
    if n < 0:
        raise ValueError("bad problem")
    return u, hist[:n]


def gmres(p: Problem, b, x0, maxkspdim, nrestarts=1, tol=1e-15):
    b = np.asfortranarray(b, dtype=np.float64)
    x = np.array(x0, dtype=np.float64, order="F", copy=True)
    st = lib().orc_gmres(p.ref, _ptr(b), _ptr(x), int(maxkspdim), int(nrestarts), float(tol))
    if st != 0:
        raise FloatingPointError("NaN in Arnoldi")
    return x


def rms(r):
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).ravel(order="F"))
    return lib().orc_rms(r.size, _ptr(r))


def sg_flux(cL, cR, DL, DR, uL, uR, dx):
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    lib().orc_sg_flux(cL, cR, DL, DR, uL, uR, dx, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value
