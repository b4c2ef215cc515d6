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
