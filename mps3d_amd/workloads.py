"""Synthetic inputs for the field-solve workloads (host side, numpy only).

Every generator returns a :class:`Case` holding reference-layout inputs:
interior float64 arrays of shape (nx, ny, nz) (Fortran order), the six BC
codes and face values in the order of ``mgridlib/globalvars.f90:12-17``
(left/right = x-/x+, bottom/top = y-/y+, back/front = z-/z+), and the
per-run constants. The operator they feed is the reference's
``find_AX``/``find_bvec`` (``mgridlib/discretepde.f90:382-1119``):
``-div(D grad u) + div(v u) - reac*u (+ u/dt) = source``.

Seeds use ``numpy.random.default_rng(seed)`` (PCG64); random parameters are
drawn once, in a fixed order, and returned in ``Case.params`` so fixtures
can store them and fields can be rebuilt from the closed-form formulas.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

E_CHARGE = 1.602176634e-19
EPS0 = 8.8541878128e-12
DEFAULT_SEED = 20261202


@dataclass
class Case:
    n: tuple
    length: tuple
    bc_codes: tuple
    bc_vals: list
    dcoeff: np.ndarray | None = None
    velx: np.ndarray | None = None
    vely: np.ndarray | None = None
    velz: np.ndarray | None = None
    reac: np.ndarray | None = None
    source: np.ndarray | None = None
    timederiv: bool = False
    dt: float = 1.0
    params: dict = field(default_factory=dict)

    def centres(self):
        """Cell-centre coordinates (x, y, z) as 1D arrays (generic_pde/pdedefs.f90:329-331)."""
        out = []
        for n, L in zip(self.n, self.length):
            h = L / n
            out.append((np.arange(n) + 0.5) * h)
        return out


def _faces(n, vals):
    nx, ny, nz = n
    shapes = [(ny, nz), (ny, nz), (nx, nz), (nx, nz), (nx, ny), (nx, ny)]
    return [np.asfortranarray(np.broadcast_to(np.asarray(v, dtype=np.float64), s)) for v, s in zip(vals, shapes)]


def laplace_deck(n=64, length=1.0):
    """Laplace deck of tests/Laplace_test (INPUTS:13-27): phi=0 at x=0, phi=1
    at x=L (DIRC), zero gradient elsewhere, D=1, no RHS. The exact discrete
    solution is phi = x at cell centres (tests/Laplace_test/README:1-4)."""
    n3 = (n, n, n) if np.isscalar(n) else tuple(n)
    L3 = (length,) * 3 if np.isscalar(length) else tuple(length)
    return Case(n=n3, length=L3, bc_codes=("DIRC", "DIRC", "ZGRD", "ZGRD", "ZGRD", "ZGRD"),
                bc_vals=_faces(n3, [0.0, 1.0, 0.0, 0.0, 0.0, 0.0]),
                dcoeff=np.ones(n3, order="F"), params={"kind": "laplace"})


def blob_params(seed=DEFAULT_SEED, nblobs=8, length=0.02):
    """Charge blobs: centres ~ U[0.1L, 0.9L]^3, sigma ~ U[1,3] mm,
    log10(A) ~ U[15,16] (m^-3), sign = +-1 with p=1/2; drawn in that order."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.1 * length, 0.9 * length, size=(nblobs, 3))
    sig = rng.uniform(1e-3, 3e-3, size=nblobs)
    amp = 10.0 ** rng.uniform(15.0, 16.0, size=nblobs)
    sgn = np.where(rng.uniform(size=nblobs) < 0.5, -1.0, 1.0)
    return np.column_stack([c, sig, amp * sgn])


def varcoef_poisson(n=129, length=0.02, seed=DEFAULT_SEED, v_top=1000.0, eps_slab=4.0,
                    slab_frac=0.25, blobs=None):
    """Variable-coefficient Poisson with space charge on a cube of side
    ``length``: D = eps_r (eps_slab for z < slab_frac*length, else 1),
    source = rho/eps0 with rho = e * sum_k s_k A_k exp(-|x-c_k|^2/(2 sigma_k^2)),
    phi = 0 V at z=0 and v_top at z=length (DIRC), zero gradient on x and y.
    Unscaled SI units, as the generic driver uses (generic_pde/pdedefs.f90)."""
    n3 = (n, n, n) if np.isscalar(n) else tuple(n)
    L3 = (length,) * 3
    if blobs is None:
        blobs = blob_params(seed, length=length)
    case = Case(n=n3, length=L3, bc_codes=("ZGRD", "ZGRD", "ZGRD", "ZGRD", "DIRC", "DIRC"),
                bc_vals=_faces(n3, [0.0, 0.0, 0.0, 0.0, 0.0, v_top]),
                params={"kind": "varcoef", "seed": seed, "blobs": blobs, "v_top": v_top})
    x, y, z = case.centres()
    eps = np.where(z < slab_frac * length, eps_slab, 1.0)
    case.dcoeff = np.asfortranarray(np.broadcast_to(eps[None, None, :], n3)).copy(order="F")
    rho = np.zeros(n3, order="F")
    for cx, cy, cz, s, a in blobs:
        gx = np.exp(-((x - cx) ** 2) / (2 * s * s))
        gy = np.exp(-((y - cy) ** 2) / (2 * s * s))
        gz = np.exp(-((z - cz) ** 2) / (2 * s * s))
        rho += a * gx[:, None, None] * gy[None, :, None] * gz[None, None, :]
    case.source = np.asfortranarray(rho * (E_CHARGE / EPS0))
    return case


def mms_cdr(n=(80, 48, 32), length=(1.0, 0.6, 0.4), dt=0.1, u_old=None):
    """Convection-diffusion-reaction deck of tests/gpde_test: dphi/dt +
    div(-x_hat phi) = div(0.5 grad phi) + 2 phi - x^2, phi=0 at x=0 (DIRC),
    outward flux -0.25 at x=L (FLUX), zero gradient elsewhere; steady state
    phi = (x^2 - x)/2 (tests/gpde_test/README:1-6; coefficients of
    generic_pde/pdedefs.f90:333-349)."""
    n3 = tuple(n)
    case = Case(n=n3, length=tuple(length), bc_codes=("DIRC", "FLUX", "ZGRD", "ZGRD", "ZGRD", "ZGRD"),
                bc_vals=_faces(n3, [0.0, -0.25, 0.0, 0.0, 0.0, 0.0]), timederiv=True, dt=dt,
                params={"kind": "mms"})
    x, _, _ = case.centres()
    case.velx = np.full(n3, -1.0, order="F")
    case.vely = np.zeros(n3, order="F")
    case.velz = np.zeros(n3, order="F")
    case.dcoeff = np.full(n3, 0.5, order="F")
    case.reac = np.full(n3, 2.0, order="F")
    src = np.asfortranarray(np.broadcast_to(-(x ** 2)[:, None, None], n3)).copy(order="F")
    if u_old is not None:
        src = src + np.asarray(u_old) / dt
    case.source = src
    return case


def random_cdr(n, seed, length=(1.2, 0.7, 0.9), codes=("DIRC", "FLUX", "ZGRD", "DIRC", "DIRC", "ZGRD"),
               timederiv=True, dt=0.37):
    """Random coefficients and BC faces for operator-level checks."""
    rng = np.random.default_rng(seed)
    n3 = tuple(n)
    nx, ny, nz = n3
    faces = [rng.standard_normal(s) for s in [(ny, nz), (ny, nz), (nx, nz), (nx, nz), (nx, ny), (nx, ny)]]
    c = Case(n=n3, length=tuple(length), bc_codes=tuple(codes), bc_vals=_faces(n3, faces),
             timederiv=timederiv, dt=dt, params={"kind": "random", "seed": seed})
    c.dcoeff = np.asfortranarray(np.exp(rng.standard_normal(n3)))
    c.velx = np.asfortranarray(rng.standard_normal(n3))
    c.vely = np.asfortranarray(rng.standard_normal(n3))
    c.velz = np.asfortranarray(rng.standard_normal(n3))
    c.reac = np.asfortranarray(rng.standard_normal(n3))
    c.source = np.asfortranarray(rng.standard_normal(n3))
    return c
