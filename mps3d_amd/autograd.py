"""Differentiable linear solves through the hand-written adjoint.

``LinearSolve`` is a thin ctypes ``torch.autograd.Function``: its forward is
one multigrid solve ``x = A(v)^-1 b`` (``mps_solve``); its backward is the
adjoint solve ``lam = A(v)^-T xbar`` (``mps_solve_transpose``: the transposed
operator with the transposed V-cycle, or the forward cycle itself for the
self-adjoint diffusion operator) plus the velocity sensitivity
``-lam^T (dA/dv) x`` (``mps_grad_velocity``). Gradients flow to the
right-hand side and to the three cell-velocity fields; everything else about
the operator (diffusion coefficient, reaction, time step, BC codes) is fixed
per :class:`Operator`. Dirichlet or flux boundary data enter through the
right-hand side, where torch differentiates them.

Tensors are float64 on the GPU. Interior fields have shape (nz, ny, nx)
(C order, i.e. the reference's Fortran (nx, ny, nz) layout); the library's
ghosted arrays are (nz+2, ny+2, nx+2).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as M
from .solver import FieldSolver


def pad(t):
    """Interior (nz,ny,nx) -> zero-ghost (nz+2,ny+2,nx+2), contiguous."""
    return torch.nn.functional.pad(t, (1, 1, 1, 1, 1, 1)).contiguous()


def interior(g):
    return g[1:-1, 1:-1, 1:-1]


class Operator:
    """A fixed operator form for LinearSolve: grid, hierarchy, flux scheme,
    diffusion coefficient, reaction, time term and BC codes (homogeneous
    faces). Velocities and right-hand sides are supplied per solve."""

    def __init__(self, n, length, dcoeff, bc_codes, flux="upwind", reac=None, timederiv=False, dt=1.0,
                 hierarchy="ref", p=(1, 1, 1), tol=1e-13, max_cycles=80, precision="fp64", cycles=None,
                 method="vcycle", gmres_m=30, gmres_restarts=20, start="guess", arith="exact"):
        """cycles: a fixed number of V-cycles per solve instead of a
        tolerance, which makes the discrete solution map a smooth function
        of the inputs (needed when finite differences probe it).
        method: "vcycle" iterates V-cycles (mps_solve); "gmres" runs
        FGMRES(gmres_m) preconditioned by one V-cycle (mps_gmres, precond 1)
        to the same relative tolerance. The adjoint solve iterates V-cycles
        of the transposed operator either way. start: "fmg" begins each
        V-cycle solve with a full-multigrid pass (aligned hierarchy); arith:
        the solver's own rows in the reference's rounding ("exact") or with
        precomputed factors ("fast", the same operator)."""
        self.n = tuple(int(v) for v in n)
        if method not in ("vcycle", "gmres"):
            raise ValueError("method must be 'vcycle' or 'gmres'")
        self.method, self.gmres_m, self.gmres_restarts = method, int(gmres_m), int(gmres_restarts)
        self.fixed = cycles is not None
        if self.fixed and method == "gmres":
            raise ValueError("a fixed cycle count needs method='vcycle'")
        self.tol = float(tol)
        if self.fixed:
            tol, max_cycles = 0.0, int(cycles)
        self.solver = FieldSolver(self.n, length, p=p, hierarchy=hierarchy, precision=precision, flux=flux,
                                  max_cycles=max_cycles, tol=tol, tol_kind="rel", start=start, arith=arith)
        self.solver.set_bc(bc_codes, (0.0,) * 6)
        self.D = pad(dcoeff.to(torch.float64))
        self.reac = None if reac is None else pad(reac.to(torch.float64))
        self.timederiv, self.dt = bool(timederiv), float(dt)
        self._zero = torch.zeros_like(self.D)
        self._static = False  # the coefficients without velocities are set
        self._rhs = None
        self.cycles = []

    def configure(self, vel=None, rhs=None):
        """Sets the operator for velocities vel (3 interior tensors or None)
        and the right-hand side rhs (interior, or None for zero). Without
        velocities the operator is the same every time, so its coefficients
        (and the hierarchy built from them) are set once; the right-hand
        side goes to the solve itself."""
        if vel is None:
            if not self._static:
                self.solver.set_coeffs(self.D, None, None, None, self.reac, self._zero, self.timederiv, self.dt)
                self._static = True
        else:
            vg = tuple(pad(v.detach()) if v is not None else self._zero for v in vel)
            self.solver.set_coeffs(self.D, *vg, self.reac, self._zero, self.timederiv, self.dt)
            self._static = False
        self._rhs = None if rhs is None else rhs.detach().contiguous()

    def solve(self, x0=None):
        x = torch.zeros_like(self.D) if x0 is None else pad(x0.detach())
        b = self._rhs if self._rhs is not None else torch.zeros(self.n[::-1], dtype=torch.float64, device=x.device)
        if self.method == "gmres":
            k, r, st = self.solver.gmres(b, x, m=self.gmres_m, restarts=self.gmres_restarts, tol=self.tol)
        else:
            k, r, _, st = self.solver.solve_rhs(b, x)
        if st != M.OK and not (self.fixed and st == M.NOT_CONVERGED):
            raise M.MpsError(st, "forward solve did not converge (%g after %d %s)" % (
                r, k, "iterations" if self.method == "gmres" else "cycles"))
        self.cycles.append(k)
        return x

    def solve_transpose(self, g):
        lam = torch.zeros_like(self.D)
        k, r, st = self.solver.solve_transpose(g.detach().contiguous(), lam)
        if st != M.OK and not (self.fixed and st == M.NOT_CONVERGED):
            raise M.MpsError(st, "adjoint solve did not converge (%g after %d cycles)" % (r, k))
        return lam

    def grad_velocity(self, lam, x):
        out = [torch.empty(self.n[::-1], dtype=torch.float64, device=lam.device) for _ in range(3)]
        lib = self.solver.lib
        self.solver.order_torch_in()
        try:
            M.check(lib.mps_grad_velocity(self.solver._h, C.c_void_p(lam.data_ptr()), C.c_void_p(x.data_ptr()),
                                          *[C.c_void_p(o.data_ptr()) for o in out]))
        finally:
            self.solver.order_torch_out()
        return out

    def close(self):
        self.solver.close()


class LinearSolve(torch.autograd.Function):
    """x = A(v)^-1 b with the hand-written adjoint in backward."""

    @staticmethod
    def forward(ctx, op: Operator, b, vx=None, vy=None, vz=None, x0=None):
        vel = None if (vx is None and vy is None and vz is None) else (vx, vy, vz)
        op.configure(vel, b)
        xg = op.solve(x0)
        ctx.op = op
        ctx.has_vel = vel is not None
        ctx.save_for_backward(xg, *(v if v is not None else torch.empty(0) for v in (vx, vy, vz)))
        return interior(xg).clone()

    @staticmethod
    def backward(ctx, xbar):
        op = ctx.op
        xg, vx, vy, vz = ctx.saved_tensors
        vel = None
        if ctx.has_vel:
            vel = tuple(v if v.numel() else None for v in (vx, vy, vz))
        op.configure(vel, None)  # the operator of this solve
        lam = op.solve_transpose(xbar.contiguous())
        gb = interior(lam).clone()
        gv = [None, None, None]
        if ctx.has_vel and any(ctx.needs_input_grad[2:5]):
            full = op.grad_velocity(lam, xg)
            # a gradient only for the components that were passed (and want
            # one): a component given as None has no gradient slot
            gv = [full[d] if (v.numel() and ctx.needs_input_grad[2 + d]) else None
                  for d, v in enumerate((vx, vy, vz))]
        return None, gb, gv[0], gv[1], gv[2], None


def linear_solve(op, b, vel=None, x0=None):
    vx, vy, vz = vel if vel is not None else (None, None, None)
    return LinearSolve.apply(op, b, vx, vy, vz, x0)
