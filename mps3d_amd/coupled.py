"""Coupled implicit drift-diffusion step: electrons, one immobile positive
ion and Poisson, with a differentiable objective.

This is the driver the field-solve path serves (the per-step order of
plasma_pde/solve_manager.f90:405-477 -- Poisson, field, species), reduced to
what the adjoint check needs:

* Poisson: -div(eps_r grad phi) = e (n_i - n_e)/eps0, phi = 0 at z = 0 and
  V_app at z = L (DIRC), zero gradient on the sides (as poisson_update_bcs
  uses for the electrodes and walls);
* field: E = -grad phi by central differences at cell centres with the
  ghost rules of update_boundary_ghostvalues (DIRC 2g - u, else u);
* electrons: implicit Scharfetter-Gummel step
  n_e/dt + div(v n_e - D_e grad n_e) = n_e_old/dt + k_iz N_g n_e_old + q,
  with v = -s_mu mu_e E, absorbing (DIRC 0) electrodes, zero-gradient walls;
* ions gain what ionisation produces (immobile on this time scale);
* ionisation: k_iz = A exp(-Ea / eps_bar), with a local-field mean energy
  eps_bar = eps0_eV (1 + |E|/E_ref).

The objective is the volume-integrated ionisation rate at the final step,
J = sum k_iz(|E|) N_g n_e h^3. Gradients with respect to V_app, the mobility
scale s_mu and the per-cell source q come from torch autograd over the
step, with every linear solve differentiated by the hand-written adjoint
(mps3d_amd.autograd.LinearSolve) and steps checkpointed in segments.
"""
from __future__ import annotations

import math

import torch
from torch.utils.checkpoint import checkpoint

from .autograd import Operator, linear_solve

E_CHARGE = 1.602176634e-19
EPS0 = 8.8541878128e-12
KB = 1.380649e-23


class GlowModel:
    def __init__(self, n=96, length=0.02, dt=1e-11, mu_e=20.0, D_e=100.0, p_torr=1.0, T=300.0,
                 A_iz=1e-13, Ea_eV=15.0, eps0_eV=1.0, E_ref=5e4, device="cuda", hierarchy="ref", p=(1, 1, 1),
                 tol=1e-13, cycles=None):
        """cycles = (poisson, electrons): fixed V-cycle counts per solve
        instead of the tolerance (smooth solution maps for finite
        differences)."""
        self.n = (n, n, n) if isinstance(n, int) else tuple(n)
        nx, ny, nz = self.n
        self.L = float(length)
        self.h = self.L / nz
        self.dt = float(dt)
        self.mu_e, self.D_e = float(mu_e), float(D_e)
        self.Ng = p_torr * 133.322 / (KB * T)
        self.A_iz, self.Ea, self.eps0_eV, self.E_ref = A_iz, Ea_eV, eps0_eV, E_ref
        self.dev = torch.device(device)
        shape = (nz, ny, nx)
        ones = torch.ones(shape, dtype=torch.float64, device=self.dev)
        z_codes = ("ZGRD", "ZGRD", "ZGRD", "ZGRD", "DIRC", "DIRC")
        L3 = (self.L, self.L, self.L)
        cp, ce = cycles if cycles is not None else (None, None)
        self.poisson = Operator(self.n, L3, ones, z_codes, flux="upwind", hierarchy=hierarchy, p=p, tol=tol,
                                cycles=cp)
        self.electrons = Operator(self.n, L3, self.D_e * ones, z_codes, flux="sg", timederiv=True, dt=self.dt,
                                  hierarchy=hierarchy, p=p, tol=tol, cycles=ce)
        h = self.h
        c = (torch.arange(nx, dtype=torch.float64, device=self.dev) + 0.5) * h
        self.x = c
        self.top = torch.zeros(shape, dtype=torch.float64, device=self.dev)
        self.top[-1] = 2.0 / (h * h)  # DIRC row weight of the z = L plate (D = 1)

    # ------------------------------------------------------------------
    def initial_state(self, background=1e13, peak=1e16, sigma=1e-3):
        c = self.x
        r2 = ((c - self.L / 2) ** 2)[None, None, :] + ((c - self.L / 2) ** 2)[None, :, None] + \
             ((c - self.L / 2) ** 2)[:, None, None]
        ne = background + peak * torch.exp(-r2 / (2 * sigma * sigma))
        return ne.clone(), ne.clone()

    # the two linear solves (overridable, e.g. by dense reference solves)
    def solve_poisson(self, rhs):
        return linear_solve(self.poisson, rhs)

    def solve_electrons(self, rhs, vel, x0):
        return linear_solve(self.electrons, rhs, vel, x0=x0)

    def potential(self, ne, ni, V):
        rhs = (E_CHARGE / EPS0) * (ni - ne) + self.top * V
        return self.solve_poisson(rhs)

    def efield(self, phi, V):
        """E = -grad phi at cell centres (central differences with ghosts)."""
        h = self.h
        # x and y walls: zero gradient ghosts; z plates: DIRC ghosts
        px = torch.cat([phi[:, :, :1], phi, phi[:, :, -1:]], dim=2)
        py = torch.cat([phi[:, :1, :], phi, phi[:, -1:, :]], dim=1)
        pz = torch.cat([-phi[:1], phi, 2.0 * V - phi[-1:]], dim=0)
        ex = -(px[:, :, 2:] - px[:, :, :-2]) / (2 * h)
        ey = -(py[:, 2:, :] - py[:, :-2, :]) / (2 * h)
        ez = -(pz[2:] - pz[:-2]) / (2 * h)
        return ex, ey, ez

    def k_iz(self, ex, ey, ez):
        emag = torch.sqrt(ex * ex + ey * ey + ez * ez + 1e-30)
        eps = self.eps0_eV * (1.0 + emag / self.E_ref)
        return self.A_iz * torch.exp(-self.Ea / eps)

    def step(self, ne, ni, V, s_mu, q):
        phi = self.potential(ne, ni, V)
        ex, ey, ez = self.efield(phi, V)
        vel = (-s_mu * self.mu_e * ex, -s_mu * self.mu_e * ey, -s_mu * self.mu_e * ez)
        prod = self.k_iz(ex, ey, ez) * self.Ng * ne + q
        ne_new = self.solve_electrons(ne / self.dt + prod, vel, ne)
        ni_new = ni + self.dt * prod
        return ne_new, ni_new

    def objective(self, ne, ni, V):
        phi = self.potential(ne, ni, V)
        ex, ey, ez = self.efield(phi, V)
        return (self.k_iz(ex, ey, ez) * self.Ng * ne).sum() * self.h ** 3

    def run(self, V, s_mu, q, nsteps, checkpoint_every=0, state=None):
        """J after nsteps implicit steps; with checkpoint_every > 0 the steps
        are recomputed segment by segment in the backward pass."""
        ne, ni = self.initial_state() if state is None else state

        def segment(ne_, ni_, V_, s_, q_, m):
            for _ in range(m):
                ne_, ni_ = self.step(ne_, ni_, V_, s_, q_)
            return ne_, ni_

        done = 0
        while done < nsteps:
            m = min(checkpoint_every, nsteps - done) if checkpoint_every > 0 else nsteps - done
            if checkpoint_every > 0 and torch.is_grad_enabled():
                ne, ni = checkpoint(segment, ne, ni, V, s_mu, q, m, use_reentrant=False)
            else:
                ne, ni = segment(ne, ni, V, s_mu, q, m)
            done += m
        return self.objective(ne, ni, V)

    def close(self):
        self.poisson.close()
        self.electrons.close()


def source_pulse(model, amplitude=1e23, sigma=1e-3):
    """A localised Gaussian source at the domain centre (m^-3 s^-1)."""
    c = model.x
    L = model.L
    r2 = ((c - L / 2) ** 2)[None, None, :] + ((c - L / 2) ** 2)[None, :, None] + ((c - L / 2) ** 2)[:, None, None]
    return amplitude * torch.exp(-r2 / (2 * sigma * sigma))


def seeded_source(shape, seed=20261202, scale=1e13, device="cuda"):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return (scale * torch.randn(shape, generator=g, dtype=torch.float64)).to(device)


def central_difference(f, p, rel=1e-6):
    """(f(p + d) - f(p - d)) / 2d with d = rel * |p| (or rel when p = 0)."""
    d = rel * abs(p) if p != 0 else rel
    return (f(p + d) - f(p - d)) / (2 * d)


__all__ = ["GlowModel", "seeded_source", "source_pulse", "central_difference", "math"]
