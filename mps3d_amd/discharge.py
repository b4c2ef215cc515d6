"""Multi-species discharge driver: Poisson, drift-diffusion species with
Scharfetter-Gummel fluxes, explicit chemistry and an optional electron
energy equation, advanced by implicit steps whose linear solves all run on
the HIP field-solve path.

One step, in the order of plasma_pde/solve_manager.f90:405-477:

1. Poisson: -div(eps_r grad phi) = (e/eps0) sum_s Z_s n_s, electrode
   potentials on the z faces (DIRC), zero gradient on the side walls
   (poisson_update_bcs, plasma_pde/poissoneq.f90);
2. the field E = -grad phi at cell centres, central differences with the
   ghost rules of update_boundary_ghostvalues (update_efield,
   poissoneq.f90:565-680);
3. reaction rates from the old state (getspecproduction,
   chemistry/ion_elec_chem.f90);
4. every species: an implicit step n/dt + div(v n - D grad n) =
   n_old/dt + S with v = sign(Z) mu E and Scharfetter-Gummel face fluxes
   (mps1d/source/convdiff.f90:10-64 taken to 3D; pure diffusion for neutral
   species), thermal outflow n cbar / 4 through FLUX faces computed from the
   old state (eden_update_bcs, plasma_pde/edeneq.f90:203-358);
5. optionally the electron energy density n_eps = n_e eps_bar, with the
   mean energy eps_bar = 3/2 T_e (eV, so n_eps is in eV m^-3), with
   v = 5/3 v_e, D = 5/3 D_e, Joule heating mu_e n_e |E|^2, elastic loss
   3 (m_e/M) nu_m n_e (T_e - T_g) and the reactions' energy losses
   (plasma_pde/eenergyeq.f90:394-456, 559-673, 714-764); without it the mean
   energy follows the local field.

Species solves are FGMRES(30) preconditioned by one V-cycle (``method=
"gmres"``), Poisson iterates V-cycles; with ``precision="fp32dc"`` both
smooth in FP32 under FP64 residuals. Boundary data enter through the
right-hand sides, which torch builds, so torch autograd differentiates the
whole run, every linear solve by the hand-written adjoint
(:class:`mps3d_amd.autograd.LinearSolve`). Fields are float64 (nz, ny, nx)
tensors on the GPU; units are SI with energies and temperatures in eV.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from .autograd import Operator, linear_solve

E_CHARGE = 1.602176634e-19
EPS0 = 8.8541878128e-12
KB = 1.380649e-23
M_E = 9.1093837015e-31
AMU = 1.66053906660e-27
M_AR = 39.948 * AMU
TORR = 133.322368


@dataclass
class Species:
    """A transported species: charge number, mobility magnitude (m^2/Vs),
    diffusion coefficient (m^2/s), particle mass (kg, for the thermal speed
    at FLUX faces) and the BC codes of the z faces (electrodes) and of the
    side walls."""
    name: str
    charge: int = 0
    mu: float = 0.0
    D: float = 0.0
    mass: float = M_AR
    electrode: str = "FLUX"
    wall: str = "ZGRD"


@dataclass
class Reaction:
    """reactants -> products at rate k * prod(reactant densities); "gas" is
    the background density. ``rate`` is a constant (m^3/s) or a function of
    the mean electron energy (eV); ``energy`` is the electron energy lost per
    event (eV)."""
    reactants: tuple
    products: tuple
    rate: object
    energy: float = 0.0

    def coefficient(self, eps_bar):
        return self.rate(eps_bar) if callable(self.rate) else self.rate


def arrhenius(A, Ea):
    """k(eps) = A exp(-Ea / eps) (A in m^3/s, Ea and eps in eV)."""
    return lambda eps: A * torch.exp(-Ea / eps)


def plate_electrodes(nx, ny, device):
    """z = 0 grounded, z = L at V[0] (a parallel-plate gap)."""
    one = torch.ones((ny, nx), dtype=torch.float64, device=device)

    def faces(V):
        return 0.0 * one, V[0] * one
    return faces, 1


def split_electrodes(nx, ny, device):
    """z = 0 split at x = L/2 into two segments at V[0] (x < L/2) and V[1]
    (x >= L/2); z = L grounded."""
    left = torch.zeros((ny, nx), dtype=torch.float64, device=device)
    left[:, : nx // 2] = 1.0
    right = 1.0 - left

    def faces(V):
        return V[0] * left + V[1] * right, 0.0 * left
    return faces, 2


class Discharge:
    def __init__(self, n, length, species, reactions, electrodes, dt=1e-11, p_torr=1.0, T_gas=300.0,
                 energy=False, eps_lfa=(1.0, 5e4), eps_r=None, precision="fp64", hierarchy="ref", p=(1, 1, 1),
                 tol=1e-10, poisson_tol=None, gmres_m=30, device="cuda"):
        """electrodes: (faces, nV) with faces(V) -> (phi on z = 0, phi on
        z = L) as (ny, nx) tensors of the voltage parameters V (nV of them).
        eps_lfa = (eps0, E_ref): mean energy eps0 (1 + |E| / E_ref) when
        there is no energy equation."""
        self.n = (n, n, n) if isinstance(n, int) else tuple(n)
        nx, ny, nz = self.n
        if not (nx == ny == nz):
            raise ValueError("a cube of cells is assumed (one spacing h)")
        self.L = float(length)
        self.h = self.L / nz
        self.dt = float(dt)
        self.Ng = p_torr * TORR / (KB * T_gas)
        self.Tg_eV = KB * T_gas / E_CHARGE
        self.species = list(species)
        self.names = [s.name for s in self.species]
        self.reactions = list(reactions)
        self.energy = bool(energy)
        if self.energy and "e" not in self.names:
            raise ValueError("the energy equation needs a species named 'e'")
        self.eps_lfa = eps_lfa
        self.faces, self.nV = electrodes
        self.dev = torch.device(device)
        shape = (nz, ny, nx)
        self.shape = shape
        ones = torch.ones(shape, dtype=torch.float64, device=self.dev)
        self.eps_r = ones if eps_r is None else eps_r.to(self.dev, torch.float64)
        L3 = (self.L,) * 3
        common = dict(hierarchy=hierarchy, p=p, precision=precision, arith="fast")
        ptol = tol if poisson_tol is None else poisson_tol
        self.poisson = Operator(self.n, L3, self.eps_r, ("ZGRD",) * 4 + ("DIRC", "DIRC"), flux="upwind",
                                tol=ptol, start="fmg" if hierarchy == "aligned" else "guess", **common)
        self._phi = None  # the last potential: the next Poisson solve starts from it
        self.ops = []
        for s in self.species:
            codes = (s.wall,) * 4 + (s.electrode,) * 2
            drift = s.charge != 0 and s.mu > 0
            self.ops.append(Operator(self.n, L3, s.D * ones, codes, flux="sg" if drift else "upwind",
                                     timederiv=True, dt=self.dt, tol=tol, method="gmres", gmres_m=gmres_m,
                                     **common))
        if self.energy:
            e = self.species[self.names.index("e")]
            codes = (e.wall,) * 4 + (e.electrode,) * 2
            self.energy_op = Operator(self.n, L3, (5.0 / 3.0) * e.D * ones, codes, flux="sg", timederiv=True,
                                      dt=self.dt, tol=tol, method="gmres", gmres_m=gmres_m, **common)
        self.step_times = []

    # ------------------------------------------------------------ pieces
    def gaussian(self, sigma=1e-3, centre=None):
        c = (torch.arange(self.n[0], dtype=torch.float64, device=self.dev) + 0.5) * self.h
        x0 = self.L / 2 if centre is None else centre
        g = torch.exp(-((c - x0) ** 2) / (2 * sigma * sigma))
        return g[:, None, None] * g[None, :, None] * g[None, None, :]

    def solve_poisson(self, rhs):
        phi = linear_solve(self.poisson, rhs, x0=self._phi)
        self._phi = phi.detach()
        return phi

    def solve_species(self, k, rhs, vel, x0):
        return linear_solve(self.ops[k], rhs, vel, x0=x0)

    def solve_energy(self, rhs, vel, x0):
        return linear_solve(self.energy_op, rhs, vel, x0=x0)

    def potential(self, dens, V):
        """phi from the charge of the state and the electrode potentials."""
        rho = torch.zeros(self.shape, dtype=torch.float64, device=self.dev)
        for s, nsp in zip(self.species, dens):
            if s.charge:
                rho = rho + s.charge * nsp
        rhs = (E_CHARGE / EPS0) * rho
        lo, hi = self.faces(V)
        w = 2.0 / (self.h * self.h)  # DIRC row weight 2 D / h^2 (find_bvec)
        rhs = torch.cat([(rhs[:1] + w * self.eps_r[:1] * lo[None]), rhs[1:-1],
                         (rhs[-1:] + w * self.eps_r[-1:] * hi[None])], dim=0)
        return self.solve_poisson(rhs)

    def efield(self, phi, V):
        h = self.h
        lo, hi = self.faces(V)
        px = torch.cat([phi[:, :, :1], phi, phi[:, :, -1:]], dim=2)
        py = torch.cat([phi[:, :1, :], phi, phi[:, -1:, :]], dim=1)
        pz = torch.cat([2.0 * lo[None] - phi[:1], phi, 2.0 * hi[None] - phi[-1:]], dim=0)
        return (-(px[:, :, 2:] - px[:, :, :-2]) / (2 * h), -(py[:, 2:, :] - py[:, :-2, :]) / (2 * h),
                -(pz[2:] - pz[:-2]) / (2 * h))

    def mean_energy(self, dens, neps, emag):
        if self.energy:
            ne = dens[self.names.index("e")]
            return torch.clamp(neps / ne, min=0.05, max=500.0)  # eps_bar = 3/2 T_e (eV)
        e0, eref = self.eps_lfa
        return e0 * (1.0 + emag / eref)

    def rates(self, dens, eps_bar):
        out = []
        for r in self.reactions:
            R = r.coefficient(eps_bar)
            for name in r.reactants:
                R = R * (self.Ng if name == "gas" else dens[self.names.index(name)])
            out.append(R)
        return out

    def thermal_outflow(self, s, nsp, T_eV):
        """n cbar / 4 / h in the boundary layers of FLUX electrodes, the
        explicit outward flux find_bvec subtracts (b -= q / h)."""
        if s.electrode != "FLUX":
            return None
        cbar = torch.sqrt(8.0 * E_CHARGE * T_eV / (math.pi * s.mass)) if torch.is_tensor(T_eV) else \
            math.sqrt(8.0 * E_CHARGE * T_eV / (math.pi * s.mass))
        q = 0.25 * nsp * cbar / self.h
        z = torch.zeros_like(nsp[1:-1])
        return torch.cat([q[:1], z, q[-1:]], dim=0)

    # ------------------------------------------------------------ one step
    def step(self, dens, neps, V):
        """dens: tuple of species densities; neps: electron energy density
        (or None). Returns the new (dens, neps)."""
        phi = self.potential(dens, V)
        ex, ey, ez = self.efield(phi, V)
        e2 = ex * ex + ey * ey + ez * ez
        emag = torch.sqrt(e2 + 1e-30)
        eps_bar = self.mean_energy(dens, neps, emag)
        Te = (2.0 / 3.0) * eps_bar
        R = self.rates(dens, eps_bar)
        new = []
        for k, (s, nsp) in enumerate(zip(self.species, dens)):
            S = torch.zeros_like(nsp)
            for r, Rr in zip(self.reactions, R):
                nu = r.products.count(s.name) - r.reactants.count(s.name)
                if nu:
                    S = S + nu * Rr
            rhs = nsp / self.dt + S
            q = self.thermal_outflow(s, nsp, Te if s.name == "e" else self.Tg_eV)
            if q is not None:
                rhs = rhs - q
            vel = None
            if s.charge != 0 and s.mu > 0:
                sg = math.copysign(s.mu, s.charge)
                vel = (sg * ex, sg * ey, sg * ez)
            new.append(self.solve_species(k, rhs, vel, nsp))
        neps_new = None
        if self.energy:
            ie = self.names.index("e")
            e = self.species[ie]
            ne = dens[ie]
            nu_m = E_CHARGE / (M_E * e.mu)
            loss = 3.0 * (M_E / M_AR) * nu_m * ne * (Te - self.Tg_eV)
            for r, Rr in zip(self.reactions, R):
                if r.energy:
                    loss = loss + r.energy * Rr
            rhs = neps / self.dt + e.mu * ne * e2 - loss
            q = self.thermal_outflow(e, ne, Te)
            if q is not None:
                rhs = rhs - 2.0 * Te * q  # energy flux 2 T_e per electron
            ve = -(5.0 / 3.0) * e.mu
            neps_new = self.solve_energy(rhs, (ve * ex, ve * ey, ve * ez), neps)
        return tuple(new), neps_new

    # ------------------------------------------------------------ runs
    def run(self, V, nsteps, state, checkpoint_every=0, timed=False):
        """The state after nsteps steps; state = (dens tuple, neps or None).
        With checkpoint_every > 0 (and grad enabled) the steps are recomputed
        segment by segment in the backward pass."""
        dens, neps = state
        ns = len(dens)

        def segment(m, V_, *flat):
            d, ne_ = tuple(flat[:ns]), (flat[ns] if self.energy else None)
            for _ in range(m):
                d, ne_ = self.step(d, ne_, V_)
            return d + ((ne_,) if self.energy else ())

        done = 0
        flat = tuple(dens) + ((neps,) if self.energy else ())
        while done < nsteps:
            m = min(checkpoint_every, nsteps - done) if checkpoint_every > 0 else 1
            t = time.perf_counter()
            if checkpoint_every > 0 and torch.is_grad_enabled():
                flat = checkpoint(segment, m, V, *flat, use_reentrant=False)
            else:
                flat = segment(m, V, *flat)
            if timed:
                torch.cuda.synchronize()
                self.step_times.append((time.perf_counter() - t) / m)
            done += m
        return tuple(flat[:ns]), (flat[ns] if self.energy else None)

    def field_energy(self, state, V):
        """1/2 int eps0 eps_r |E|^2 dV (J) of the state."""
        phi = self.potential(state[0], V)
        ex, ey, ez = self.efield(phi, V)
        return 0.5 * EPS0 * (self.eps_r * (ex * ex + ey * ey + ez * ez)).sum() * self.h ** 3

    def potential_of(self, state, V):
        return self.potential(state[0], V)

    def solve_counts(self):
        """Per operator: the iteration (GMRES) or cycle counts of its solves."""
        out = {"poisson": list(self.poisson.cycles)}
        for s, op in zip(self.species, self.ops):
            out[s.name] = list(op.cycles)
        if self.energy:
            out["energy"] = list(self.energy_op.cycles)
        return out

    def close(self):
        self.poisson.close()
        for op in self.ops:
            op.close()
        if self.energy:
            self.energy_op.close()


# ------------------------------------------------------------------ workloads
def glow_discharge(n=128, length=0.02, dt=1e-11, precision="fp64", hierarchy="ref", p=(1, 1, 1), tol=1e-10,
                   device="cuda"):
    """Coupled parallel-plate glow discharge: electrons, one positive ion,
    electron energy and Poisson over a 2 cm gap with a square cross-section.
    Argon-like gas at 1 Torr and 300 K; mu_e = 20, D_e = 100, mu_i = 0.2,
    D_i = 0.01 (chemistry/ion_elec_chem.f90:213-240); ionisation Arrhenius in
    the mean electron energy, k = 1e-13 exp(-15 eV / eps) m^3/s, 15 eV lost
    per event. Electrodes: z = 0 grounded, z = L at the applied voltage;
    thermal outflow of both species at the electrodes (edeneq.f90:229-237),
    zero gradient on the side walls. Returns (model, initial state): a
    uniform 1e13 m^-3 background plus a centred Gaussian of 1e16 m^-3
    (sigma 1 mm) for both species, mean energy 1 eV."""
    nx = n if isinstance(n, int) else n[0]
    species = [Species("e", -1, 20.0, 100.0, mass=M_E), Species("Ar+", +1, 0.2, 0.01, mass=M_AR)]
    reactions = [Reaction(("e", "gas"), ("e", "e", "Ar+"), arrhenius(1e-13, 15.0), energy=15.0)]
    model = Discharge(n, length, species, reactions, plate_electrodes(nx, nx, device), dt=dt, energy=True,
                      precision=precision, hierarchy=hierarchy, p=p, tol=tol, device=device)
    ne = 1e13 + 1e16 * model.gaussian(1e-3)
    return model, ((ne.clone(), ne.clone()), 1.0 * ne)


def five_species_params(seed=20261202):
    """Seeded chemistry of the five-species case, drawn in this order:
    three Arrhenius electron-impact rates (A ~ 10^U[-14, -13] m^3/s,
    Ea ~ U[11, 16] eV: ionisation, excitation to A*1, excitation to A*2),
    then the attachment and detachment rate constants, log-uniform in
    [1e-17, 1e-15] m^3/s."""
    rng = np.random.default_rng(seed)
    A = 10.0 ** rng.uniform(-14.0, -13.0, size=3)
    Ea = np.sort(rng.uniform(11.0, 16.0, size=3))[::-1]  # ionisation the highest
    k_att, k_det = 10.0 ** rng.uniform(-17.0, -15.0, size=2)
    return {"A": A.tolist(), "Ea": Ea.tolist(), "k_att": float(k_att), "k_det": float(k_det)}


def five_species(n=257, length=0.04, dt=1e-11, seed=20261202, precision="fp32dc", hierarchy="aligned",
                 p=(1, 1, 1), tol=1e-10, poisson_tol=None, E_ref=5e3, device="cuda"):
    """Five species over a cube of side 0.04 m: e-, one positive ion A+, one
    negative ion A- and two excited neutrals A*1, A*2, at 1 Torr and 300 K.
    Reactions: e + A -> 2e + A+, e + A -> e + A*1, e + A -> e + A*2 (seeded
    Arrhenius in the local-field mean energy 1 eV (1 + |E| / E_ref)),
    attachment e + A -> A-, detachment A- + A*1 -> e + 2A (seeded constants).
    Electrodes: z = 0 split at x = L/2 into two segments (voltage parameters
    V[0], V[1]), z = L grounded; thermal outflow of every species at the z
    faces, zero gradient on the side walls. Returns (model, initial state,
    params): e- 1e13 + 1e16 G, A- 1e12 + 1e14 G, A+ = e- + A- (neutral), A*1
    and A*2 1e14 + 1e16 G, with G a centred Gaussian of sigma 2 mm."""
    nx = n if isinstance(n, int) else n[0]
    prm = five_species_params(seed)
    A, Ea = prm["A"], prm["Ea"]
    species = [Species("e", -1, 20.0, 100.0, mass=M_E), Species("A+", +1, 0.2, 0.01),
               Species("A-", -1, 0.2, 0.01), Species("A*1", 0, 0.0, 0.01), Species("A*2", 0, 0.0, 0.01)]
    reactions = [Reaction(("e", "gas"), ("e", "e", "A+"), arrhenius(A[0], Ea[0])),
                 Reaction(("e", "gas"), ("e", "A*1"), arrhenius(A[1], Ea[1])),
                 Reaction(("e", "gas"), ("e", "A*2"), arrhenius(A[2], Ea[2])),
                 Reaction(("e", "gas"), ("A-",), prm["k_att"]),
                 Reaction(("A-", "A*1"), ("e",), prm["k_det"])]
    model = Discharge(n, length, species, reactions, split_electrodes(nx, nx, device), dt=dt, energy=False,
                      eps_lfa=(1.0, E_ref), precision=precision, hierarchy=hierarchy, p=p, tol=tol,
                      poisson_tol=poisson_tol, device=device)
    G = model.gaussian(2e-3)
    ne = 1e13 + 1e16 * G
    nn = 1e12 + 1e14 * G
    nx1 = 1e14 + 1e16 * G
    state = ((ne, ne + nn, nn, nx1.clone(), nx1.clone()), None)
    return model, state, prm


__all__ = ["Species", "Reaction", "Discharge", "arrhenius", "plate_electrodes", "split_electrodes",
           "glow_discharge", "five_species", "five_species_params"]
