"""The plasma solver's time-step loop on the device.

:class:`PlasmaRun` is ``timestepping`` of plasma_pde/solve_manager.f90:363-502
for a chemistry of :mod:`mps3d_amd.chemistry` (ion_elec: one ion; H2_O2:
five ions and two solved neutral species): per step the Poisson solve and
the field, the electron energy solve and the electron temperature, the
electron solve, the ion solves and the neutral solves, each as the
reference's ``*_solve`` does it -- transport assembly,
boundary update (``*_update_bcs``), the loop of at most 20 reference V-cycles
to an RMS residual of 1e-11, floors and ghost values -- with every field
resident in HBM between steps. Each equation has its own context (the
reference hierarchy on the deck's process grid, exact arithmetic), so the
solves follow perform_vcycle's trajectory.

The deck is the dictionary form of the reference's GEOM_INPUTS and
PLASMA_INPUTS (``ION_ELEC_DECK`` and ``H2_O2_DECK`` below). Fields are
float64 device tensors of shape (nz+2, ny+2, nx+2), the ghosted Fortran
layout; :meth:`PlasmaRun.fields` returns the physical fields under the
names of the reference's output files.
"""
from __future__ import annotations

from .chemistry import K_BMAN, etempmin_default, ion_elec
from .solver import FieldSolver

ERRTOLR = 1.0e-11        # mgridlib/globalvars.f90:29
NVCYCLES = 20            # the *_solve loops
FACES = ("left", "right", "bottom", "top", "back", "front")


# tests/ion_elec/{GEOM_INPUTS,PLASMA_INPUTS} of the reference as a deck
ION_ELEC_DECK = {
    "origin": (0.0, 0.0, 0.0), "length": (0.01, 0.01, 0.01), "n": (32, 32, 32), "p": (2, 4, 1),
    "timestep": 1.0e-11, "pressure": 133.0, "temperature": 300.0,
    "scales": {"LENGTH": 0.01, "TIME": 1.0e-11, "ELECDEN": 1.0e14, "IONDEN": 1.0e14, "SPECDEN": 1.0e15,
               "POT": 100.0, "ETEMP": 5800.0},
    "init": {"N": 3.21e22, "E": 1.0e14, "I+": 1.0e14, "POT": 0.0, "ELECTEMP": 50000.0},
    # face: codes and values of (electron density, species density, potential, electron temperature)
    "bcs": [("left", "ZGRD", "ZGRD", "ZGRD", "ZGRD", 0.0, 0.0, 0.0, 0.0),
            ("right", "ZGRD", "ZGRD", "ZGRD", "ZGRD", 0.0, 0.0, 0.0, 0.0),
            ("bottom", "FLUX", "FLUX", "DIRC", "FLUX", 0.0, 0.0, 100.0, 0.0),
            ("top", "FLUX", "FLUX", "DIRC", "FLUX", 0.0, 0.0, 0.0, 0.0),
            ("back", "ZGRD", "ZGRD", "ZGRD", "ZGRD", 0.0, 0.0, 0.0, 0.0),
            ("front", "ZGRD", "ZGRD", "ZGRD", "ZGRD", 0.0, 0.0, 0.0, 0.0)],
}


# tests/H2_O2/{GEOM_INPUTS,PLASMA_INPUTS} of the reference (the H2_O2_chem deck)
H2_O2_DECK = {
    "origin": (0.0, 0.0, 0.0), "length": (0.004, 0.004, 0.004), "n": (32, 32, 32), "p": (2, 4, 1),
    "timestep": 1.0e-11, "pressure": 13300.0, "temperature": 300.0,
    "scales": {"LENGTH": 0.004, "TIME": 1.0e-11, "ELECDEN": 1.0e14, "IONDEN": 1.0e14, "SPECDEN": 1.0e15,
               "POT": 1000.0, "ETEMP": 5800.0},
    "init": {"H2": 2.0e24, "O2": 1.0e24, "E": 1.0e14, "H+": 0.25e14, "H2+": 0.25e14, "O+": 0.25e14,
             "O2+": 0.5e14, "O-": 0.25e14, "H": 1.0e12, "O": 1.0e12, "POT": 0.0, "ELECTEMP": 5802.0},
    "bcs": [("left", "ZGRD", "ZGRD", "ZGRD", "ZGRD", 0.0, 0.0, 0.0, 0.0),
            ("right", "ZGRD", "ZGRD", "ZGRD", "ZGRD", 0.0, 0.0, 0.0, 0.0),
            ("bottom", "ZGRD", "ZGRD", "ZGRD", "ZGRD", 0.0, 0.0, 0.0, 0.0),
            ("top", "ZGRD", "ZGRD", "ZGRD", "ZGRD", 0.0, 0.0, 0.0, 0.0),
            ("back", "FLUX", "FLUX", "DIRC", "FLUX", 0.0, 0.0, 1750.0, 0.0),
            ("front", "FLUX", "FLUX", "DIRC", "FLUX", 0.0, 0.0, 0.0, 0.0)],
}


# tests/parallel_plate/{GEOM_INPUTS,PLASMA_INPUTS} of the reference (ion_elec
# chemistry): a 0.625 mm x 1 cm x 0.625 mm column, 32 x 512 x 32 cells on a
# 2 x 4 x 2 process grid, electrodes on the y faces at 0 V and -250 V
PARALLEL_PLATE_DECK = {
    "origin": (0.0, 0.0, 0.0), "length": (0.000625, 0.01, 0.000625), "n": (32, 512, 32), "p": (2, 4, 2),
    "timestep": 1.0e-10, "pressure": 133.0, "temperature": 300.0,
    "scales": {"LENGTH": 0.01, "TIME": 1.0e-11, "ELECDEN": 1.0e14, "IONDEN": 1.0e14, "SPECDEN": 1.0e15,
               "POT": 100.0, "ETEMP": 5800.0},
    "init": {"N": 3.21e22, "E": 1.0e14, "I+": 1.0e14, "POT": 0.0, "ELECTEMP": 1160.0},
    "bcs": [("left", "ZGRD", "ZGRD", "ZGRD", "ZGRD", 0.0, 0.0, 0.0, 0.0),
            ("right", "ZGRD", "ZGRD", "ZGRD", "ZGRD", 0.0, 0.0, 0.0, 0.0),
            ("bottom", "FLUX", "FLUX", "DIRC", "FLUX", 0.0, 0.0, 0.0, 0.0),
            ("top", "FLUX", "FLUX", "DIRC", "FLUX", 0.0, 0.0, -250.0, 0.0),
            ("back", "ZGRD", "ZGRD", "ZGRD", "ZGRD", 0.0, 0.0, 0.0, 0.0),
            ("front", "ZGRD", "ZGRD", "ZGRD", "ZGRD", 0.0, 0.0, 0.0, 0.0)],
}


def output_names(chem):
    """The reference's output scalars in its order (all_scalarnames,
    solve_manager.f90:270-289): species densities, potential, field, electron
    temperature and energy, electron and solved-neutral production."""
    names = ["%s_density (#/m3)" % nm for nm in chem.names] + ["Potential (V)"]
    names += ["Electric_field_%s (V/m)" % a for a in "XYZ"]
    names += ["Electron_Temp (K)", "Electron_energy (J/m3)", "E_prod (#/m3/s)"]
    names += ["%s_prod (#/m3/s)" % chem.names[sp] for sp in range(chem.neutralspecmin - 1, chem.neutralspecmax)]
    return names


def write_deck_output(dirname, fields, chem, deck):
    """writeoutputfile (solve_manager.f90:353-361) for physical fields
    {output name: (nx, ny, nz) array}: the reference's VTK files, one piece
    per rank of the deck's process grid, coordinates in units of the deck's
    length scale (scalelengths, solve_manager.f90:152-172)."""
    from . import vtk
    names = output_names(chem)
    return vtk.write_output(dirname, [fields[nm] for nm in names], names, tuple(deck["n"]), tuple(deck["length"]),
                            tuple(deck["p"]), origin=tuple(deck.get("origin", (0.0, 0.0, 0.0))),
                            length_scale=float(deck["scales"]["LENGTH"]))


class PlasmaRun:
    """arith "exact" (default): every solve is the reference's V-cycle loop in
    the reference's rounding (the coarse GMRES included), so the fields follow
    the reference run's trajectory. arith "fast": the same loops and
    operators with the rows of the equations without convection in
    fast arithmetic and the coarsest level solved by a device-built dense
    inverse (the reference's coarse GMRES with m = n and tolerance 1e-15 is
    an exact solve; mgridsteps.f90:342-355)."""

    def __init__(self, deck, chem=None, device="cuda", arith="exact"):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.chem = chem or ion_elec()
        ch = self.chem
        self.ions = list(range(ch.ionspecmin - 1, ch.ionspecmax))
        self.neutrals = list(range(ch.neutralspecmin - 1, ch.neutralspecmax))
        self.n = tuple(int(v) for v in deck["n"])
        self.p = tuple(int(v) for v in deck["p"])
        sc = deck["scales"]
        self.L0, self.t0 = float(sc["LENGTH"]), float(sc["TIME"])
        self.escale_n, self.iscale = float(sc["ELECDEN"]), float(sc["IONDEN"])
        self.nscale = float(sc["SPECDEN"])
        self.potscale, self.etempscale = float(sc["POT"]), float(sc["ETEMP"])
        # solversetup (solve_manager.f90:316)
        self.eescale = self.escale_n * K_BMAN * self.etempscale
        self.dt = float(deck["timestep"])
        self.Tg, self.Pg = float(deck["temperature"]), float(deck["pressure"])
        self.deck = dict(deck)
        length = tuple(float(L) / self.L0 for L in deck["length"])   # scalelengths
        self.length = length

        def ctx():
            s = FieldSolver(self.n, length, self.p, hierarchy="ref", precision="fp64", flux="upwind",
                            max_cycles=NVCYCLES, tol=ERRTOLR, tol_kind="rms", arith=arith)
            s.set_chemistry(ch)
            return s

        self.pois, self.ener, self.elec, self.ion = ctx(), ctx(), ctx(), ctx()
        self.neut = ctx() if self.neutrals else None
        # boundary codes and parameters per equation (readsolverinputfile)
        bcs = {b[0]: b[1:] for b in deck["bcs"]}
        self.codes = {eq: tuple(bcs[f][q] for f in FACES) for q, eq in enumerate(("eden", "spec", "pot", "etemp"))}
        self.params = {eq: tuple(float(bcs[f][4 + q]) for f in FACES)
                       for q, eq in enumerate(("eden", "spec", "pot", "etemp"))}
        # scalings of each module's *_initialize
        tl, tl2 = self.t0 / self.L0, self.t0 / (self.L0 * self.L0)
        self.scal = {"pot": (tl, tl2, self.t0, self.t0 / self.potscale),
                     "eden": (tl, tl2, self.t0, self.t0 / self.escale_n),
                     "ion": (tl, tl2, self.t0, self.t0 / self.iscale),
                     "neut": (tl, tl2, self.t0, self.t0 / self.nscale),
                     "ener": (tl, tl2, self.t0, self.t0 / self.eescale)}
        self.vel_scaling = tl
        self.nemin = 1.0e11 / self.escale_n                 # eden_initialize
        self.etempmin = etempmin_default()                  # eenergy_initialize
        # initial state (solversetup, solve_manager.f90:294-339)
        ini = deck["init"]
        g = tuple(v + 2 for v in self.n)[::-1]
        full = lambda v: torch.full(g, float(v), dtype=torch.float64, device=self.dev)  # noqa: E731
        self.numden = torch.stack([full(ini[nm]) for nm in ch.names])
        self.pot = full(ini["POT"])
        self.E = torch.zeros((3,) + g, dtype=torch.float64, device=self.dev)
        self.Te = full(ini["ELECTEMP"])
        e = ch.especnum - 1
        self.eps = 1.5 * self.numden[e] * K_BMAN * self.Te     # real*4 1.5 is exact
        self.eprod = torch.zeros(g, dtype=torch.float64, device=self.dev)
        self.nprod = [torch.zeros(g, dtype=torch.float64, device=self.dev) for _ in self.neutrals]
        # module solutions (scaled, ghosted): poisson_init, eden_init, ...
        # (IEEE divisions, as the reference divides: a tensor divided by a
        # Python number goes through its reciprocal)
        self.phi_s = self._div(self.pot, self.potscale)
        self.ne_s = self._div(self.numden[e], self.escale_n)
        self.ni_s = [self._div(self.numden[i], self.iscale) for i in self.ions]
        self.nn_s = [self._div(self.numden[i], self.nscale) for i in self.neutrals]
        self.eps_s = self._div(self.eps, self.eescale)
        # *_set_bcs: face values bc_params / scale on every face
        self.faces = {
            "pot": [self._face_full(f, self.params["pot"][q] / self.potscale) for q, f in enumerate(FACES)],
            "eden": [self._face_full(f, self.params["eden"][q] / self.escale_n) for q, f in enumerate(FACES)],
            "ion": [[self._face_full(f, self.params["spec"][q] / self.iscale) for q, f in enumerate(FACES)]
                    for _ in self.ions],
            "neut": [[self._face_full(f, self.params["spec"][q] / self.nscale) for q, f in enumerate(FACES)]
                     for _ in self.neutrals],
            "ener": [self._face_full(f, 0.0) for f in FACES],
        }
        self.t = 0.0
        self.steps = 0
        self.cycles = []

    def _div(self, t, v):
        return t / self.torch.full_like(t, float(v))

    # ------------------------------------------------------------ faces
    def _face_shape(self, f):
        nx, ny, nz = self.n
        return {"left": (nz, ny), "right": (nz, ny), "bottom": (nz, nx), "top": (nz, nx),
                "back": (ny, nx), "front": (ny, nx)}[f]

    def _face_full(self, f, v):
        return self.torch.full(self._face_shape(f), float(v), dtype=self.torch.float64, device=self.dev)

    def _bc(self, G, f):
        """Boundary-cell values of a ghosted (nz+2, ny+2, nx+2) tensor on face
        f, as the reference's face arrays ((ny,nz), (nx,nz), (nx,ny) in
        Fortran order)."""
        nx, ny, nz = self.n
        sl = {"left": (slice(1, -1), slice(1, -1), 1), "right": (slice(1, -1), slice(1, -1), nx),
              "bottom": (slice(1, -1), 1, slice(1, -1)), "top": (slice(1, -1), ny, slice(1, -1)),
              "back": (1, slice(1, -1), slice(1, -1)), "front": (nz, slice(1, -1), slice(1, -1))}[f]
        return G[sl]

    @staticmethod
    def _axis(f):
        return {"left": 0, "right": 0, "bottom": 1, "top": 1, "back": 2, "front": 2}[f]

    # ------------------------------------------------------------ one step
    def _solve(self, s, u, codes, faces, co, timederiv, dt_s, faces_changed=True, coeffs_changed=True):
        if faces_changed:
            s.set_bc(codes, [x.contiguous() for x in faces])
        if coeffs_changed:
            V = co["vel"].view(3, -1)
            s.set_coeffs(co["dcoeff"], V[0], V[1], V[2], co["reac"], co["source"], timederiv, dt_s)
        else:
            s.set_source(co["source"])
        k, _, _, _ = s.solve(u)
        s.fill_ghosts(u)
        return k

    def step(self):
        ch = self.chem
        e = ch.especnum - 1
        dt_s = self.dt / self.t0
        cyc = []
        # poisson_solve (poissoneq.f90:387-470); the potential's faces stay
        # (poisson_update_bcs changes only 'DZGD' electrodes)
        ions = self.numden[self.ions[0]:self.ions[-1] + 1].reshape(-1)
        co = self.pois.poisson_transport(self.numden[e].reshape(-1), ions, self.scal["pot"], timederiv=False, dt=dt_s)
        # (its coefficients never change: the permittivity-free operator
        # dcoeff = tscale/lscale^2 of poisson_update_transport; only the source)
        cyc.append(self._solve(self.pois, self.phi_s, self.codes["pot"], self.faces["pot"], co, False, dt_s,
                               faces_changed=self.steps == 0, coeffs_changed=self.steps == 0))
        self.pot = self.phi_s * self.potscale
        # update_efield (poissoneq.f90:565-680)
        self.pois.efield(self.phi_s, self.potscale / self.L0, False, out=self.E)
        # eenergy_solve (eenergyeq.f90:776-840) and update_electrontemp
        co = self.ener.eenergy_transport(self.E.reshape(-1), self.Te.reshape(-1), self.numden.reshape(-1),
                                         self.scal["ener"], lscale=self.L0, soln=self.eps_s.reshape(-1),
                                         dt=dt_s, Tg=self.Tg, Pg=self.Pg)
        self._eenergy_bcs()
        cyc.append(self._solve(self.ener, self.eps_s, self.codes["etemp"], self.faces["ener"], co, True, dt_s))
        self.ener.electron_temp(self.eps_s, self.numden[e], self.eescale, self.etempmin, self.Te)
        self.eps = self.eps_s * self.eescale
        # eden_solve (edeneq.f90:446-523)
        co = self.elec.species_transport(ch.especnum, self.E.reshape(-1), self.Te.reshape(-1),
                                         self.numden.reshape(-1), self.scal["eden"], soln=self.ne_s.reshape(-1),
                                         dt=dt_s, Tg=self.Tg, Pg=self.Pg, out={"prod": self.eprod.view(-1)})
        self._eden_bcs()
        cyc.append(self._solve(self.elec, self.ne_s, self.codes["eden"], self.faces["eden"], co, True, dt_s))
        self.elec.floor(self.ne_s, self.nemin)
        self.numden[e].copy_(self.ne_s * self.escale_n)
        # ions_solve (ionseq.f90:560-629): every ion from the same densities,
        # which take the new ion densities after the last one
        for q, sp in enumerate(self.ions):
            co = self.ion.species_transport(sp + 1, self.E.reshape(-1), self.Te.reshape(-1),
                                            self.numden.reshape(-1), self.scal["ion"],
                                            soln=self.ni_s[q].reshape(-1), dt=dt_s, Tg=self.Tg, Pg=self.Pg)
            self._ion_bcs(q, sp)
            cyc.append(self._solve(self.ion, self.ni_s[q], self.codes["spec"], self.faces["ion"][q], co, True, dt_s))
        for q, sp in enumerate(self.ions):
            self.numden[sp].copy_(self.ni_s[q] * self.iscale)
        # neutrals_solve (neutralseq.f90:438-508), after it the same update
        for q, sp in enumerate(self.neutrals):
            co = self.neut.species_transport(sp + 1, self.E.reshape(-1), self.Te.reshape(-1),
                                             self.numden.reshape(-1), self.scal["neut"],
                                             soln=self.nn_s[q].reshape(-1), dt=dt_s, Tg=self.Tg, Pg=self.Pg,
                                             out={"prod": self.nprod[q].view(-1)})
            # neutrals_update_transport leaves vel at zero
            co["vel"].zero_()
            self._neutral_bcs(q, sp)
            cyc.append(self._solve(self.neut, self.nn_s[q], self.codes["spec"], self.faces["neut"][q], co, True,
                                   dt_s))
        for q, sp in enumerate(self.neutrals):
            self.numden[sp].copy_(self.nn_s[q] * self.nscale)
        self.t = self.t + self.dt
        self.steps += 1
        self.cycles.append(tuple(cyc))

    def run(self, nsteps):
        for _ in range(nsteps):
            self.step()
        return self

    # ------------------------------------------------------------ BC updaters
    def _eden_bcs(self):
        """eden_update_bcs (edeneq.f90:203-358): FLUX faces get the thermal
        outflow 0.25 n cbar, cbar = sqrt(8 k_B T_e / (pi m_e)), in one launch
        (mps_eden_faces)."""
        self.elec.eden_faces(self.codes["eden"], self.ne_s.reshape(-1), self.Te.reshape(-1), self.vel_scaling,
                             self.faces["eden"])

    def _eenergy_bcs(self):
        """eenergy_update_bcs (eenergyeq.f90:208-392): DIRC faces 3/2 k_B n_e
        T_wall, FLUX faces 4/3 n_eps cbar, in one launch (mps_eenergy_faces)."""
        ne = self.numden[self.chem.especnum - 1]
        self.ener.eenergy_faces(self.codes["etemp"], self.eps_s.reshape(-1), ne.reshape(-1), self.eescale,
                                self.params["etemp"], self.vel_scaling, self.faces["ener"])

    def _ion_bcs(self, q, sp):
        """ions_update_bcs (ionseq.f90:206-476) for ion sp: FLUX faces carry
        the drift outflow mu n E_normal where it leaves the domain, else
        zero."""
        faces = self.faces["ion"][q]
        flux = [None] * 6
        for k, f in enumerate(FACES):
            code = self.codes["spec"][k]
            if code == "DIRC":
                # the reference divides by ionscale on every call (ionseq.f90:243)
                faces[k] = self._div(faces[k], self.iscale)
            elif code == "FLUX":
                flux[k] = faces[k]    # rewritten in place: mps_set_bc copies the faces
        if any(t is not None for t in flux):
            # all FLUX faces of this ion in one launch (mps_ion_drift_faces)
            self.ion.ion_drift_faces(sp + 1, self.E.reshape(-1), self.numden.reshape(-1),
                                     self.ni_s[q].reshape(-1), self.vel_scaling, flux)

    def _neutral_bcs(self, q, sp):
        """neutrals_update_bcs (neutralseq.f90:206-358): FLUX faces get the
        thermal outflow 0.25 n cbar, cbar = sqrt(8 k_B T_g / (pi m)), in one
        launch (mps_neutral_faces)."""
        self.neut.neutral_faces(self.codes["spec"], self.nn_s[q].reshape(-1), float(self.chem.molmass[sp]), self.Tg,
                                self.vel_scaling, self.faces["neut"][q])

    # ------------------------------------------------------------ output
    def fields(self):
        """{reference output name: physical (nx, ny, nz) numpy array}."""
        ch = self.chem

        def host(G):
            return G[1:-1, 1:-1, 1:-1].detach().cpu().numpy().T

        out = {"%s_density (#/m3)" % nm: host(self.numden[s]) for s, nm in enumerate(ch.names)}
        out["Potential (V)"] = host(self.pot)
        for q, a in enumerate("XYZ"):
            out["Electric_field_%s (V/m)" % a] = host(self.E[q])
        out["Electron_Temp (K)"] = host(self.Te)
        out["Electron_energy (J/m3)"] = host(self.eps)
        out["E_prod (#/m3/s)"] = host(self.eprod)
        for q, sp in enumerate(self.neutrals):
            out["%s_prod (#/m3/s)" % ch.names[sp]] = host(self.nprod[q])
        return out

    def output_names(self):
        return output_names(self.chem)

    def write_output(self, dirname):
        """writeoutputfile: the current fields in the reference's VTK files
        (write_deck_output)."""
        return write_deck_output(dirname, self.fields(), self.chem, self.deck)

    def close(self):
        for s in (self.pois, self.ener, self.elec, self.ion, self.neut):
            if s is not None:
                s.close()
