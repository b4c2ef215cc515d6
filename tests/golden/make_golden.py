"""Writes the golden fixtures in tests/golden/ from the reference itself.

Runs the reference Fortran/MPI library (built from the reference sources by
``make -C oracle ref`` into oracle/_ref/, driven by our own
oracle/ref_driver/*.f90) on seeded inputs and stores inputs and outputs as
compressed .npz files. Run it in the build container (where the reference
tree and MPICH are present):

    make -C oracle ref plasma chem && python tests/golden/make_golden.py [vtk] [plasma] [plasma32] [parallel_plate] [chem]

(``vtk`` / ``plasma`` / ``plasma32`` / ``parallel_plate`` / ``chem`` write only those fixtures.)

The fixtures are data only: inputs, expected outputs and residual histories.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import refrun  # noqa: E402
from mps3d_amd import workloads as W  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def case_dict(c: W.Case, **kw):
    d = dict(n=c.n, len=c.length, bc_codes=list(c.bc_codes), bc_vals=c.bc_vals,
             dcoeff=c.dcoeff, velx=c.velx, vely=c.vely, velz=c.velz, reac=c.reac,
             source=c.source, timederiv=c.timederiv, dt=c.dt)
    d.update(kw)
    return d


def inputs_of(c: W.Case):
    z = np.zeros(c.n, order="F")
    d = {"n": np.array(c.n), "length": np.array(c.length),
         "bc_codes": np.array(["%-4s" % s for s in c.bc_codes]),
         "timederiv": np.array(int(c.timederiv)), "dt": np.array(c.dt)}
    for i, f in enumerate(c.bc_vals):
        d["bc%d" % i] = f
    for k in ("dcoeff", "velx", "vely", "velz", "reac", "source"):
        v = getattr(c, k)
        d[k] = z if v is None else v
    return d


def save(name, **arrays):
    path = os.path.join(OUT, name + ".npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


def vtk_fixture():
    """writeoutput (mgridlib/cartwrite.f90:455-481) on an 8x6x4 grid over a
    2x1x2 process grid, a 6x4x10 grid over 1x2x5 and a 5x3x2 grid on one
    rank, two seeded cell scalars each: the .pvtr and .vtr files' bytes."""
    rng = np.random.default_rng(20261203)
    out = {}
    for tag, n, L, p in (("a", (8, 6, 4), (0.8, 0.3, 0.4), (2, 1, 2)),
                         ("b", (6, 4, 10), (0.01, 0.01, 0.01), (1, 2, 5)),
                         ("c", (5, 3, 2), (1.0, 0.7, 0.3), (1, 1, 1))):
        u0 = rng.standard_normal(n)
        d = rng.uniform(1e14, 1e16, n)
        pv, vt = refrun.run_vtk(dict(n=n, len=L, p=p, bc_codes=["ZGRD"] * 6, u0=u0, dcoeff=d))
        out[tag + "_n"] = np.array(n)
        out[tag + "_len"] = np.array(L)
        out[tag + "_p"] = np.array(p)
        out[tag + "_u0"] = u0
        out[tag + "_d"] = d
        out[tag + "_pvtr"] = np.frombuffer(pv, dtype=np.uint8)
        for r, v in enumerate(vt):
            out["%s_vtr%d" % (tag, r)] = np.frombuffer(v, dtype=np.uint8)
    save("vtk_output", **out)


CHEM_STATES = 64


def chem_fixture():
    """The reference's chem_modules (H2_O2_chem.f90, ion_elec_chem.f90,
    driven by oracle/ref_driver/ref_chem.f90; O2_chem.f90 and H2_chem.f90,
    driven by oracle/ref_driver/ref_chem4.f90) on 64 seeded states each:
    their tables, and per state every reaction rate, every species'
    production, mobility and diffusion coefficient, the inelastic term and
    the electron collision frequency."""
    rng = np.random.default_rng(20261204)
    out = {}
    for name, ns, Pg in (("H2_O2", 10, 13332.2), ("ion_elec", 3, 133.0), ("O2", 6, 133.0), ("H2", 5, 133.0)):
        Tg = 300.0
        lo = {"H2_O2": [21, 21, 14, 12, 12, 12, 12, 12, 17, 17], "ion_elec": [21, 14, 14],
              "O2": [21, 14, 12, 14, 12, 17], "H2": [21, 14, 12, 14, 17]}[name]
        hi = {"H2_O2": [23, 23, 17, 16, 16, 16, 16, 16, 21, 21], "ion_elec": [23, 17, 17],
              "O2": [23, 17, 16, 17, 16, 21], "H2": [23, 17, 16, 17, 21]}[name]
        Te = 10.0 ** rng.uniform(np.log10(3000.0), np.log10(1.0e5), CHEM_STATES)
        ef = 10.0 ** rng.uniform(2, 6, CHEM_STATES)
        sd = 10.0 ** rng.uniform(lo, hi, (CHEM_STATES, ns))
        d = refrun.run_chem(name, Tg, Pg, list(zip(Te, ef, sd)))
        for k, v in d.items():
            out["%s_%s" % (name, k)] = np.asarray(v)
        out[name + "_Te"], out[name + "_ef"], out[name + "_sd"] = Te, ef, sd
        out[name + "_TgPg"] = np.array([Tg, Pg])
    save("chem_reference", **out)


PLASMA_STEPS = (1, 5)


def plasma_fixture(chem="ion_elec"):
    """The reference plasma solver (plasma_pde/main.f90 with the ion_elec or
    the H2_O2 chemistry) on its tests/ion_elec or tests/H2_O2 deck at 16^3
    cells over a 2x2x1 process grid: the fields of its VTK output after 1 and
    after 5 time steps."""
    from mps3d_amd.plasma import H2_O2_DECK, ION_ELEC_DECK
    deck = dict(ION_ELEC_DECK if chem == "ion_elec" else H2_O2_DECK)
    deck.update(n=(16, 16, 16), p=(2, 2, 1))
    out = {"n": np.array(deck["n"]), "p": np.array(deck["p"])}
    names = None
    for k in PLASMA_STEPS:
        ini, fin = refrun.run_plasma(deck, k, chem=chem)
        names = sorted(fin)
        for i, nm in enumerate(names):
            out["s%d_f%d" % (k, i)] = fin[nm]
    out["names"] = np.array(names)
    out["steps"] = np.array(PLASMA_STEPS)
    save("plasma_ion_elec16" if chem == "ion_elec" else "plasma_h2_o2_16", **out)


def plasma_deck_fixture(chem="ion_elec"):
    """The reference plasma solver on its own deck (tests/ion_elec or
    tests/H2_O2: 32^3 cells on the 2x4x1 process grid) after 1 and 5 time
    steps. Per field of its VTK output: the three lines through the centre
    (x at j = k = 16, y at i = k = 16, z at i = j = 16, 1-based), the xz
    plane at j = 16 and the field's sum, sum of squares, min and max."""
    from mps3d_amd.plasma import H2_O2_DECK, ION_ELEC_DECK
    deck = dict(ION_ELEC_DECK if chem == "ion_elec" else H2_O2_DECK)
    out = {"n": np.array(deck["n"]), "p": np.array(deck["p"])}
    names = None
    for k in PLASMA_STEPS:
        _, fin = refrun.run_plasma(deck, k, chem=chem)
        names = sorted(fin)
        for i, nm in enumerate(names):
            F = fin[nm]
            h = F.shape[0] // 2 - 1
            out["s%d_f%d_x" % (k, i)] = F[:, h, h]
            out["s%d_f%d_y" % (k, i)] = F[h, :, h]
            out["s%d_f%d_z" % (k, i)] = F[h, h, :]
            out["s%d_f%d_xz" % (k, i)] = F[:, h, :]
            out["s%d_f%d_stats" % (k, i)] = np.array([F.sum(), (F * F).sum(), F.min(), F.max()])
    out["names"] = np.array(names)
    out["steps"] = np.array(PLASMA_STEPS)
    save("plasma_ion_elec32_p241" if chem == "ion_elec" else "plasma_h2_o2_32_p241", **out)


def parallel_plate_fixture():
    """The reference plasma solver on its tests/parallel_plate deck (ion_elec
    chemistry, 32 x 512 x 32 cells on the 2x4x2 process grid) after 1 and 5
    time steps. Per field of its VTK output: the lines through the centre in
    x, y and z (i = k = 16, j = 256, 1-based), four more y lines (i, k in
    {4, 28}), and the field's sum, sum of squares, min and max."""
    from mps3d_amd.plasma import PARALLEL_PLATE_DECK
    deck = dict(PARALLEL_PLATE_DECK)
    out = {"n": np.array(deck["n"]), "p": np.array(deck["p"])}
    names = None
    for k in PLASMA_STEPS:
        _, fin = refrun.run_plasma(deck, k, chem="ion_elec", timeout=3600)
        names = sorted(fin)
        for i, nm in enumerate(names):
            F = fin[nm]
            key = "s%d_f%d" % (k, i)
            out[key + "_x"] = F[:, 255, 15]
            out[key + "_y"] = F[15, :, 15]
            out[key + "_z"] = F[15, 255, :]
            out[key + "_ylines"] = np.stack([F[a, :, b] for a in (3, 27) for b in (3, 27)])
            out[key + "_stats"] = np.array([F.sum(), (F * F).sum(), F.min(), F.max()])
    out["names"] = np.array(names)
    out["steps"] = np.array(PLASMA_STEPS)
    save("plasma_parallel_plate", **out)


def plasma_output_fixture():
    """The bytes of the reference plasma solver's VTK output (output_final:
    the .pvtr and the four ranks' .vtr) after one step of each deck at 16^3
    on 2x2x1 ranks."""
    from mps3d_amd.plasma import H2_O2_DECK, ION_ELEC_DECK
    out = {}
    for chem, base in (("ion_elec", ION_ELEC_DECK), ("H2_O2", H2_O2_DECK)):
        deck = dict(base)
        deck.update(n=(16, 16, 16), p=(2, 2, 1))
        _, _, files = refrun.run_plasma(deck, 1, chem=chem, raw=True)
        for name, data in files.items():
            key = "%s__%s" % (chem, name.replace("/", "__").replace(".", "_"))
            out[key] = np.frombuffer(data, dtype=np.uint8)
    save("plasma_output_files", **out)


def main(argv=()):
    assert refrun.available(), "build the reference first: make -C oracle ref"
    if set(argv) & {"vtk", "plasma", "chem", "plasma32", "parallel_plate"}:
        if "vtk" in argv:
            vtk_fixture()
        if "chem" in argv:
            chem_fixture()
        if "plasma" in argv:
            plasma_fixture("ion_elec")
            plasma_fixture("H2_O2")
            plasma_output_fixture()
        if "plasma32" in argv:
            plasma_deck_fixture("ion_elec")
            plasma_deck_fixture("H2_O2")
        if "parallel_plate" in argv:
            parallel_plate_fixture()
        return
    timings = {}

    # --- operator, RHS and smoother on a mixed-BC random problem -------------
    c = W.random_cdr((12, 10, 8), seed=11)
    rng = np.random.default_rng(12)
    u0 = rng.standard_normal(c.n)
    st = rng.standard_normal(c.n)
    o = refrun.run(case_dict(c, mode=refrun.MODE_APPLY, u0=u0, sterm=st))
    s = refrun.run(case_dict(c, mode=refrun.MODE_SMOOTH, u0=u0, sterm=st, niter=2))
    save("operator_mixed", u0=u0, sterm=st, AX=o["AX"], DIAG=o["DIAG"], OFFDIAG=o["OFFDIAG"],
         BVEC=o["BVEC"], BVECNOBC=o["BVECNOBC"], SMOOTH_U=s["U"], SMOOTH_RES=s["RES"], **inputs_of(c))

    # Poisson-only variant (no convection, reaction or time term)
    c2 = W.random_cdr((9, 7, 6), seed=21, codes=("DIRC", "DIRC", "ZGRD", "FLUX", "DIRC", "ZGRD"),
                      timederiv=False)
    c2.velx = c2.vely = c2.velz = c2.reac = None
    o = refrun.run(case_dict(c2, mode=refrun.MODE_APPLY, u0=u0[:9, :7, :6], sterm=st[:9, :7, :6]))
    save("operator_poisson", u0=u0[:9, :7, :6], sterm=st[:9, :7, :6], AX=o["AX"], DIAG=o["DIAG"],
         OFFDIAG=o["OFFDIAG"], BVEC=o["BVEC"], **inputs_of(c2))

    # --- grid transfers and ghost rules -------------------------------------
    ct = W.random_cdr((8, 6, 4), seed=31, codes=("DIRC", "ZGRD", "FLUX", "DIRC", "ZGRD", "DIRC"))
    fine = rng.standard_normal(ct.n)
    coarse_g = rng.standard_normal((4 + 2, 3 + 2, 2 + 2))
    o = refrun.run(case_dict(ct, mode=refrun.MODE_TRANSFER, u0=fine, extra=coarse_g))
    ghost_in = rng.standard_normal((8 + 2, 6 + 2, 4 + 2))
    g = refrun.run(case_dict(ct, mode=refrun.MODE_GHOST, extra=ghost_in))
    save("transfer", fine=fine, coarse_g=coarse_g, RESTR=o["RESTR"].reshape((4, 3, 2), order="F"),
         PROLONG=o["PROLONG"], BCLEFT=o["BCLEFT"], BCRIGHT=o["BCRIGHT"], BCBOTTOM=o["BCBOTTOM"],
         BCTOP=o["BCTOP"], BCBACK=o["BCBACK"], BCFRONT=o["BCFRONT"], ghost_in=ghost_in,
         GHOSTED=g["GHOSTED"].reshape((10, 8, 6), order="F"), **inputs_of(ct))

    # --- restarted GMRES with the GS preconditioner --------------------------
    cg = W.random_cdr((4, 4, 4), seed=41)
    cg.dcoeff = cg.dcoeff + 2.0  # keep the one-sweep GS preconditioner well defined
    b = rng.standard_normal(cg.n)
    x0 = rng.standard_normal(cg.n)
    runs = {}
    for tag, (m, nr, tol) in {"full": (64, 1, 1e-15), "restarted": (6, 8, 1e-12)}.items():
        o = refrun.run(case_dict(cg, mode=refrun.MODE_GMRES, u0=x0, b=b, maxkspdim=m,
                                 nrestarts=nr, tol=tol))
        runs["X_" + tag] = o["X"]
        runs["cfg_" + tag] = np.array([m, nr, tol])
    save("gmres_cdr", b=b, x0=x0, **runs, **inputs_of(cg))

    # --- V-cycle solve trajectories ------------------------------------------
    def solve_fixture(name, c, p, ncyc, tol, full=True):
        t = time.time()
        o = refrun.run(case_dict(c, mode=refrun.MODE_SOLVE, p=p, nvcycles=ncyc, tol=tol), timeout=3600)
        timings[name] = {"ranks": int(np.prod(p)), "solve_s": float(o["TIME"][0]),
                         "wall_s": time.time() - t, "n": list(c.n),
                         "cycles": int(np.sum(o["HIST"] >= 0))}
        hist = o["HIST"][o["HIST"] >= 0]
        U = o["U"]
        nx, ny, nz = c.n
        extra = {"U": U} if full else {
            "line_x": U[:, ny // 2, nz // 2], "line_y": U[nx // 2, :, nz // 2], "line_z": U[nx // 2, ny // 2, :],
            "stats": np.array([U.sum(), (U * U).sum(), U.min(), U.max()])}
        params = {}
        if "blobs" in c.params:
            params["blobs"] = c.params["blobs"]
        save(name, hist=hist, p=np.array(p), tol=np.array(tol), ncyc=np.array(ncyc),
             kind=np.array(c.params.get("kind", "")), **params, **extra,
             **({} if not full else inputs_of(c)))

    tol_generic = 1e-11 * float(np.float32(0.001))  # generic_pde/pdedefs.f90:368
    solve_fixture("solve_laplace16_p1", W.laplace_deck(16), (1, 1, 1), 20, tol_generic)
    solve_fixture("solve_laplace16_p8", W.laplace_deck(16), (2, 2, 2), 20, tol_generic)
    vc16 = W.varcoef_poisson(16)
    solve_fixture("solve_varcoef16_p1", vc16, (1, 1, 1), 20, 0.0)
    solve_fixture("solve_varcoef16_p8", vc16, (2, 2, 2), 20, 0.0)
    solve_fixture("solve_mms20x12x8_p512", W.mms_cdr((20, 12, 8)), (5, 1, 2), 10, 0.0)
    solve_fixture("solve_mms16x8x8_p1", W.mms_cdr((16, 8, 8)), (1, 1, 1), 10, 0.0)
    solve_fixture("solve_laplace64_p1", W.laplace_deck(64), (1, 1, 1), 20, tol_generic, full=False)
    solve_fixture("solve_laplace64_p8", W.laplace_deck(64), (2, 2, 2), 20, tol_generic, full=False)
    solve_fixture("solve_varcoef64_p8", W.varcoef_poisson(64), (2, 2, 2), 20, 0.0, full=False)
    solve_fixture("solve_varcoef128_p8", W.varcoef_poisson(128), (2, 2, 2), 20, 0.0, full=False)

    # --- 1D Scharfetter-Gummel flux ------------------------------------------
    rows = []
    for cL in (-3.0, -1e-3, -1e-9, 0.0, 2e-11, 1e-6, 0.7, 50.0):
        for cR in (-2.0, 0.0, 1e-3, 4.0):
            for DL, DR in ((1.0, 1.0), (0.3, 2.0), (1e-12, 1e-12), (5.0, 1e-3)):
                rows.append([cL, cR, DL, DR, 1.3, -0.4, 0.05])
    rows = np.array(rows)
    save("sg1d", rows=rows, out=refrun.run_sg1d(rows))

    vtk_fixture()
    plasma_fixture("ion_elec")
    plasma_fixture("H2_O2")
    plasma_output_fixture()
    plasma_deck_fixture("ion_elec")
    plasma_deck_fixture("H2_O2")
    parallel_plate_fixture()
    chem_fixture()

    with open(os.path.join(OUT, "ref_timings.json"), "w") as fh:
        json.dump({"host": os.uname().nodename, "nproc": os.cpu_count(), "runs": timings}, fh, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])
