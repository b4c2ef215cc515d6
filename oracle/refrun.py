"""Runner for the reference Fortran/MPI drivers built by ``make -C oracle ref``.

Test infrastructure only: used by ``tests/golden/make_golden.py`` to write
golden fixtures and by ``bench.py``'s CPU-baseline leg when the prebuilt
binary and an MPI launcher are present. Nothing in the product path imports
this module.

The binary file layout written here is the one read by
``oracle/ref_driver/ref_driver.f90``.
"""
from __future__ import annotations

import os
import shutil
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
REF_SG1D = os.path.join(HERE, "_ref", "ref_sg1d")
MAGIC = 861098061  # b"MPS3" as little-endian int32

MODE_APPLY, MODE_SMOOTH, MODE_SOLVE, MODE_GMRES, MODE_TRANSFER, MODE_GHOST = 1, 2, 3, 4, 5, 6


def mpiexec_path():
    for cand in ("/opt/conda/bin/mpiexec", shutil.which("mpiexec") or ""):
        if cand and os.path.exists(cand):
            return cand
    return None


def available():
    return os.path.exists(REF_DRIVER) and mpiexec_path() is not None


def _f(a, shape):
    a = np.zeros(shape) if a is None else np.broadcast_to(np.asarray(a, dtype=np.float64), shape)
    return np.asfortranarray(a).tobytes(order="F")


def write_input(path, case):
    nx, ny, nz = case["n"]
    px, py, pz = case.get("p", (1, 1, 1))
    hdr = [MAGIC, case["mode"], nx, ny, nz, px, py, pz,
           int(bool(case.get("timederiv", False))), case.get("nvcycles", 20),
           case.get("niter", 2), case.get("maxkspdim", 1), case.get("nrestarts", 1), 0, 0, 0]
    L = case.get("len", (1.0, 1.0, 1.0))
    rhdr = [L[0], L[1], L[2], case.get("dt", 1.0), case.get("tol", 0.0)]
    codes = "".join(c.ljust(4)[:4] for c in case["bc_codes"])
    faces = case.get("bc_vals", [0.0] * 6)
    fshapes = [(ny, nz), (ny, nz), (nx, nz), (nx, nz), (nx, ny), (nx, ny)]
    with open(path, "wb") as fh:
        fh.write(struct.pack("<16i", *hdr))
        fh.write(struct.pack("<5d", *rhdr))
        fh.write(codes.encode("ascii"))
        for v, s in zip(faces, fshapes):
            fh.write(_f(v, s))
        g = (nx, ny, nz)
        for key in ("u0", "dcoeff", "velx", "vely", "velz", "reac", "source", "sterm", "b"):
            fh.write(_f(case.get(key), g))
        if case["mode"] == MODE_TRANSFER:
            fh.write(_f(case["extra"], (nx // 2 + 2, ny // 2 + 2, nz // 2 + 2)))
        elif case["mode"] == MODE_GHOST:
            fh.write(_f(case["extra"], (nx + 2, ny + 2, nz + 2)))


def read_output(path):
    out = {}
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0
    while pos < len(data):
        name = data[pos:pos + 8].decode("ascii").strip()
        (n,) = struct.unpack("<q", data[pos + 8:pos + 16])
        pos += 16
        out[name] = np.frombuffer(data[pos:pos + 8 * n], dtype="<f8").copy()
        pos += 8 * n
    return out


MODE_VTK = 7


def run_vtk(case, timeout=120):
    """writeoutput of the reference (mode 7) on ``case`` (u0 and dcoeff as
    the two scalars): returns (pvtr bytes, [vtr bytes of rank 0, 1, ...])."""
    px, py, pz = case.get("p", (1, 1, 1))
    nranks = px * py * pz
    case = dict(case, mode=MODE_VTK)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_input(fin, case)
        cmd = [mpiexec_path(), "-n", str(nranks), REF_DRIVER, fin, fout]
        subprocess.run(["bash", "-c", "ulimit -s unlimited && exec \"$@\"", "_"] + cmd, check=True,
                       timeout=timeout, cwd=td, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
        with open(os.path.join(td, "vtkout.pvtr"), "rb") as fh:
            pvtr = fh.read()
        vtr = []
        for r in range(nranks):
            with open(os.path.join(td, "vtkout", "file_%08d.vtr" % r), "rb") as fh:
                vtr.append(fh.read())
    return pvtr, vtr


def run_chem(name, Tg, Pg, states, timeout=60):
    """The reference chem_module ``name`` ("H2_O2", "ion_elec"; built by
    ``make -C oracle chem``) on sample states [(Te, efield_mag,
    specden[nspecies]), ...]: its tables and, per state, every reaction's
    rate and every species' production, mobility, diffusion coefficient,
    the inelastic term and the electron collision frequency."""
    exe = os.path.join(HERE, "_ref", "ref_chem_" + name)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as fh:
            fh.write(struct.pack("<i", len(states)))
            fh.write(struct.pack("<2d", Tg, Pg))
            for Te, ef, sd in states:
                fh.write(struct.pack("<2d", Te, ef))
                fh.write(np.asarray(sd, dtype="<f8").tobytes())
        subprocess.run([exe, fin, fout], check=True, timeout=timeout, stdout=subprocess.DEVNULL,
                       stderr=subprocess.PIPE)
        with open(fout, "rb") as fh:
            raw = fh.read()
    ns, nr, e, bg, imin, imax = struct.unpack("<6i", raw[:24])
    pos = 24

    def take(dtype, count, shape=None):
        nonlocal pos
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=pos).copy()
        pos += a.nbytes
        return a.reshape(shape, order="F") if shape else a

    out = {"nspecies": ns, "nreac": nr, "especnum": e, "bgspecnum": bg, "ionspecmin": imin, "ionspecmax": imax}
    out["reactants"] = take("<i4", nr * ns, (nr, ns))
    out["products"] = take("<i4", nr * ns, (nr, ns))
    out["isratearrh"] = take("<i4", nr)
    out["k_arrh"] = take("<f8", 3 * nr, (3, nr)).T.copy()
    out["elecenergy"] = take("<f8", nr)
    out["molmass"] = take("<f8", ns)
    out["charge"] = take("<f8", ns)
    per = {k: [] for k in ("rate", "prod", "mob", "dco", "inel", "collfreq")}
    for _ in states:
        per["rate"].append(take("<f8", nr))
        per["prod"].append(take("<f8", ns))
        per["mob"].append(take("<f8", ns))
        per["dco"].append(take("<f8", ns))
        per["inel"].append(take("<f8", 1)[0])
        per["collfreq"].append(take("<f8", 1)[0])
    out.update({k: np.array(v) for k, v in per.items()})
    return out


REF_PLASMA = os.path.join(HERE, "_ref", "mps3d_ion_elec")

from mps3d_amd.plasma import ION_ELEC_DECK  # noqa: E402,F401  (the deck layout)


def write_deck(td, deck, nsteps):
    """GEOM_INPUTS and PLASMA_INPUTS in the layout the reference reads
    (plasma_pde/inputs.f90:6-45, plasma_pde/solve_manager.f90:58-150), for a
    run of exactly nsteps steps (finaltime between nsteps-1 and nsteps
    steps) whose only outputs are output_0000000 and output_final."""
    sep = "=" * 70
    o, L, n, p = deck["origin"], deck["length"], deck["n"], deck["p"]
    geom = [sep, "Geometry", sep] + ["ORIGIN_%s %r" % (a, v) for a, v in zip("XYZ", o)] + \
        ["LENGTH_%s %r" % (a, v) for a, v in zip("XYZ", L)] + [sep, "Grid_and_processors", sep] + \
        ["CELLS_%s %d" % (a, v) for a, v in zip("XYZ", n)] + ["PROC_%s %d" % (a, v) for a, v in zip("XYZ", p)] + [sep]
    dt = deck["timestep"]
    sc, ini = deck["scales"], deck["init"]
    plasma = [sep, "Time information", sep, "timestep %r" % dt, "finaltime %r" % ((nsteps - 0.5) * dt),
              "maxtimesteps 1000000", "outputstep 1000000", "stdoutstep 1000000",
              sep, "Background gas parameters", sep, "PRESSURE %r" % deck["pressure"],
              "TEMPERATURE %r" % deck["temperature"], sep, "Scales", sep] + \
        ["%s %r" % (k, sc[k]) for k in ("LENGTH", "TIME", "ELECDEN", "IONDEN", "SPECDEN", "POT", "ETEMP")] + \
        [sep, "Initial conditions (species,potential,electemp)", sep] + \
        ["%s %r" % (k, v) for k, v in ini.items()] + \
        [sep, "Boundary_conditions(elecden,specden,potential,electemp)", sep] + \
        ["%s %s %s %s %s %r %r %r %r" % b for b in deck["bcs"]] + \
        [sep, "Problem_specific_parameters", sep, " ".join(["0.0"] * 20)]
    with open(os.path.join(td, "GEOM_INPUTS"), "w") as fh:
        fh.write("\n".join(geom) + "\n")
    with open(os.path.join(td, "PLASMA_INPUTS"), "w") as fh:
        fh.write("\n".join(plasma) + "\n")


def _gather(td, name, n, p):
    """Global (nx, ny, nz) fields of output <name> from its per-rank .vtr files."""
    from mps3d_amd import vtk
    px, py, pz = p
    l = (n[0] // px, n[1] // py, n[2] // pz)
    out = {}
    for r in range(px * py * pz):
        pi = (r % px, (r // px) % py, r // (px * py))
        d = vtk.read_vtr(os.path.join(td, name, "file_%08d.vtr" % r))
        for k, v in d.items():
            if k in ("X", "Y", "Z"):
                continue
            a = out.setdefault(k, np.zeros(n, order="F"))
            a[pi[0] * l[0]:(pi[0] + 1) * l[0], pi[1] * l[1]:(pi[1] + 1) * l[1], pi[2] * l[2]:(pi[2] + 1) * l[2]] = v
    return out


def run_plasma(deck, nsteps, timeout=600, chem="ion_elec", raw=False):
    """The reference plasma solver (oracle/_ref/mps3d_ion_elec) on ``deck``
    for nsteps time steps: (initial fields, fields after nsteps steps), each
    {scalar name: global (nx, ny, nz) array} as its VTK output holds them."""
    p = deck["p"]
    with tempfile.TemporaryDirectory() as td:
        write_deck(td, deck, nsteps)
        cmd = [mpiexec_path(), "-n", str(int(np.prod(p))), os.path.join(HERE, "_ref", "mps3d_" + chem)]
        subprocess.run(["bash", "-c", "ulimit -s unlimited && exec \"$@\"", "_"] + cmd, check=True,
                       timeout=timeout, cwd=td, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
        res = (_gather(td, "output_0000000", deck["n"], p), _gather(td, "output_final", deck["n"], p))
        if raw:
            files = {"output_final.pvtr": open(os.path.join(td, "output_final.pvtr"), "rb").read()}
            for r in range(int(np.prod(p))):
                name = "output_final/file_%08d.vtr" % r
                files[name] = open(os.path.join(td, name), "rb").read()
            res = res + (files,)
        return res


def run(case, timeout=600):
    """Run the reference driver on ``case``; returns a dict of float64 arrays.

    3D arrays come back in Fortran order, reshaped to (nx, ny, nz)."""
    nx, ny, nz = case["n"]
    px, py, pz = case.get("p", (1, 1, 1))
    nranks = px * py * pz
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_input(fin, case)
        cmd = [mpiexec_path(), "-n", str(nranks), REF_DRIVER, fin, fout]
        env = dict(os.environ)
        subprocess.run(["bash", "-c", "ulimit -s unlimited && exec \"$@\"", "_"] + cmd,
                       check=True, timeout=timeout, env=env,
                       stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
        out = read_output(fout)
    full = nx * ny * nz
    for k, v in list(out.items()):
        if v.size == full:
            out[k] = v.reshape((nx, ny, nz), order="F")
    return out


def run_sg1d(rows, timeout=60):
    rows = np.ascontiguousarray(np.asarray(rows, dtype=np.float64).reshape(-1, 7))
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as fh:
            fh.write(struct.pack("<i", rows.shape[0]))
            fh.write(rows.tobytes())
        subprocess.run([REF_SG1D, fin, fout], check=True, timeout=timeout,
                       stdout=subprocess.DEVNULL)
        return np.fromfile(fout, dtype="<f8").reshape(-1, 3)
