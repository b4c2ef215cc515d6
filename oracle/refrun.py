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
