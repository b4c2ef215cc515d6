"""VTK rectilinear-grid output in the reference's on-disk format.

``write_output`` writes what ``writeoutput`` (mgridlib/cartwrite.f90:455-481)
writes for a run on a px x py x pz process grid: ``<dirname>.pvtr`` (the
parallel header, mgridlib/cartwrite.f90:338-453) and one
``<dirname>/file_<rank>.vtr`` per rank (Writevtrfile, :178-336), each holding
that rank's block of every cell scalar as raw little-endian float64 in an
appended-data section, and the block's vertex coordinates. Here the whole
domain sits in one array (on the GPU or the host); the blocks are the
virtual ranks' slices, so the files are the ones an MPI run with that
process grid writes, byte for byte. ``write_serial_vtr`` is
Writeserialvtrfile (:5-176), one file for a whole grid.

Fields are interior (nx, ny, nz) or ghosted (nx+2, ny+2, nz+2) Fortran-order
arrays, or device tensors of (nz[+2], ny[+2], nx[+2]) elements; only the
interior is written. ``read_vtr`` reads a file back (cell scalars and
coordinates), for checks and for comparing fields between runs.
"""
from __future__ import annotations

import os
import re
import struct

import numpy as np

INTSIZE = 4       # the record length prefix: a default Fortran integer
DOUBLESIZE = 8


def _i8(v):
    """Fortran's (i8) edit descriptor: right-justified in 8 columns."""
    return "%8d" % v


def _extent(lo, hi):
    # x, y and z ranges; Fortran's trim() keeps the leading blanks of each
    # (i8) field and there is no separator before zmin, only its padding
    return (_i8(lo[0]) + " " + _i8(hi[0]) + " " + _i8(lo[1]) + " " + _i8(hi[1]) + _i8(lo[2]) + " " +
            _i8(hi[2]))


def _interior(a, n):
    """Interior (nx, ny, nz) float64 Fortran array of a field."""
    if hasattr(a, "detach"):   # torch tensor, (nz[+2], ny[+2], nx[+2]) elements
        a = a.detach().to("cpu").numpy().reshape(-1)
        g = tuple(v + 2 for v in n)
        if a.size == int(np.prod(g)):
            a = a.reshape(g[::-1]).T
        else:
            a = a.reshape(tuple(n)[::-1]).T
    a = np.asarray(a, dtype=np.float64)
    if a.shape == tuple(v + 2 for v in n):
        a = a[1:-1, 1:-1, 1:-1]
    if a.shape != tuple(n):
        raise ValueError("field of shape %s for a %s grid" % (a.shape, tuple(n)))
    return a


def _vtr_bytes(blocks, names, lo, hi, offs, dxs):
    """One .vtr: blocks are the (lx, ly, lz) interiors of each scalar."""
    lx, ly, lz = (h - l for l, h in zip(lo, hi))
    ext = _extent(lo, hi)
    scalars_size = lx * ly * lz * DOUBLESIZE
    head = ['<?xml version="1.0"?>\n',
            '<VTKFile type="RectilinearGrid" version="0.1" byte_order="LittleEndian">\n',
            '  <RectilinearGrid WholeExtent="%s">\n' % ext,
            '    <Piece Extent="%s">\n' % ext,
            '      <PointData>  </PointData>\n',
            '      <CellData> \n']
    off = 0
    for nm in names:
        head.append('         <DataArray type="Float64" Name="%s" format="appended" offset="%s"           />\n'
                    % (nm.rstrip(), _i8(off)))
        off += INTSIZE + scalars_size
    head.append('      </CellData>\n')
    head.append('      <Coordinates>\n')
    for axis, m in zip("XYZ", (lx, ly, lz)):
        head.append('<DataArray type="Float64" Name="%s"  format="appended" offset="%s" />\n' % (axis, _i8(off)))
        off += INTSIZE + (m + 1) * DOUBLESIZE
    head += ['      </Coordinates>\n', '    </Piece>\n', '  </RectilinearGrid>\n',
             '  <AppendedData encoding="raw">\n', '_']
    out = ["".join(head).encode("ascii")]
    for b in blocks:
        out.append(struct.pack("<i", scalars_size))
        out.append(np.asfortranarray(b, dtype="<f8").tobytes(order="F"))
    for m, o, d in zip((lx, ly, lz), offs, dxs):
        # local_xc(i) = offx + (i-1)*dx, i = 1..lx+1
        c = np.array([o + float(i) * d for i in range(m + 1)], dtype="<f8")
        out.append(struct.pack("<i", (m + 1) * DOUBLESIZE))
        out.append(c.tobytes())
    out.append(b"\n  </AppendedData>\n</VTKFile>\n")
    return b"".join(out)


def write_output(dirname, fields, names, n, length, p=(1, 1, 1), origin=(0.0, 0.0, 0.0), length_scale=1.0):
    """writeoutput: ``<dirname>.pvtr`` and ``<dirname>/file_<rank>.vtr`` for
    the px*py*pz blocks (rank r at block (r mod px, (r / px) mod py,
    r / (px py)), find3dindex, mgridlib/globalvars.f90:73-87). Coordinates
    are divided by length_scale the way the plasma solver scales its
    geometry (scalelengths, plasma_pde/solve_manager.f90:152-172: block
    offsets and lengths in metres, then divided). Returns the paths
    written."""
    n = tuple(int(v) for v in n)
    px, py, pz = (int(v) for v in p)
    if any(nv % pv for nv, pv in zip(n, (px, py, pz))):
        raise ValueError("grid %s does not divide over process grid %s" % (n, (px, py, pz)))
    if len(fields) != len(names):
        raise ValueError("one name per field")
    l = (n[0] // px, n[1] // py, n[2] // pz)
    data = [_interior(f, n) for f in fields]
    # domaindecompose (mgridlib/par_decompose.f90:15-42): d = len/N,
    # off = origin + (pind-1)*l*d, block length l*d; Writevtrfile spacing
    # = block length / l
    d = [float(L) / N for L, N in zip(length, n)]
    os.makedirs(dirname, exist_ok=True)
    paths = []
    nprocs = px * py * pz
    pieces = []
    for r in range(nprocs):
        pi = (r % px, (r // px) % py, r // (px * py))
        lo = tuple(q * m for q, m in zip(pi, l))
        hi = tuple((q + 1) * m for q, m in zip(pi, l))
        offs = tuple((float(o) + float(q * m) * dd) / length_scale for o, q, m, dd in zip(origin, pi, l, d))
        dxs = tuple(((m * dd) / length_scale) / m for m, dd in zip(l, d))
        blocks = [a[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] for a in data]
        src = "%s/file_%08d.vtr" % (dirname, r)
        with open(src, "wb") as fh:
            fh.write(_vtr_bytes(blocks, names, lo, hi, offs, dxs))
        paths.append(src)
        pieces.append('<Piece Extent="%s" Source="%s"/>\n' % (_extent(lo, hi), src))
    head = ['<?xml version="1.0"?>\n',
            '<VTKFile type="PRectilinearGrid" version="0.1" byte_order="LittleEndian">\n',
            '<PRectilinearGrid WholeExtent="%s" GhostLevel="0">\n' % _extent((0, 0, 0), n),
            '<PPointData>  </PPointData>\n', '<PCellData> \n']
    head += ['<PDataArray type="Float64" Name="%s"/>\n' % nm.rstrip() for nm in names]
    head += ['</PCellData>\n', '<PCoordinates>\n', '<PDataArray type="Float64" Name="X"/>\n',
             '<PDataArray type="Float64" Name="Y"/>\n', '<PDataArray type="Float64" Name="Z"/>\n',
             '</PCoordinates>\n']
    head += pieces + ['</PRectilinearGrid>\n', '</VTKFile>\n']
    with open(dirname + ".pvtr", "wb") as fh:
        fh.write("".join(head).encode("ascii"))
    return [dirname + ".pvtr"] + paths


def write_serial_vtr(filename, fields, names, n, length, offset=(0.0, 0.0, 0.0)):
    """Writeserialvtrfile: ``<filename>.vtr`` with the whole grid as one piece."""
    n = tuple(int(v) for v in n)
    data = [_interior(f, n) for f in fields]
    dxs = tuple(float(L) / m for L, m in zip(length, n))
    with open(filename + ".vtr", "wb") as fh:
        fh.write(_vtr_bytes(data, names, (0, 0, 0), n, tuple(float(o) for o in offset), dxs))
    return filename + ".vtr"


def read_vtr(path):
    """Cell scalars {name: (lx, ly, lz) array} and coordinates {"X","Y","Z"}
    of a .vtr in this format."""
    with open(path, "rb") as fh:
        raw = fh.read()
    cut = raw.index(b"<AppendedData")
    head = raw[:cut].decode("ascii")
    base = raw.index(b"_", cut) + 1
    ext = [int(v) for v in re.search(r'<Piece Extent="([^"]*)"', head).group(1).split()]
    lx, ly, lz = ext[1] - ext[0], ext[3] - ext[2], ext[5] - ext[4]
    out = {}
    for name, off in re.findall(r'<DataArray type="Float64" Name="([^"]*)"\s+format="appended" offset="\s*(\d+)"',
                                head):
        pos = base + int(off)
        (nbytes,) = struct.unpack("<i", raw[pos:pos + INTSIZE])
        a = np.frombuffer(raw[pos + INTSIZE:pos + INTSIZE + nbytes], dtype="<f8").copy()
        out[name] = a.reshape((lx, ly, lz), order="F") if name not in ("X", "Y", "Z") else a
    return out
