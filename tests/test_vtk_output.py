"""VTK output (mps3d_amd/vtk.py) against the reference's own writer.

The fixture tests/golden/vtk_output.npz holds the bytes that writeoutput
(mgridlib/cartwrite.f90:455-481), built from the reference sources and run
under MPI by tests/golden/make_golden.py, wrote for three grids and process
grids (2x1x2, 1x2x5, one rank) with two seeded cell scalars. The writer here
must produce the same .pvtr and .vtr files byte for byte from the whole-domain
arrays; the reader must return the blocks and coordinates.
"""
import os

import numpy as np
import pytest

from conftest import golden
from mps3d_amd import vtk

NAMES = ("Potential (V)", "Electron_density (#/m3)")


@pytest.mark.parametrize("tag", ["a", "b", "c"])
def test_write_output_matches_reference_bytes(tag, tmp_path, monkeypatch):
    d = golden("vtk_output")
    n, L, p = tuple(d[tag + "_n"]), tuple(d[tag + "_len"]), tuple(d[tag + "_p"])
    u0, dc = d[tag + "_u0"], d[tag + "_d"]
    monkeypatch.chdir(tmp_path)        # the reference writes relative to its working directory
    paths = vtk.write_output("vtkout", [u0, dc], NAMES, n, L, p)
    nranks = int(np.prod(p))
    assert len(paths) == 1 + nranks
    with open("vtkout.pvtr", "rb") as fh:
        assert fh.read() == d[tag + "_pvtr"].tobytes()
    for r in range(nranks):
        with open(os.path.join("vtkout", "file_%08d.vtr" % r), "rb") as fh:
            assert fh.read() == d["%s_vtr%d" % (tag, r)].tobytes(), r


def test_ghosted_fields_and_reader(tmp_path):
    """Ghosted inputs write their interior; read_vtr returns each block and
    the block's vertex coordinates."""
    n, L, p = (8, 6, 4), (0.8, 0.3, 0.4), (2, 1, 2)
    rng = np.random.default_rng(3)
    a = rng.standard_normal(tuple(v + 2 for v in n))
    b = rng.standard_normal(n)
    base = str(tmp_path / "out")
    vtk.write_output(base, [a, b], ("A", "B"), n, L, p)
    r3 = vtk.read_vtr(base + "/file_00000003.vtr")      # block (1, 0, 1)
    assert np.array_equal(r3["A"], a[1:-1, 1:-1, 1:-1][4:8, 0:6, 2:4])
    assert np.array_equal(r3["B"], b[4:8, 0:6, 2:4])
    assert np.allclose(r3["X"], np.linspace(0.4, 0.8, 5), rtol=1e-15)
    assert np.allclose(r3["Z"], np.linspace(0.2, 0.4, 3), rtol=1e-15)
    s = vtk.write_serial_vtr(str(tmp_path / "whole"), [b], ("B",), n, L)
    assert np.array_equal(vtk.read_vtr(s)["B"], b)


def test_uneven_decomposition_is_rejected(tmp_path):
    with pytest.raises(ValueError):
        vtk.write_output(str(tmp_path / "x"), [np.zeros((5, 4, 4))], ("A",), (5, 4, 4), (1, 1, 1), (2, 1, 1))


@pytest.mark.parametrize("chem", ["ion_elec", "H2_O2"])
def test_deck_output_matches_the_reference_plasma_output(chem, tmp_path, monkeypatch):
    """output_final of the reference plasma solver after one step of its
    deck (16^3, 2x2x1 ranks; tests/golden/plasma_output_files.npz): read back
    and written again by write_deck_output with the deck's scalar names and
    length scale, every file comes out byte for byte."""
    from mps3d_amd.chemistry import h2_o2, ion_elec
    from mps3d_amd.plasma import H2_O2_DECK, ION_ELEC_DECK, write_deck_output
    d = golden("plasma_output_files")
    deck = dict(ION_ELEC_DECK if chem == "ion_elec" else H2_O2_DECK)
    deck.update(n=(16, 16, 16), p=(2, 2, 1))
    ch = ion_elec() if chem == "ion_elec" else h2_o2()
    ref = {}
    fields = {}
    for r in range(4):
        key = "%s__output_final__file_%08d_vtr" % (chem, r)
        ref[r] = d[key].tobytes()
        path = tmp_path / ("ref_%d.vtr" % r)
        path.write_bytes(ref[r])
        blk = vtk.read_vtr(str(path))
        pi = (r % 2, (r // 2) % 2, 0)
        for nm, a in blk.items():
            if nm in ("X", "Y", "Z"):
                continue
            g = fields.setdefault(nm, np.zeros((16, 16, 16), order="F"))
            g[pi[0] * 8:(pi[0] + 1) * 8, pi[1] * 8:(pi[1] + 1) * 8, :] = a
    monkeypatch.chdir(tmp_path)
    write_deck_output("output_final", fields, ch, deck)
    with open("output_final.pvtr", "rb") as fh:
        assert fh.read() == d["%s__output_final_pvtr" % chem].tobytes()
    for r in range(4):
        with open(os.path.join("output_final", "file_%08d.vtr" % r), "rb") as fh:
            assert fh.read() == ref[r], r
