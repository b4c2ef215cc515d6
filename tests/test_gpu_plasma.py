"""The device time-step loop (mps3d_amd/plasma.py, timestepping of
plasma_pde/solve_manager.f90:363-502) against the reference plasma solver
itself: tests/golden/plasma_ion_elec16.npz holds the fields of the
reference's VTK output after 1 and after 5 steps of its tests/ion_elec deck
at 16^3 cells on a 2x2x1 process grid (plasma_pde/main.f90 with the ion_elec
chemistry, built from the reference sources and run under MPI by
tests/golden/make_golden.py).

Every solve takes the reference's V-cycles on the same hierarchy, so the
fields follow the reference's trajectory; what separates them is rounding
(the cell width from the scaled lengths, exp and pow of the device library,
the MPI sum order of the RMS norm), bounded here by 1e-12 relative to each
field's largest magnitude (the electric field: to the largest component);
measured: bit-equal after one step except the production (3e-16), 2e-15 at
most after five.
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("k", [1, 5])
def test_time_steps_follow_the_reference_run(gpu, k):
    from mps3d_amd.plasma import ION_ELEC_DECK, PlasmaRun
    d = golden("plasma_ion_elec16")
    names = [str(s) for s in d["names"]]
    deck = dict(ION_ELEC_DECK)
    deck.update(n=tuple(int(v) for v in d["n"]), p=tuple(int(v) for v in d["p"]))
    run = PlasmaRun(deck)
    try:
        run.run(k)
        got = run.fields()
    finally:
        run.close()
    emax = max(np.abs(d["s%d_f%d" % (k, i)]).max() for i, nm in enumerate(names) if nm.startswith("Electric"))
    worst = {}
    for i, nm in enumerate(names):
        ref = d["s%d_f%d" % (k, i)]
        scale = emax if nm.startswith("Electric") else np.abs(ref).max()
        err = np.abs(got[nm] - ref).max() / scale
        worst[nm] = err
    print(k, worst, run.cycles)
    for nm, err in worst.items():
        assert err <= TOL, (nm, err)
