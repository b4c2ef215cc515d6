"""The Fortran drop-in boundary (mps3d_amd/fortran/mps3d_gpu.f90).

The module compiles with the image's Fortran compiler and the test driver
tests/fortran/laplace_gpu.f90 links against the C ABI library. Without a GPU
the driver stops at the device check (status MPS_ERR_NOLIB, exit 2); with a
GPU it solves the reference's Laplace deck (128^3, 2x2x2 process grid)
through gpu_perform_vcycle and gpu_solve and checks phi = x.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

DRIVER = os.path.join(ROOT, "tests", "fortran", "_build", "laplace_gpu")


def _built():
    """The driver, brought up to date with its sources when the compiler is
    here (make rebuilds nothing when it is current: a driver built against
    an older mps_opts layout would corrupt its stack)."""
    have_fc = shutil.which("amdflang") is not None or os.path.exists("/opt/rocm/bin/amdflang")
    if have_fc:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "mps3d_amd", "fortran")], check=True)
    elif not os.path.exists(DRIVER):
        pytest.skip("no Fortran compiler and no prebuilt driver")
    return DRIVER


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU device node is present")
def test_fortran_driver_reaches_the_device_boundary():
    r = subprocess.run([_built()], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2, r.stdout + r.stderr
    assert "no usable GPU" in r.stdout


@pytest.mark.gpu
def test_fortran_host_solves_the_laplace_deck(gpu):
    r = subprocess.run([_built()], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout
