"""The opacity query's host side on the CPU (python -m ocotillo_amd.opacity,
the command line of tools/ocotillo-py/solve_opacity.py): its wavelength grid
is the Fortran calc_wavelength's bit for bit (the oracle's restatement,
pinned to the reference's wavegrid outputs), and without a GPU it fails
loudly instead of computing on the host."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import REPO, golden


@pytest.mark.parametrize("w0,w1,nw", [(3000.0, 10000.0, 2), (3000.0, 5000.0, 5), (100.0, 1e5, 137),
                                      (4000.0, 4000.0, 1), (912.0, 1e4, 1000)])
def test_grid_is_calc_wavelength(w0, w1, nw):
    from ocotillo_amd.opacity import wavelengths
    from oracle.oracle import calc_wavelength
    assert np.array_equal(wavelengths(w0, w1, nw), calc_wavelength(w1, w0, nw))


def test_grid_matches_reference_sample():
    """NongreyAGN's wavegrid.bin (w0 = 3000, w1 = 5000, nw = 2 from its input)."""
    from ocotillo_amd.opacity import wavelengths
    g = golden("sample_NongreyAGN.npz")
    assert np.array_equal(wavelengths(3000.0, 5000.0, int(g["nw"])), g["wavegrid"])


def test_options():
    from ocotillo_amd.opacity import parse
    a = parse(["-d", "1e-12", "-t", "8000", "-n", "7", "--lambdaMin", "500", "--lambdaMax", "2e4", "-v"])
    assert (a.density, a.temperature, a.nlambda, a.lambdaMin, a.lambdaMax, a.verbose) == (1e-12, 8000.0, 7, 500.0,
                                                                                          2e4, True)
    assert parse([]).nlambda == 2 and parse([]).temperature == 1000.0


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_fails_without_gpu():
    r = subprocess.run([sys.executable, "-m", "ocotillo_amd.opacity", "-t", "5000"], cwd=REPO,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "no HIP device" in r.stderr
    assert "Absorp. Coeff." not in r.stdout
