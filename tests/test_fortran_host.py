"""The Fortran drop-in host (ocotillo_amd/fortran/flux_feautrier_gpu.x) on the
CPU: run-time resolution.in and namelist parsing, and a loud failure when
no GPU is present (it never computes on the host)."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, REPO

EXE = os.path.join(REPO, "ocotillo_amd", "fortran", "flux_feautrier_gpu.x")
SAMPLES = {"GreyIsothermal": (1, 1, 1, 10, 1, 1, 1, 1, 1),
           "NongreyAGN": (2, 1, 1, 10, 1, 1, 1, 1, 1),
           "ReadFromAthena": (5, 31, 32, 33, 1, 1, 1, 31, 32),
           "multiInTest": (5, 32, 34, 36, 2, 1, 2, 16, 34)}

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built (no amdflang)")


@needs_exe
@pytest.mark.parametrize("sample", sorted(SAMPLES))
def test_resolution_and_namelists_parsed_at_run_time(sample):
    r = subprocess.run([EXE, "--print-config"], cwd=os.path.join(GOLDEN, "samples", sample),
                       capture_output=True, text=True, timeout=60)
    line = [l for l in r.stdout.splitlines() if l.startswith("shapes")][0]
    assert tuple(int(t) for t in line.split(":")[1].split()) == SAMPLES[sample]


@needs_exe
def test_resolution_expressions(tmp_path):
    (tmp_path / "resolution.in").write_text(
        "integer, parameter :: nw=3   ! c\n"
        "integer, parameter :: nx=2*(3+1),ny=12,nz=40-8/2\n"
        "integer, parameter :: nprocx=2,nprocy=3,nprocz=1\n"
        "integer, parameter :: nyloc=ny/nprocy,nxloc=nx/nprocx\n")
    shutil.copy(os.path.join(GOLDEN, "samples", "NongreyAGN", "input.in"), tmp_path / "input.in")
    r = subprocess.run([EXE, "--print-config"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    line = [l for l in r.stdout.splitlines() if l.startswith("shapes")][0]
    assert tuple(int(t) for t in line.split(":")[1].split()) == (3, 8, 12, 36, 2, 3, 1, 4, 4)


@needs_exe
@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_fails_loudly_without_gpu(tmp_path):
    for f in ("input.in", "resolution.in"):
        shutil.copy(os.path.join(GOLDEN, "samples", "NongreyAGN", f), tmp_path / f)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "no HIP device" in r.stdout
    assert not (tmp_path / "output" / "diagnostics.txt").exists()
