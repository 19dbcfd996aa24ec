"""The reference's autotest (bin/rt_autotest) run through the Fortran drop-in
host on the GPU: diagnostics.txt must equal each sample's reference.out, and
the full-precision binary outputs must match the reference's (golden
sample_*.npz, written by the compiled reference) to 1e-12."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, REPO, flux_err, golden, relerr

pytestmark = pytest.mark.gpu
EXE = os.path.join(REPO, "ocotillo_amd", "fortran", "flux_feautrier_gpu.x")
SAMPLES = ["GreyIsothermal", "NongreyAGN", "ReadFromAthena", "multiInTest"]


@pytest.mark.parametrize("sample", SAMPLES)
def test_sample(sample, tmp_path):
    from ocotillo_amd.fortio import read_block
    print("autotest sample", sample, flush=True)
    src = os.path.join(GOLDEN, "samples", sample)
    for f in ("input.in", "resolution.in"):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    if os.path.isdir(os.path.join(src, "input_athena")):
        os.symlink(os.path.join(src, "input_athena"), tmp_path / "input_athena")
    (tmp_path / "output").mkdir()
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    diag = (tmp_path / "output" / "diagnostics.txt").read_text()
    assert diag == open(os.path.join(src, "reference.out")).read()
    g = golden("sample_%s.npz" % sample)
    nz, nw, nyloc, nxloc = int(g["nz"]), int(g["nw"]), int(g["nyloc"]), int(g["nxloc"])
    cols = g["cols"]
    for px in range(int(g["nprocx"])):
        for py in range(int(g["nprocy"])):
            U, A, V = read_block(str(tmp_path / "output"), px, py, "0000", nw, nxloc, nyloc, nz)
            tag = "p%d_%d" % (px, py)
            ncol = nxloc * nyloc
            assert relerr(A.reshape(nw, ncol, nz)[:, cols], g["A_" + tag]).max() <= 1e-12
            assert relerr(U.reshape(nw, ncol, nz + 6)[:, cols], g["U_" + tag]).max() <= 1e-12
            assert flux_err(V.reshape(nw, ncol, nz)[:, cols], g["V_" + tag]).max() <= 1e-12
            for nm, arr in (("U", U), ("A", A), ("V", V)):
                ref = g["sum%s_%s" % (nm, tag)]
                assert abs(np.abs(arr).sum() - ref[1]) <= 1e-12 * ref[1]
