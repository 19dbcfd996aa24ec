"""Block outputs streamed to files (oco_solve_block_files).

Bars:
  * the three files hold exactly what oco_solve_block returns (bit for bit,
    both precision modes), whatever the wavelength chunking: one chunk, one
    wavelength per chunk, and chunks that do not divide nw;
  * they are one record each in the reference's layout (read back with the
    package's reader, sizes checked), under the reference's names;
  * the printed column (nxloc, nyloc) and the diagnostics equal the block
    path's;
  * on the reference's zero-pivot stop no file is written and earlier files
    of the same name are left as they were; the status and the failing
    (ix, iy, iw, row) are oco_solve_block's.
"""
import os

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import ocotillo_amd as oco
    with oco.Context(0) as c:
        c.set_physics(2e4, False, 0.4)
        yield c


def per_wavelength_bytes(nx, ny, nz):
    return nx * ny * (3 * nz + 6) * 8


# (nx, ny, nz, nw, chunk wavelengths: 0 = default staging)
CASES = [(3, 2, 64, 70, 1), (4, 4, 100, 130, 3), (2, 3, 300, 64, 0), (5, 1, 513, 17, 4), (1, 1, 16, 5, 2)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%dx%d_nw%d_chunk%d" % c)
@pytest.mark.parametrize("mode_name", ["fp64", "mixed"])
def test_files_equal_block_path(ctx, tmp_path, case, mode_name):
    import ocotillo_amd as oco
    from ocotillo_amd.fortio import read_block
    from ocotillo_amd.workloads import disk_columns, log_waves
    nx, ny, nz, nw, chunk = case
    mode = oco.MODE_MIXED if mode_name == "mixed" else oco.MODE_FP64
    print("case", case, mode_name, flush=True)
    rho, T, dz = disk_columns(nx, ny, nz, seed=nz + nw)
    ctx.set_wavelengths(log_waves(nw))
    U, V, A, d = ctx.solve_block(rho, T, dz, mode=mode)
    out = tmp_path / "output"
    blk = out / "procx0_procy0"
    blk.mkdir(parents=True)
    stage = chunk * per_wavelength_bytes(nx, ny, nz)
    Uc, Ac, df = ctx.solve_block_files(rho, T, dz, str(blk), "0007", mode=mode, stage_bytes=stage)
    Uf, Af, Vf = read_block(str(out), 0, 0, "0007", nw, nx, ny, nz)
    assert np.array_equal(Uf, U) and np.array_equal(Af, A) and np.array_equal(Vf, V)
    assert sorted(os.listdir(blk)) == ["absorption_coefficients_0007.bin", "flux_0007.bin",
                                       "mean_intensity_0007.bin"]
    assert os.path.getsize(blk / "mean_intensity_0007.bin") == U.nbytes + 8
    assert np.array_equal(Uc, U[:, -1, -1, :]) and np.array_equal(Ac, A[:, -1, -1, :])
    for f in ("sum_a", "sum_b", "sum_c", "sum_d", "minU", "maxU", "minV", "maxV", "minA", "maxA"):
        assert getattr(df, f) == getattr(d, f), f
    assert df.status == 0 and df.total_ms > 0


def test_grey_block(tmp_path):
    import ocotillo_amd as oco
    from ocotillo_amd.fortio import read_block
    from ocotillo_amd.workloads import disk_columns
    with oco.Context(0) as c:
        c.set_physics(2e4, True, 0.4)
        c.set_wavelengths(np.array([4000.0]))
        rho, T, dz = disk_columns(3, 3, 40, seed=5)
        U, V, A, d = c.solve_block(rho, T, dz)
        blk = tmp_path / "procx0_procy0"
        blk.mkdir()
        c.solve_block_files(rho, T, dz, str(blk), "0000")
        Uf, Af, Vf = read_block(str(tmp_path), 0, 0, "0000", 1, 3, 3, 40)
        assert np.array_equal(Uf, U) and np.array_equal(Af, A) and np.array_equal(Vf, V)


def test_zero_pivot_writes_nothing(ctx, tmp_path):
    import ocotillo_amd as oco
    g = golden("dump_pivot0.npz")
    ncol, nz = int(g["ncol"]), int(g["nz"])
    rho, T = g["rho"].reshape(1, ncol, nz), g["T"].reshape(1, ncol, nz)
    ctx.set_wavelengths(g["waves"])
    _, _, _, d = ctx.solve_block(rho, T, float(g["dz"]), raise_on_pivot=False)
    assert d.status == int(g["ref_stop_stage"])
    blk = tmp_path / "procx0_procy0"
    blk.mkdir()
    old = blk / "flux_0000.bin"
    old.write_bytes(b"earlier run")
    for stage in (0, per_wavelength_bytes(1, ncol, nz) * 5):   # one chunk; chunks of 5 wavelengths
        _, _, df = ctx.solve_block_files(rho, T, float(g["dz"]), str(blk), "0000", stage_bytes=stage,
                                         raise_on_pivot=False)
        assert (df.status, df.fail_ix, df.fail_iy, df.fail_iw, df.fail_j) == (d.status, d.fail_ix, d.fail_iy,
                                                                            d.fail_iw, d.fail_j)
        assert sorted(os.listdir(blk)) == ["flux_0000.bin"]
        assert old.read_bytes() == b"earlier run"
    with pytest.raises(oco.OcoError) as e:
        ctx.solve_block_files(rho, T, float(g["dz"]), str(blk), "0000")
    assert e.value.status == d.status


def test_missing_directory(ctx, tmp_path):
    import ocotillo_amd as oco
    from ocotillo_amd.workloads import disk_columns, log_waves
    rho, T, dz = disk_columns(2, 2, 32, seed=1)
    ctx.set_wavelengths(log_waves(8))
    with pytest.raises(oco.OcoError) as e:
        ctx.solve_block_files(rho, T, dz, str(tmp_path / "nope"), "0000")
    assert e.value.status == oco.E_IO
