"""OCO_CHECK_FINITE: the device check of a block solve's outputs.

The reference's debug build traps the first invalid, divide-by-zero or
overflow operation (-ffpe-trap, source/Makefile.source:8) and stops.  With
the check on, a NaN or infinity anywhere in U, V or absorp turns a block
solve into OCO_E_NONFINITE, located at the first such (column, wavelength)
system in the reference's loop order (ix, then iy, then iw) and its first
such cell (absorp first, then U, then V); the streamed block outputs then
write no file.  A NaN density in one cell makes every wavelength's alpha
there NaN, so the first is that column's first wavelength, that cell of
absorp.  Off (the default), the same solve returns OK with the NaNs in its
outputs, as the reference built without traps does; on a clean block the
check changes no bit.
"""
import os
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = ("fp64", "mixed")


def setup(c, nz=48, nw=12):
    from ocotillo_amd.workloads import disk_columns, log_waves
    rho, T, dz = disk_columns(3, 4, nz)
    c.set_wavelengths(log_waves(nw))
    c.set_physics(2e4, False, 0.4)
    return rho, T, dz


def mode_code(name):
    import ocotillo_amd as oco
    return {"fp64": oco.MODE_FP64, "mixed": oco.MODE_MIXED}[name]


@pytest.mark.parametrize("mode", MODES)
def test_clean_block_unchanged(mode):
    import ocotillo_amd as oco
    with oco.Context(0) as c:
        rho, T, dz = setup(c)
        U0, V0, A0, d0 = c.solve_block(rho, T, dz, mode=mode_code(mode))
        c.set_checks(finite=True)
        U1, V1, A1, d1 = c.solve_block(rho, T, dz, mode=mode_code(mode))
    assert d0.status == 0 and d1.status == 0
    assert np.array_equal(U0, U1) and np.array_equal(V0, V1) and np.array_equal(A0, A1)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("cell", [(1, 0, 5), (2, 3, 0), (0, 2, 47)])
def test_nan_input_located(mode, cell):
    import ocotillo_amd as oco
    ix, iy, z = cell
    print("case", mode, cell, flush=True)
    with oco.Context(0) as c:
        rho, T, dz = setup(c)
        rho = rho.copy()
        rho[ix, iy, z] = np.nan
        U, V, A, d = c.solve_block(rho, T, dz, mode=mode_code(mode))   # unchecked: OK, NaNs kept
        assert d.status == 0 and np.isnan(A[:, ix, iy, z]).all() and np.isfinite(A[:, 0, 0]).all()
        c.set_checks(finite=True)
        with pytest.raises(oco.OcoError) as e:
            c.solve_block(rho, T, dz, mode=mode_code(mode))
        assert e.value.status == oco.E_NONFINITE
        g = e.value.diag
        assert (g.fail_ix, g.fail_iy, g.fail_iw, g.fail_j) == (ix + 1, iy + 1, 1, z + 1)


def test_nan_files_not_written():
    import ocotillo_amd as oco
    with oco.Context(0) as c, tempfile.TemporaryDirectory() as td:
        rho, T, dz = setup(c)
        c.set_checks(finite=True)
        c.solve_block_files(rho, T, dz, td, "0000")
        assert os.path.exists(os.path.join(td, "flux_0000.bin"))
        bad = rho.copy()
        bad[1, 2, 10] = np.nan
        with pytest.raises(oco.OcoError) as e:
            c.solve_block_files(bad, T, dz, td, "0001", stage_bytes=4 * 12 * 48 * 8 * 3)   # several chunks
        assert e.value.status == oco.E_NONFINITE
        g = e.value.diag
        assert (g.fail_ix, g.fail_iy, g.fail_iw, g.fail_j) == (2, 3, 1, 11)
        assert not any(n.endswith("_0001.bin") for n in os.listdir(td))


def test_nan_noalloc():
    """The allocation-free entry point checks too (its key lives in the
    static scratch and is cleared on every call)."""
    import ocotillo_amd as oco
    from conftest import DeviceArray
    with oco.Context(0) as c:
        rho, T, dz = setup(c)
        bad = rho.copy()
        bad[2, 1, 7] = np.nan
        nx, ny, nz = rho.shape
        nw = c.nw
        # device arrays in the library's layouts: rho, T (nz, ny, nx) Fortran = [nx, ny, nz] C
        d_rho, d_T = DeviceArray(bad.nbytes).upload(bad), DeviceArray(T.nbytes).upload(np.ascontiguousarray(T))
        d_U = DeviceArray(8 * nw * nx * ny * (nz + 6))
        d_V = DeviceArray(8 * nw * nx * ny * nz)
        d_A = DeviceArray(8 * nw * nx * ny * nz)
        try:
            c.set_checks(finite=True)
            args = (nz, ny, nx, dz, d_rho.ptr, d_T.ptr, d_U.ptr, d_V.ptr, d_A.ptr)
            with pytest.raises(oco.OcoError) as e:
                c.solve_block_device(*args, mode=oco.MODE_MIXED, noalloc=True)
            assert e.value.status == oco.E_NONFINITE
            g = e.value.diag
            assert (g.fail_ix, g.fail_iy, g.fail_iw, g.fail_j) == (3, 2, 1, 8)
            d_rho.upload(np.ascontiguousarray(rho))
            d = c.solve_block_device(*args, mode=oco.MODE_MIXED, noalloc=True)
            assert d.status == 0
        finally:
            for x in (d_rho, d_T, d_U, d_V, d_A):
                x.free()
