"""Snapshot job source on the CPU: command-line tags, the run-wide settings
(wavelength grid, analytic disk) against the oracle and the reference's
sample outputs, the Athena slab writer, and diagnostics.txt's E15.5 text
against the reference's own reference.out files."""
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN, golden

SAMPLES = ["GreyIsothermal", "NongreyAGN", "ReadFromAthena", "multiInTest"]


def test_parse_snapshots():
    from ocotillo_amd.snapshots import parse_snapshots
    assert parse_snapshots([]) == ["0000"]
    assert parse_snapshots(["0:2", "0007", "9999:10000"]) == ["0000", "0001", "0002", "0007", "9999", "10000"]
    with pytest.raises(ValueError):
        parse_snapshots(["5:3"])


@pytest.mark.parametrize("sample", SAMPLES)
def test_diagnostics_text_is_reference_out(sample):
    """The column (nxloc, nyloc) of the compiled reference's outputs, written
    as E15.5 text, is reference.out byte for byte."""
    from ocotillo_amd.fortio import diagnostics_text
    g = golden("sample_%s.npz" % sample)
    k = list(g["cols"]).index(int(g["nxloc"]) * int(g["nyloc"]) - 1)
    txt = diagnostics_text(g["U_p0_0"][:, k, :], g["A_p0_0"][:, k, :])
    assert txt == open(os.path.join(GOLDEN, "samples", sample, "reference.out")).read()


def test_format_e_edges():
    """Three-digit exponents drop the E (the standard's Ew.d form, which
    amdflang and gfortran print), rounding may carry into the exponent."""
    from ocotillo_amd.fortio import format_e
    assert format_e(1.2345678e-299) == "    0.12346-298"
    assert format_e(9.87654e150) == "    0.98765+151"
    assert format_e(1e-100) == "    0.10000E-99"
    assert format_e(2.5e-310) == "    0.25000-309"
    assert format_e(0.0) == "    0.00000E+00"
    assert format_e(9.999996e9) == "    0.10000E+11"
    assert format_e(-3.25e-120) == "   -0.32500-119"


@pytest.mark.parametrize("sample", SAMPLES)
def test_run_settings(sample):
    from ocotillo_amd.snapshots import Run
    run = Run(os.path.join(GOLDEN, "samples", sample))
    g = golden("sample_%s.npz" % sample)
    assert np.array_equal(run.waves(), g["wavegrid"])
    assert run.fully_ionized_T == 2e4


@pytest.mark.parametrize("sample", ["GreyIsothermal", "NongreyAGN"])
def test_analytic_disk_block(sample):
    """The analytic disk's pillar equals the oracle's calc_grid /
    calc_density / calc_temperature bit for bit, its grid the reference's
    zgrid.bin."""
    import ctypes as C
    from oracle.oracle import lib
    from ocotillo_amd.snapshots import Run
    run = Run(os.path.join(GOLDEN, "samples", sample))
    rho, T, z, dz = run.disk_block()
    g = golden("sample_%s.npz" % sample)
    assert np.array_equal(z, g["zgrid"])
    L = lib(False)
    nz = run.shapes["nz"]
    zo = np.zeros(nz + 6)
    dzo = C.c_double()
    L.oo_calc_grid(nz, run.z1, run.z0, zo, C.byref(dzo))
    ro = np.zeros(nz)
    L.oo_calc_density(nz, zo, run.rho0, run.H, run.rho_floor, ro)
    To = np.zeros(nz)
    L.oo_calc_temperature(nz, zo, run.mid_temp, run.sigma, run.floor_temp, run.switch_temp, None, 0, To)
    assert np.array_equal(zo, z) and dzo.value == dz
    assert np.array_equal(rho[0, 0], ro) and np.array_equal(T[-1, -1], To)
    assert (rho == rho[0, 0]).all() and (T == T[0, 0]).all()


def test_write_slab_round_trip(tmp_path):
    from ocotillo_amd.athena import SlabHeader, write_slab
    rng = np.random.default_rng(3)
    f = [rng.normal(size=(4, 3, 5)) for _ in range(5)]
    write_slab(tmp_path / "a.bin", f, np.arange(5.0), np.arange(3.0), np.arange(4.0) * 0.5, 0.4)
    h = SlabHeader(str(tmp_path / "a.bin"))
    assert (h.nxloc, h.nyloc, h.nzl, h.gamma1) == (5, 3, 4, 0.4)
    assert all(np.array_equal(a, b) for a, b in zip(h.fields(), f))
    assert np.array_equal(h.z, np.arange(4.0) * 0.5)


def test_sample_file_rewritten_identically(tmp_path):
    """Re-writing a reference sample slab from its parsed header and fields
    reproduces the file byte for byte (the writer is the reader's inverse)."""
    from ocotillo_amd.athena import SlabHeader, write_slab
    src = os.path.join(GOLDEN, "samples", "multiInTest", "input_athena", "distributed", "id1",
                       "agn_snova_3d-id1.0000.bin")
    h = SlabHeader(src)
    ints = np.fromfile(src, "<i4", 8)[4:]
    write_slab(tmp_path / "b.bin", [np.asarray(a) for a in h.fields()], h.x, h.y, h.z, h.gamma1,
               coordsys=h.coordsys, header_ints=tuple(int(v) for v in ints), cs=h.cs, t=h.t, dt=h.dt)
    assert open(src, "rb").read() == open(tmp_path / "b.bin", "rb").read()


def test_snapshot_listing(tmp_path):
    from ocotillo_amd.athena import AthenaRun
    src = os.path.join(GOLDEN, "samples", "ReadFromAthena")
    for f in ("input.in", "resolution.in"):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    d = tmp_path / "input_athena" / "distributed" / "id0"
    d.mkdir(parents=True)
    for t in ("0002", "0000", "0010"):
        (d / ("agn_snova_3d.%s.bin" % t)).write_bytes(b"")
    (d / "other.0001.bin").write_bytes(b"")
    assert AthenaRun(str(tmp_path)).snapshots() == ["0000", "0002", "0010"]
