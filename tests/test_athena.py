"""Athena snapshot input on the CPU: the host-side reader (namelists,
resolution.in, .bin headers, code units, z grid) and the oracle's
restatement of the field arithmetic, pinned end to end against the
reference's own Athena samples (ReadFromAthena: one z slab; multiInTest:
two x blocks of two z slabs each).

The sample fixtures hold what the compiled reference wrote for these inputs
(zgrid.bin and, per block, U, alpha, V at a seeded set of columns).  The
oracle's path is: oo_athena_units -> oo_athena_slab_raw per slab ->
oo_athena_postprocess -> the glibc oracle's block solve; the reference's
libm is glibc, so every value must be bit-identical.  A dropped kinetic
term, a wrong transpose, a swapped unit or a wrong temperature branch
changes T and with it every alpha of the column.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden

SAMPLES = ["ReadFromAthena", "multiInTest"]


def run(sample):
    from ocotillo_amd.athena import AthenaRun
    return AthenaRun(os.path.join(GOLDEN, "samples", sample))


def oracle_block(r, px, py, units=None, avoid=None):
    """rho3d, T3d, z, dz of one block through the oracle's restatement."""
    from oracle.oracle import lib
    L = lib(False)
    heads, offs, z, dz, gamma1 = r.block_layout(px, py)
    s = r.shapes
    nz, nyloc, nxloc = s["nz"], s["nyloc"], s["nxloc"]
    rho = np.zeros((nxloc, nyloc, nz))
    temp = np.zeros_like(rho)
    for h, iz0 in zip(heads, offs):
        f = [np.ascontiguousarray(a) for a in h.fields()]
        L.oo_athena_slab_raw(nz, nyloc, nxloc, h.nzl, iz0, *[a.reshape(-1) for a in f],
                             int(r.avoid_momentum_underflow if avoid is None else avoid),
                             rho.reshape(-1), temp.reshape(-1))
    ul, ud, ut = units if units is not None else oracle_units(r)
    L.oo_athena_postprocess(nz, nyloc, nxloc, ud, r.rho_floor, gamma1, ut, r.tgauss(z), r.switch_temp,
                            r.floor_temp, rho.reshape(-1), temp.reshape(-1))
    return rho, temp, z, dz


def oracle_units(r):
    from oracle.oracle import lib
    from ocotillo_amd.athena import read_namelist
    a = read_namelist(os.path.join(r.dir, "input.in"), "athena_input")
    out = [C.c_double() for _ in range(3)]
    lib(False).oo_athena_units(a["mbh_solarmasses"], a["r0ref_rg"], a["aspect_ratio"],
                               a["mean_molecular_weight"], a["rho0"], *[C.byref(o) for o in out])
    return tuple(o.value for o in out)


@pytest.mark.parametrize("sample", SAMPLES)
def test_reader_layout_and_grid(sample):
    """Shapes, slab offsets and the physical z grid equal the reference's
    zgrid.bin bit for bit (pins unit_length, nint offsets and the ghosts)."""
    r = run(sample)
    g = golden("sample_%s.npz" % sample)
    assert (r.shapes["nz"], r.shapes["nyloc"], r.shapes["nxloc"]) == (int(g["nz"]), int(g["nyloc"]),
                                                                     int(g["nxloc"]))
    for px in range(r.shapes["nprocx"]):
        heads, offs, z, dz, _ = r.block_layout(px, 0)
        assert offs == [sum(h.nzl for h in heads[:i]) for i in range(len(heads))]
        assert np.array_equal(z, g["zgrid"])
        assert dz == (z[-4] - z[3]) / (r.shapes["nz"] - 1)


@pytest.mark.parametrize("sample", SAMPLES)
def test_host_units_equal_oracle(sample):
    r = run(sample)
    assert (r.unit_length, r.unit_density, r.unit_temperature) == oracle_units(r)


@pytest.mark.parametrize("sample", SAMPLES)
def test_oracle_athena_path_bit_exact(sample):
    """Athena input construction + solve through the oracle == the compiled
    reference's outputs, every kept column, every block."""
    from oracle.oracle import Oracle
    r = run(sample)
    g = golden("sample_%s.npz" % sample)
    waves = g["wavegrid"]
    o = Oracle(waves)
    cols = g["cols"]
    dz1 = dz2 = None
    for px in range(int(g["nprocx"])):
        for py in range(int(g["nprocy"])):
            rho, T, z, dz = oracle_block(r, px, py)
            if dz1 is None:                 # set on the root block (source/flux_feautrier.f90:151-154)
                dz1, dz2 = 1.0 / dz, dz * dz
            U, V, A, d = o.solve_block(rho, T, dz, dz1=dz1, dz2=dz2)
            assert d.status == 0
            nw, nx, ny = U.shape[:3]
            tag = "p%d_%d" % (px, py)
            for nm, arr in (("U", U), ("A", A), ("V", V)):
                got = arr.reshape(nw, nx * ny, -1)[:, cols]
                assert np.array_equal(got, g["%s_%s" % (nm, tag)]), (nm, tag)
                assert arr.sum() == g["sum%s_%s" % (nm, tag)][0]


def test_sample_branch_coverage():
    """What the bit-exact sample test pins: ReadFromAthena has every cell
    above switchTemp (the kinetic/internal energy arithmetic and both units
    decide every T); multiInTest has every cell below it with the Gaussian
    profile above floorTemp (the profile branch); both hold zero momenta.
    The density floor, the floorTemp branch and nonzero momenta below
    epsilon are not in the samples: test_oracle_branch_properties pins them."""
    from ocotillo_amd.athena import EPSI
    for sample, below_all in (("ReadFromAthena", False), ("multiInTest", True)):
        r = run(sample)
        heads, offs, z, dz, gamma1 = r.block_layout(0, 0)
        Tphys = raw_temperature(r, heads, offs) * gamma1 * r.unit_temperature
        assert (Tphys < r.switch_temp).all() == below_all and (Tphys < r.switch_temp).any() == below_all
        assert (r.tgauss(z) > r.floor_temp).all()
        f = [np.concatenate([np.asarray(h.fields()[i]).reshape(-1) for h in heads]) for i in range(5)]
        assert any((m == 0).any() for m in f[1:4])
        assert not any(((np.abs(m) < EPSI) & (m != 0)).any() for m in f[1:4])
        assert (f[0] * r.unit_density >= r.rho_floor).all()


def raw_temperature(r, heads, offs):
    """Code-unit internal energy per mass of a block (the oracle's raw pass)."""
    from oracle.oracle import lib
    s = r.shapes
    rho = np.zeros((s["nxloc"], s["nyloc"], s["nz"]))
    temp = np.zeros_like(rho)
    for h, iz0 in zip(heads, offs):
        fl = [np.ascontiguousarray(a).reshape(-1) for a in h.fields()]
        lib(False).oo_athena_slab_raw(s["nz"], s["nyloc"], s["nxloc"], h.nzl, iz0, *fl, 1, rho.reshape(-1),
                                      temp.reshape(-1))
    return temp


def synthetic_slab(seed=7, nxloc=5, nyloc=4, nzl=6):
    """A slab that fires every branch: densities across the floor, momenta
    of ordinary size, zero, and nonzero below epsilon, internal energies
    across switchTemp.  Fields C-ordered [nzl, nyloc, nxloc]."""
    rng = np.random.default_rng(seed)
    shape = (nzl, nyloc, nxloc)
    rho = 10.0 ** rng.uniform(-8, 2, shape)
    mom = []
    for _ in range(3):
        m = rng.normal(0, 1, shape) * rho
        pick = rng.integers(0, 3, shape)
        mom.append(np.where(pick == 0, 0.0, np.where(pick == 1, rng.uniform(-1, 1, shape) * 1e-17, m)))
    eint = 10.0 ** rng.uniform(-2, 2, shape)          # per unit mass, code units
    e = sum(0.5 * m * m / rho for m in mom) + rho * eint
    # one cell whose only kinetic term is a sub-epsilon momentum of the
    # order of its energy, so the guard is visible in the result
    rho[0, 0, 0], e[0, 0, 0] = 1e-8, 2e-26
    mom[0][0, 0, 0], mom[1][0, 0, 0], mom[2][0, 0, 0] = 1e-17, 0.0, 0.0
    return [rho] + mom + [e]


def synthetic_units(**kw):
    """The samples' units with switchTemp inside the synthetic slab's range."""
    from ocotillo_amd import AthenaUnits
    r = run("multiInTest")
    u = AthenaUnits(r.unit_density, r.unit_temperature, 0.4, r.rho_floor, 2.0e5, r.floor_temp, 1)
    for k, v in kw.items():
        setattr(u, k, v)
    return u


def synthetic_tgauss(nz, floor_temp):
    tg = np.geomspace(0.1, 30.0, nz) * floor_temp
    tg[nz // 2] = floor_temp                            # a tie: the strict > gives floorTemp
    return tg


def oracle_slab(fields, tg, u, nz, iz0):
    """The oracle's two stages on one slab of a block whose other z rows hold
    -1 (rho) and -2 (T)."""
    from oracle.oracle import lib
    L = lib(False)
    nzl, nyloc, nxloc = fields[0].shape
    rho = np.full((nxloc, nyloc, nz), -1.0)
    temp = np.full_like(rho, -2.0)
    L.oo_athena_slab_raw(nz, nyloc, nxloc, nzl, iz0, *[np.ascontiguousarray(a).reshape(-1) for a in fields],
                         u.avoid_momentum_underflow, rho.reshape(-1), temp.reshape(-1))
    sl = slice(iz0, iz0 + nzl)
    rs, ts = np.ascontiguousarray(rho[:, :, sl]), np.ascontiguousarray(temp[:, :, sl])
    L.oo_athena_postprocess(nzl, nyloc, nxloc, u.unit_density, u.rho_floor, u.gamma1, u.unit_temperature,
                            np.ascontiguousarray(tg[sl]), u.switch_temp, u.floor_temp, rs.reshape(-1),
                            ts.reshape(-1))
    rho[:, :, sl], temp[:, :, sl] = rs, ts
    return rho, temp


def test_oracle_branch_properties():
    """The branches the samples leave open, by their definitions:
    * density: rho_floor exactly where rho * unit_density falls below it,
      rho * unit_density elsewhere;
    * temperature below switchTemp: max(tgauss(z), floorTemp), a tie giving
      floorTemp; at or above switchTemp: untouched;
    * momentum guard: guard on == guard off with the sub-epsilon momenta set
      to zero, and != guard off where such a momentum is the only kinetic term;
    * internal energy: (E - |M|^2 / (2 rho)) / rho within 4 ulp of the exact
      rational value wherever E > 2 ekin (no cancellation);
    * the transpose puts file element (ix, iy, k) at (ix, iy, iz0 + k) and
      z rows outside the slab are untouched."""
    from fractions import Fraction as Fr
    from ocotillo_amd.athena import EPSI
    nz, iz0 = 11, 3
    f = synthetic_slab()
    u = synthetic_units()
    tg = synthetic_tgauss(nz, u.floor_temp)
    rho, T = oracle_slab(f, tg, u, nz, iz0)
    sl = slice(iz0, iz0 + f[0].shape[0])
    assert (rho[:, :, :iz0] == -1).all() and (rho[:, :, sl.stop:] == -1).all()
    assert (T[:, :, :iz0] == -2).all() and (T[:, :, sl.stop:] == -2).all()
    phys = np.transpose(f[0], (2, 1, 0)) * u.unit_density          # [ix, iy, k]
    low = phys < u.rho_floor
    assert low.any() and (~low).any()
    assert np.array_equal(rho[:, :, sl], np.where(low, u.rho_floor, phys))
    _, Tr = oracle_slab(f, tg, synthetic_units(switch_temp=-np.inf), nz, iz0)
    Tr = Tr[:, :, sl]
    below = Tr < u.switch_temp
    tgz = np.broadcast_to(tg[sl], Tr.shape)
    assert below.any() and (~below).any()
    assert (below & (tgz > u.floor_temp)).any() and (below & (tgz == u.floor_temp)).any()
    assert (below & (tgz < u.floor_temp)).any()
    assert np.array_equal(T[:, :, sl], np.where(below, np.maximum(tgz, u.floor_temp), Tr))
    tiny = [(np.abs(m) < EPSI) & (m != 0) for m in f[1:4]]
    assert all(t.any() for t in tiny)
    zeroed = [f[0]] + [np.where(t, 0.0, m) for t, m in zip(tiny, f[1:4])] + [f[4]]
    off = synthetic_units(avoid_momentum_underflow=0)
    assert np.array_equal(T, oracle_slab(zeroed, tg, off, nz, iz0)[1])
    assert not np.array_equal(oracle_slab(f, tg, synthetic_units(switch_temp=-np.inf), nz, iz0)[1],
                              oracle_slab(f, tg, synthetic_units(switch_temp=-np.inf, avoid_momentum_underflow=0),
                                          nz, iz0)[1])
    scale = Fr(u.gamma1) * Fr(u.unit_temperature)
    checked = 0
    nzl, nyloc, nxloc = f[0].shape
    for ix in range(nxloc):
        for iy in range(nyloc):
            for k in range(nzl):
                r_ = Fr(float(f[0][k, iy, ix]))
                ms = [float(m[k, iy, ix]) for m in f[1:4]]
                ek = sum(Fr(m) * Fr(m) for m in ms if abs(m) >= EPSI) / (2 * r_)
                e_ = Fr(float(f[4][k, iy, ix]))
                if e_ <= 2 * ek:
                    continue
                exact = (e_ - ek) / r_ * scale
                assert abs(Fr(float(Tr[ix, iy, k])) - exact) <= 4 * EPSI * exact, (ix, iy, k)
                checked += 1
    assert checked > nxloc * nyloc * nzl // 2


def test_namelist_and_resolution_parsing(tmp_path):
    from ocotillo_amd.athena import read_namelist, read_resolution
    (tmp_path / "input.in").write_text(
        "&other\n  x=1 /\n&athena_input\n ! comment = 3\n  RunName='a b', datadir=\"./d/e\"\n"
        "  Mbh_SolarMasses = 1d8, r0ref_rg=1e3 lavoid_momentum_underflow=.false.\n  rho_floor=2.5D-15\n/\n")
    a = read_namelist(tmp_path / "input.in", "athena_input")
    assert a == {"runname": "a b", "datadir": "./d/e", "mbh_solarmasses": 1e8, "r0ref_rg": 1e3,
                 "lavoid_momentum_underflow": False, "rho_floor": 2.5e-15}
    with pytest.raises(KeyError):
        read_namelist(tmp_path / "input.in", "temperature_input")
    (tmp_path / "resolution.in").write_text(
        "integer, parameter :: nw=3   ! c\ninteger, parameter :: nx=2*(3+1),ny=12,nz=40-8/2\n"
        "integer, parameter :: nprocx=2,nprocy=3\ninteger, parameter :: nyloc=ny/nprocy,nxloc=nx/nprocx\n")
    s = read_resolution(tmp_path / "resolution.in")
    assert (s["nw"], s["nx"], s["ny"], s["nz"], s["nprocz"], s["nxloc"], s["nyloc"]) == (3, 8, 12, 36, 1, 4, 4)


def test_missing_rho_floor_and_files(tmp_path):
    from ocotillo_amd.athena import AthenaRun
    src = os.path.join(GOLDEN, "samples", "ReadFromAthena")
    text = open(os.path.join(src, "input.in")).read()
    (tmp_path / "resolution.in").write_text(open(os.path.join(src, "resolution.in")).read())
    (tmp_path / "input.in").write_text(text.replace("rho_floor=1e-15\n/", "/").replace(
        "  rho_floor=1e-15\n/", "/"))
    if "rho_floor" not in open(tmp_path / "input.in").read().split("&athena_input")[1]:
        with pytest.raises(ValueError, match="rho_floor"):
            AthenaRun(str(tmp_path))
    (tmp_path / "input.in").write_text(text)
    r = AthenaRun(str(tmp_path))
    with pytest.raises(FileNotFoundError):
        r.block_layout(0, 0)
