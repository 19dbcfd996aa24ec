"""Athena input construction on the device (oco_athena_slab, _device).

Bars:
  * the device's rho3d, T3d equal the oracle's restatement
    (source/read_athena.f90:163-268, source/disk.f90:96-104) bit for bit, on
    synthetic slabs that fire every branch (density floor, floorTemp, the
    momentum guard on and off, z rows outside the slab untouched) and on the
    reference's two Athena samples;
  * the product path -- AthenaRun.load_block (device construction) then the
    FP64 block solve -- reproduces the compiled reference's sample outputs
    within 1e-12 (its elementary functions are the portable ones, the
    reference's glibc's; the construction itself has none).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, flux_err, golden, relerr
from test_athena import oracle_block, oracle_slab, run, synthetic_slab, synthetic_tgauss, synthetic_units

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import ocotillo_amd as oco
    with oco.Context(0) as c:
        c.set_physics(2e4, False, 0.4)
        yield c


CASES = [(7, 5, 4, 6, 11, 3), (8, 64, 48, 40, 100, 30), (9, 1, 1, 1, 1, 0), (10, 33, 17, 9, 9, 0),
         (11, 3, 200, 70, 300, 230)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "seed%d_%dx%dx%d_nz%d_iz%d" % c)
@pytest.mark.parametrize("guard", [1, 0])
def test_slab_bit_exact(ctx, case, guard):
    seed, nxloc, nyloc, nzl, nz, iz0 = case
    print("case", case, "guard", guard, flush=True)
    f = synthetic_slab(seed, nxloc, nyloc, nzl)
    u = synthetic_units(avoid_momentum_underflow=guard)
    tg = synthetic_tgauss(nz, u.floor_temp)
    rho_o, T_o = oracle_slab(f, tg, u, nz, iz0)
    rho = np.full((nxloc, nyloc, nz), -1.0)
    T = np.full_like(rho, -2.0)
    ctx.athena_slab(f, tg, u, iz0, rho, T)
    assert np.array_equal(rho, rho_o)
    assert np.array_equal(T, T_o)


class DevBuf:
    """Device memory from the HIP runtime the library itself is bound to
    (libamdhip64.so.7, already loaded by the library), not torch's."""

    def __init__(self, host):
        import ctypes as C
        import ocotillo_amd as oco
        oco.load_library()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.host.nbytes)) == 0
        self._copy(self.ptr, self.host.ctypes.data_as(C.c_void_p), 1)

    def _copy(self, dst, src, kind):
        import ctypes as C
        assert self.hip.hipMemcpy(dst, src, C.c_size_t(self.host.nbytes), kind) == 0

    def get(self):
        import ctypes as C
        out = np.empty_like(self.host)
        self._copy(out.ctypes.data_as(C.c_void_p), self.ptr, 2)
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def test_slab_device_pointers(ctx):
    seed, nxloc, nyloc, nzl, nz, iz0 = CASES[1]
    f = synthetic_slab(seed, nxloc, nyloc, nzl)
    u = synthetic_units()
    tg = synthetic_tgauss(nz, u.floor_temp)
    rho_o, T_o = oracle_slab(f, tg, u, nz, iz0)
    bufs = [DevBuf(a) for a in f]
    d_tg = DevBuf(tg)
    d_rho = DevBuf(np.full((nxloc, nyloc, nz), -1.0))
    d_T = DevBuf(np.full((nxloc, nyloc, nz), -2.0))
    try:
        ctx.athena_slab_device(nz, nyloc, nxloc, nzl, iz0, [b.ptr.value for b in bufs], d_tg.ptr.value, u,
                               d_rho.ptr.value, d_T.ptr.value)
        assert np.array_equal(d_rho.get(), rho_o)
        assert np.array_equal(d_T.get(), T_o)
    finally:
        for b in bufs + [d_tg, d_rho, d_T]:
            b.free()


def test_bad_arguments_rejected(ctx):
    import ocotillo_amd as oco
    f = synthetic_slab(7, 5, 4, 6)
    u = synthetic_units()
    tg = synthetic_tgauss(8, u.floor_temp)
    rho = np.zeros((5, 4, 8))
    T = np.zeros_like(rho)
    with pytest.raises(oco.OcoError) as e:      # slab rows 3..8 do not fit nz = 8
        ctx.athena_slab(f, tg, u, 3, rho, T)
    assert e.value.status == oco.E_ARG
    with pytest.raises(oco.OcoError):
        ctx.athena_slab(f, tg, u, -1, rho, T)
    with pytest.raises(oco.OcoError):
        ctx.athena_slab(f, tg[:5], u, 0, rho, T)


@pytest.mark.parametrize("sample", ["ReadFromAthena", "multiInTest"])
def test_sample_blocks(ctx, sample):
    """Device construction == oracle on every block of the sample; then the
    product path's FP64 solve against the reference's outputs."""
    r = run(sample)
    g = golden("sample_%s.npz" % sample)
    ctx.set_wavelengths(g["wavegrid"])
    nz, nw = int(g["nz"]), int(g["nw"])
    cols = g["cols"]
    dz1 = dz2 = None
    for px in range(int(g["nprocx"])):
        for py in range(int(g["nprocy"])):
            print("sample", sample, "block", px, py, flush=True)
            rho, T, z, dz = r.load_block(ctx, px, py)
            rho_o, T_o, z_o, dz_o = oracle_block(r, px, py)
            assert np.array_equal(rho, rho_o) and np.array_equal(T, T_o)
            assert np.array_equal(z, g["zgrid"]) and dz == dz_o
            if dz1 is None:
                dz1, dz2 = 1.0 / dz, dz * dz
            U, V, A, d = ctx.solve_block(rho, T, dz, dz1=dz1, dz2=dz2)
            assert d.status == 0
            ncol = rho.shape[0] * rho.shape[1]
            tag = "p%d_%d" % (px, py)
            assert relerr(A.reshape(nw, ncol, nz)[:, cols], g["A_" + tag]).max() <= 1e-12
            assert relerr(U.reshape(nw, ncol, nz + 6)[:, cols], g["U_" + tag]).max() <= 1e-12
            assert flux_err(V.reshape(nw, ncol, nz)[:, cols], g["V_" + tag]).max() <= 1e-12
