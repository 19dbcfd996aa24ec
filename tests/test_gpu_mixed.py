"""Difference-form modes: OCO_MODE_MIXED and OCO_MODE_ACCURATE.

Both solve the reference's rows (source/auxiliary.f90:147-197) in the
Rybicki-Hummer difference form (-A x[i-1] + (A+C+H) x[i] - C x[i+1] = R, all
terms >= 0), which keeps the row excess H = zeta that the reference's
b = -((a + c) + zeta) rounds away in optically thin cells, and form the
emergent flux V(1), V(nz) from the rows' interface differences rather than
from differences of U (source/auxiliary.f90:64-78, 96-109).  MIXED evaluates
the opacity's exponentials and logarithms with FP32 cores (~1.2e-7 each);
ACCURATE uses the FP64 drop-in's opacity, alpha bit for bit.

Reference: the exact solution of the reference's discrete problem (the
oracle's binary128 evaluation of the reference's rows, Thomas, ghosts, der
and flux, on the FP64 opacity: Oracle.solve_block(exact=True), pinned to
Fractions in test_oracle_exact.py).  Stated bounds, with no free factor:
  MIXED     alpha <= 1e-6 relative to the FP64 alpha at every point;
            U <= 1e-6 relative to the exact U at every point;
            emergent V(1), V(nz) <= 1e-6 relative to the exact V at every point;
            interior V within 1e-4 of each system's max |V|, or within the
            FP64 drop-in's own largest interior error on the same columns
            where that is larger.
  ACCURATE  alpha equal to the FP64 drop-in's; U <= 1e-12 and emergent
            V <= 1e-8 relative to the exact ones at every point; interior V
            within 1e-5 of max |V|, or the drop-in's own largest error.
Interior V is der of U (source/auxiliary.f90:64-78) as in the reference; in
an optically thin cell the U that it differences agree to ~13 digits, so
their errors (FP32-core opacity in MIXED, the last bit of U in both modes
and in the drop-in) are amplified by about 1 / (alpha dz) there.  Measured:
mixed interior V up to 1.7e-3 of max |V| on a 1088-cell disk column whose
drop-in interior V is off by 1.0e-2 (emergent by 1.7e-2); 2.7e-8 on the C4
slices.
  Against the FP64 drop-in, the comparison the headline metric names:
            |V_mixed - V_fp64| <= 1e-6 |V_fp64| + |V_fp64 - V_exact| at every
            emergent point, i.e. within 1e-6 of the drop-in up to the
            drop-in's own measured error.  On the snapshot's thin floor
            the drop-in (the reference's arithmetic) is off from the exact
            answer by up to 8.6e-3, so only about a third of the emergent
            points are within 1e-6 of it; that fraction is printed.
The batched mixed path returns the block path's bits for every column.
"""
import os

import numpy as np
import pytest

from conftest import golden, relerr

pytestmark = pytest.mark.gpu

ALPHA_TOL = 1e-6
U_TOL = 1e-6
FLUX_TOL = 1e-6
INTERIOR_TOL = 1e-4
ACC_U_TOL = 1e-12
ACC_FLUX_TOL = 1e-8
ACC_INTERIOR_TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    import ocotillo_amd as oco
    with oco.Context(0) as c:
        yield c


def env_context(lanes=None, group=None):
    """A context whose difference-form sweep takes `lanes` lanes per system
    and `group` wavefronts per workgroup (None: the library's choice)."""
    import ocotillo_amd as oco
    env = {}
    if lanes is not None:
        env["OCO_WAVE_LANES"] = str(lanes)
    if group is not None:
        env["OCO_WAVE_GROUP"] = str(group)
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return oco.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def solve_modes(c, rho, T, dz, waves, modes, fully_ionized_T=2e4, lgrey=False):
    import ocotillo_amd as oco
    c.set_wavelengths(waves)
    c.set_physics(fully_ionized_T, lgrey, 0.4)
    out = {}
    for name in modes:
        mode = {"fp64": oco.MODE_FP64, "mixed": oco.MODE_MIXED, "accurate": oco.MODE_ACCURATE}[name]
        U, V, A, d = c.solve_block(rho, T, dz, mode=mode)
        assert d.status == 0, (name, d.status)
        out[name] = (U, V, A)
    return out


def exact(rho, T, dz, waves, fully_ionized_T=2e4, lgrey=False):
    """The exact solution on the kernels' FP64 opacity (portable elementary
    functions, the same alpha as the drop-in's)."""
    from oracle.oracle import Oracle
    U, V, A, d = Oracle(waves, portable=True).solve_block(rho, T, dz, fully_ionized_T, lgrey, exact=True)
    assert d.status == 0
    return U, V, A


def emergent(V):
    return V[..., [0, -1]]


def interior_err(V, Ve):
    """Per system: the largest interior |V - V_exact| over the system's max |V_exact|."""
    sc = np.abs(Ve).max(axis=-1, keepdims=True)
    return (np.abs(V - Ve) / sc)[..., 1:-1].max(axis=-1)


def check_interior(V, Ve, V64, tol, label):
    ei = interior_err(V, Ve)
    e64 = interior_err(V64, Ve)
    print("    %s interior V: max %.2e (FP64 drop-in's own: %.2e)" % (label, ei.max(), e64.max()), flush=True)
    assert ei.max() <= max(tol, e64.max())


def check_mixed(U, V, A, ex, A64, label, V64=None):
    Ue, Ve, Ae = ex
    ea = relerr(A, A64).max()
    eu = relerr(U, Ue).max()
    ev = relerr(emergent(V), emergent(Ve))
    print("  %s vs exact: alpha %.2e  U %.2e  emergent V %.2e (frac <= 1e-6: %.6f)"
          % (label, ea, eu, ev.max(), (ev <= FLUX_TOL).mean()), flush=True)
    assert ea <= ALPHA_TOL
    assert eu <= U_TOL
    assert ev.max() <= FLUX_TOL
    if V64 is not None:
        check_interior(V, Ve, V64, INTERIOR_TOL, label)


def check_accurate(U, V, A, ex, A64, label, V64=None):
    Ue, Ve, Ae = ex
    assert np.array_equal(A, A64)
    eu = relerr(U, Ue).max()
    ev = relerr(emergent(V), emergent(Ve)).max()
    print("  %s vs exact: U %.2e  emergent V %.2e" % (label, eu, ev), flush=True)
    assert eu <= ACC_U_TOL
    assert ev <= ACC_FLUX_TOL
    if V64 is not None:
        check_interior(V, Ve, V64, ACC_INTERIOR_TOL, label)


def check_against_fp64(Vm, V64, Ve, label):
    """|V_mixed - V_fp64| <= 1e-6 |V_fp64| + |V_fp64 - V_exact| at every
    emergent point; prints the fractions within 1e-6 of the drop-in and of
    the exact answer."""
    vm, vf, ve = emergent(Vm), emergent(V64), emergent(Ve)
    ok = np.abs(vm - vf) <= FLUX_TOL * np.abs(vf) + np.abs(vf - ve)
    print("  %s: emergent points within 1e-6 of FP64 %.4f; FP64 within 1e-6 of exact %.4f; criterion holds %.6f"
          % (label, (relerr(vm, vf) <= FLUX_TOL).mean(), (relerr(vf, ve) <= FLUX_TOL).mean(), ok.mean()),
          flush=True)
    assert ok.all()


def test_mixed_alpha_goldens(ctx):
    import ocotillo_amd as oco
    for name in ("dump_mix.npz", "dump_saha.npz", "dump_saha_lowfion.npz", "dump_c2cols.npz"):
        print("case", name, flush=True)
        g = golden(name)
        ncol, nz = int(g["ncol"]), int(g["nz"])
        ctx.set_wavelengths(g["waves"])
        ctx.set_physics(float(g["fully_ionized_T"]), False, 0.4)
        rho, T = g["rho"].reshape(1, ncol, nz), g["T"].reshape(1, ncol, nz)
        _, _, A64, d0 = ctx.solve_block(rho, T, float(g["dz"]), mode=oco.MODE_FP64)
        _, _, Am, d1 = ctx.solve_block(rho, T, float(g["dz"]), mode=oco.MODE_MIXED)
        _, _, Aa, d2 = ctx.solve_block(rho, T, float(g["dz"]), mode=oco.MODE_ACCURATE)
        assert d0.status == 0 and d1.status == 0 and d2.status == 0
        err = relerr(Am, A64).max()
        print("  alpha max rel err %.3e" % err, flush=True)
        assert err <= ALPHA_TOL
        assert np.array_equal(Aa, A64)


def c4_slice(x0, y0, n):
    from ocotillo_amd.workloads import config4
    rho, T, dz, waves = config4(nx=64, ny=64)
    return rho[x0:x0 + n, y0:y0 + n].copy(), T[x0:x0 + n, y0:y0 + n].copy(), dz, waves


@pytest.mark.parametrize("where", [(0, 0), (30, 30)])
def test_c4_slice(ctx, where):
    """The headline configuration's columns (nz 512, nw 128), all three modes."""
    rho, T, dz, waves = c4_slice(where[0], where[1], 4)
    print("case c4 slice", where, rho.shape, flush=True)
    out = solve_modes(ctx, rho, T, dz, waves, ("fp64", "mixed", "accurate"))
    ex = exact(rho, T, dz, waves)
    A64 = out["fp64"][2]
    check_mixed(*out["mixed"], ex, A64, "mixed", out["fp64"][1])
    check_accurate(*out["accurate"], ex, A64, "accurate", out["fp64"][1])
    check_against_fp64(out["mixed"][1], out["fp64"][1], ex[1], "mixed vs FP64")


def test_c2_slice(ctx):
    from ocotillo_amd.workloads import config2
    rho, T, dz, waves = config2(4, 4, 256, 128)
    print("case c2 slice", rho.shape, flush=True)
    out = solve_modes(ctx, rho, T, dz, waves, ("fp64", "mixed", "accurate"))
    ex = exact(rho, T, dz, waves)
    check_mixed(*out["mixed"], ex, out["fp64"][2], "mixed", out["fp64"][1])
    check_accurate(*out["accurate"], ex, out["fp64"][2], "accurate", out["fp64"][1])
    check_against_fp64(out["mixed"][1], out["fp64"][1], ex[1], "mixed vs FP64")


def lanes_fitting(nz):
    """Lanes per system whose chunks fit: at most 9 rows per lane below 64
    lanes, 17 with 64 (rows per lane: ceil(nz / L), made odd)."""
    return [L for L in (8, 16, 32, 64) if (((nz + L - 1) // L) | 1) <= (17 if L == 64 else 9)]


@pytest.mark.parametrize("nz", [4, 5, 32, 63, 64, 72, 73, 100, 144, 145, 200, 288, 289, 512, 513, 1088])
def test_lanes_per_system(nz):
    """Every lanes-per-system variant that takes nz (8, 16, 32, 64 lanes;
    1 to 17 rows per lane) meets the bounds against the exact solution, in
    both modes; the alpha of the variants is the same bits."""
    from ocotillo_amd.workloads import config2
    rho, T, dz, waves = config2(2, 2, nz, 24)
    ex = exact(rho, T, dz, waves)
    alphas = []
    for L in lanes_fitting(nz):
        print("case nz", nz, "lanes", L, flush=True)
        with env_context(lanes=L) as c:
            out = solve_modes(c, rho, T, dz, waves, ("fp64", "mixed", "accurate"))
        check_mixed(*out["mixed"], ex, out["fp64"][2], "mixed L=%d" % L, out["fp64"][1])
        check_accurate(*out["accurate"], ex, out["fp64"][2], "accurate L=%d" % L, out["fp64"][1])
        alphas.append(out["mixed"][2])
    for a in alphas[1:]:
        assert np.array_equal(a, alphas[0])


def test_grey_difference_form(ctx):
    """Grey columns (source/continuous_opacity.f90:377-390) in both
    difference-form modes: the grey point is FP64 in both, alpha equal to the
    drop-in's; on the grey isothermal sample's thin floor (rho floor 1e-24)
    and on a resolved column."""
    from ocotillo_amd.workloads import config1
    rho, T, dz, waves = config1()
    cases = [("grey isothermal", rho, T, dz)]
    nz = 64
    z = np.linspace(-1.0, 1.0, nz)
    cases.append(("grey resolved", (1e-10 * np.exp(-z * z)).reshape(1, 1, nz),
                  (6000.0 + 2000.0 * z).reshape(1, 1, nz), 1e11))
    for label, r, t, d in cases:
        print("case", label, flush=True)
        out = solve_modes(ctx, r, t, d, waves, ("fp64", "mixed", "accurate"), lgrey=True)
        ex = exact(r, t, d, waves, lgrey=True)
        for mode in ("mixed", "accurate"):
            check_accurate(*out[mode], ex, out["fp64"][2], label + " " + mode, out["fp64"][1])


def test_long_columns_rejected(ctx):
    """Columns above 1088 cells do not fit the difference-form block sweep:
    OCO_E_ARG (the FP64 drop-in and the server take them)."""
    import ocotillo_amd as oco
    from ocotillo_amd.workloads import config2
    rho, T, dz, waves = config2(1, 1, 1100, 4)
    ctx.set_wavelengths(waves)
    ctx.set_physics(2e4, False, 0.4)
    with pytest.raises(oco.OcoError) as e:
        ctx.solve_block(rho, T, dz, mode=oco.MODE_MIXED)
    assert e.value.status == oco.E_ARG
    U, V, A, d = ctx.solve_block(rho, T, dz, mode=oco.MODE_FP64)
    assert d.status == 0


@pytest.mark.parametrize("group,nx,ny", [(1, 3, 5), (4, 4, 4)])
def test_wave_sweep_queues_cover_every_system(group, nx, ny):
    """The difference-form sweep hands out wavelength groups from work queues
    (one per XCD when its grid is a multiple of 8, else one), so which
    workgroup solves a system varies from call to call; each system is still
    solved once and by itself.  The blocks here (nz 256 on 32 lanes per
    system, 40 wavelengths: with one wavefront per workgroup, 2 systems per
    workgroup and a grid of 300; with four, 8 systems and a grid of 80) equal
    their columns solved one at a time (one queue, or seven empty ones) and
    with 37 wavelengths (one queue), bit for bit, and a repeat of the block
    after those smaller calls equals the first."""
    import ocotillo_amd as oco
    from ocotillo_amd.workloads import config2
    rho, T, dz, waves = config2(nx, ny, 256, 40)
    with env_context(group=group) as c:
        c.set_physics(2e4, False, 0.4)
        c.set_wavelengths(waves)
        U, V, A, d = c.solve_block(rho, T, dz, mode=oco.MODE_MIXED)
        assert d.status == 0 and np.isfinite(U).all() and np.isfinite(V).all()
        for ix in range(nx):
            for iy in range(ny):
                u, v, a, _ = c.solve_block(rho[ix:ix + 1, iy:iy + 1], T[ix:ix + 1, iy:iy + 1], dz,
                                           mode=oco.MODE_MIXED)
                assert np.array_equal(u[:, 0, 0], U[:, ix, iy]), (ix, iy)
                assert np.array_equal(v[:, 0, 0], V[:, ix, iy]), (ix, iy)
                assert np.array_equal(a[:, 0, 0], A[:, ix, iy]), (ix, iy)
        c.set_wavelengths(waves[:37])
        u, v, a, _ = c.solve_block(rho[2:3, 3:4], T[2:3, 3:4], dz, mode=oco.MODE_MIXED)
        assert np.array_equal(u[:, 0, 0], U[:37, 2, 3]) and np.array_equal(v[:, 0, 0], V[:37, 2, 3])
        c.set_wavelengths(waves)
        U2, V2, A2, _ = c.solve_block(rho, T, dz, mode=oco.MODE_MIXED)
        assert np.array_equal(U2, U) and np.array_equal(V2, V) and np.array_equal(A2, A)


@pytest.mark.parametrize("nz,nw", [(4, 5), (65, 40), (100, 37), (256, 37), (512, 130), (1080, 9)])
def test_wave_groups_equal(nz, nw):
    """Workgroups of 1, 2 and 4 wavefronts (each cell record a thread loads
    evaluated for the workgroup's wavelengths; columns below 128 cells split
    the workgroup into parts over the wavelengths) give the same bits: the
    points and each system's solve are the same code whichever workgroup
    runs them.  nw = 5, 37, 130 and 9 leave a last group of fewer
    wavelengths than systems per workgroup."""
    import ocotillo_amd as oco
    from ocotillo_amd.workloads import config2
    rho, T, dz, waves = config2(2, 3, nz, nw)
    print("case nz", nz, "nw", nw, flush=True)
    for mode in (oco.MODE_MIXED, oco.MODE_ACCURATE):
        out = []
        for g in (1, 2, 4):
            with env_context(group=g) as c:
                c.set_wavelengths(waves)
                c.set_physics(2e4, False, 0.4)
                U, V, A, d = c.solve_block(rho, T, dz, mode=mode)
            assert d.status == 0 and np.isfinite(U).all() and np.isfinite(V).all()
            out.append((U, V, A))
        for g, (U, V, A) in zip((2, 4), out[1:]):
            assert np.array_equal(U, out[0][0]), (mode, g)
            assert np.array_equal(V, out[0][1]), (mode, g)
            assert np.array_equal(A, out[0][2]), (mode, g)


@pytest.mark.parametrize("nz", [50, 96, 200, 400, 600, 900])
def test_batch_mixed_equals_block(nz):
    """Batched (persistent, graph-replayed) mixed solve == the block mixed
    solve, bit for bit, on every replay; nz 50, 96, 200, 400, 600 and 900 put
    every column in one of the batched kernel's six classes (8, 16, 32 lanes
    per system up to 72, 144, 288 cells; 64 lanes up to 512, 768, 1088), which
    take the block sweep's lanes per system."""
    import ocotillo_amd as oco
    from ocotillo_amd.workloads import config2
    rho, T, dz, waves = config2(2, 3, nz, 80)
    ncol = 6
    nz, nw = rho.shape[2], waves.size
    with oco.Context(0) as c:
        c.set_wavelengths(waves)
        c.set_physics(2e4, False, 0.4)
        U0, V0, A0, _ = c.solve_block(rho, T, dz, mode=oco.MODE_MIXED)
        b = oco.Batch(c, [nz] * ncol, [nw] * ncol, np.tile(waves, ncol), [dz] * ncol)
        try:
            for rep in range(3):
                U, V, A, d = b.run_host(rho.reshape(-1), T.reshape(-1), mode=oco.MODE_MIXED)
                assert d.status == 0
                Us, Vs, As = b.split(U, V, A)
                for col in range(ncol):
                    ix, iy = divmod(col, rho.shape[1])
                    assert np.array_equal(Us[col], U0[:, ix, iy]), (rep, col, "U")
                    assert np.array_equal(Vs[col], V0[:, ix, iy]), (rep, col, "V")
                    assert np.array_equal(As[col], A0[:, ix, iy]), (rep, col, "A")
                # FP64 run in between: the two modes keep separate graphs
                b.run_host(rho.reshape(-1), T.reshape(-1), mode=oco.MODE_FP64)
            v, nb, tuned = b.tuning(oco.MODE_MIXED)
            assert tuned == 1
        finally:
            b.close()


def test_batch_mixed_uneven_columns():
    """Batched mixed mode on uneven columns (nz 4..1000 across the six
    column classes, nw 1..140; each class's kernel is instantiated for its
    longest column): every column equals that column solved as a block of its
    own, bit for bit, on every graph replay, with the context's own eager work
    between replays, and meets the bounds against the exact solution."""
    import ocotillo_amd as oco
    from ocotillo_amd.workloads import config3
    from oracle.oracle import Oracle
    nz, nw, waves, dz, rho, T = config3(ncol=18, seed=78, nz_range=(4, 1000), nw_range=(1, 140))
    print("batch nz", nz.tolist(), "nw", nw.tolist(), flush=True)
    with oco.Context(0) as c:
        c.set_physics(2e4, False, 0.4)
        b = oco.Batch(c, nz, nw, waves, dz)
        try:
            ref = []
            c0 = w0 = 0
            for n, m, d in zip(nz.tolist(), nw.tolist(), dz.tolist()):
                wv = waves[w0:w0 + m]
                r1, t1 = rho[c0:c0 + n].reshape(1, 1, n), T[c0:c0 + n].reshape(1, 1, n)
                c.set_wavelengths(wv)
                u, v, a, dg = c.solve_block(r1, t1, d, mode=oco.MODE_MIXED)
                _, v64, a64, _ = c.solve_block(r1, t1, d, mode=oco.MODE_FP64)
                assert dg.status == 0
                Ue, Ve, _, _ = Oracle(wv, portable=True).solve_block(r1, t1, d, exact=True)
                check_mixed(u, v, a, (Ue, Ve, None), a64, "column nz %d nw %d" % (n, m), v64)
                ref.append((u[:, 0, 0], v[:, 0, 0], a[:, 0, 0]))
                c0 += n
                w0 += m
            for rep in range(3):
                U, V, A, d = b.run_host(rho, T, mode=oco.MODE_MIXED)
                assert d.status == 0
                Us, Vs, As = b.split(U, V, A)
                for i, (ru, rv, ra) in enumerate(ref):
                    assert np.array_equal(Us[i], ru), (rep, i, "U")
                    assert np.array_equal(Vs[i], rv), (rep, i, "V")
                    assert np.array_equal(As[i], ra), (rep, i, "A")
                c.solve_block(rho[:nz[0]].reshape(1, 1, -1), T[:nz[0]].reshape(1, 1, -1), float(dz[0]),
                              mode=oco.MODE_MIXED)
            v, nb, tuned = b.tuning(oco.MODE_MIXED)
            assert tuned == 1
        finally:
            b.close()
