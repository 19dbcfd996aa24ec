"""Serving loop (oco_server_*): continuous batching of jobs of different
shapes and wavelength grids, HBM-budget admission, emergent-flux outputs.

Bars:
  * a job's result is bit-identical alone and in a mix of concurrently
    submitted jobs (every system is solved by one lane, whatever else shares
    the launch);
  * FP64: the emergent flux V(1), V(nz) equals the block path's V[..., 0]
    and V[..., -1] and the oracle's, bit for bit;
  * mixed: the emergent flux is within 1e-6 of the exact solution of the
    reference's discrete problem (binary128 oracle) up to nz = 2048, and
    agrees with the block path (whose rows, reduced system, back-substitution
    and boundary flux the serving kernel shares) within twice that;
  * admission keeps the job arena within its budget; a job larger than the
    budget is refused at submit; the reference's zero-pivot stop is reported
    per job with the reference's (ix, iy, iw).
"""

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

GiB = 1 << 30


@pytest.fixture(scope="module")
def ctx():
    import ocotillo_amd as oco
    with oco.Context(0) as c:
        c.set_physics(2e4, False, 0.4)
        yield c


def make_job(seed, nx, ny, nz, nw):
    from ocotillo_amd.workloads import disk_columns, log_waves
    rho, T, dz = disk_columns(nx, ny, nz, seed=seed)
    return rho, T, dz, log_waves(nw)


JOBS = [(11, 3, 2, 96, 70), (12, 2, 2, 300, 16), (13, 4, 3, 64, 130), (14, 1, 5, 513, 40),
        (15, 2, 3, 200, 200), (16, 5, 2, 128, 64), (17, 1, 1, 2048, 33), (18, 3, 3, 150, 90)]


@pytest.mark.parametrize("mode_name", ["mixed", "fp64"])
def test_alone_equals_in_the_mix(ctx, mode_name):
    import ocotillo_amd as oco
    mode = oco.MODE_MIXED if mode_name == "mixed" else oco.MODE_FP64
    probe = make_job(*JOBS[0])
    print("case alone", mode_name, flush=True)
    with oco.Server(ctx, 1 * GiB, max_nz=2048, mode=mode) as srv:
        alone, r = srv.wait(srv.submit(*probe))
        assert r.status == 0
    print("case mix", mode_name, flush=True)
    with oco.Server(ctx, 1 * GiB, max_nz=2048, mode=mode) as srv:
        tickets = [srv.submit(*make_job(*j)) for j in JOBS[1:4]]
        tp = srv.submit(*probe)
        tickets += [srv.submit(*make_job(*j)) for j in JOBS[4:]]
        mixed_out, r = srv.wait(tp)
        others = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert r.status == 0 and all(o[1].status == 0 for o in others)
    assert np.array_equal(alone, mixed_out)
    assert st.completed == len(JOBS) and st.iterations >= 1
    print("  iterations %d, p50 %.1f ms, p99 %.1f ms" % (st.iterations, st.p50_ms, st.p99_ms), flush=True)


def test_emergent_equals_block_path_fp64():
    """FP64 serving: V(1), V(nz) equal the block sweep's and the oracle's, bit
    for bit."""
    import ocotillo_amd as oco
    from oracle.oracle import Oracle
    jobs = [make_job(*j) for j in JOBS[:3]]
    with oco.Context(0) as c:
        c.set_physics(2e4, False, 0.4)
        with oco.Server(c, 1 * GiB, max_nz=1024, mode=oco.MODE_FP64) as srv:
            outs = [srv.wait(srv.submit(*j)) for j in jobs]
        for (rho, T, dz, waves), (em, r) in zip(jobs, outs):
            print("case", rho.shape, waves.size, flush=True)
            assert r.status == 0
            c.set_wavelengths(waves)
            U, V, A, d = c.solve_block(rho, T, dz, mode=oco.MODE_FP64)
            assert np.array_equal(em[..., 0], V[..., 0]) and np.array_equal(em[..., 1], V[..., -1])
            _, Vo, _, _ = Oracle(waves, portable=True).solve_block(rho, T, dz)
            assert np.array_equal(em[..., 0], Vo[..., 0]) and np.array_equal(em[..., 1], Vo[..., -1])


def check_emergent(em, rho, T, dz, waves, label):
    """Mixed emergent V(1), V(nz) within 1e-6 of the exact solution of the
    reference's discrete problem at every point (the block modes' bound)."""
    from oracle.oracle import Oracle
    from test_gpu_mixed import FLUX_TOL
    _, Ve, _, d = Oracle(waves, portable=True).solve_block(rho, T, dz, exact=True)
    assert d.status == 0
    err = 0.0
    for i, k in ((0, 0), (1, -1)):
        e = np.abs(em[..., i] - Ve[..., k]) / np.abs(Ve[..., k])
        err = max(err, float(e.max()))
    print("  %s: emergent V vs exact %.2e" % (label, err), flush=True)
    assert err <= FLUX_TOL, label
    return err


def test_emergent_equals_block_path_mixed():
    """Mixed serving against the block path: both meet the bound against the
    exact solution, so they agree to within twice it.  They are not compared
    bit for bit: the serving kernel partitions a column into 64 chunks of at
    least 5 rows, the block sweep into its own lanes per system and rows."""
    import ocotillo_amd as oco
    from test_gpu_mixed import FLUX_TOL
    jobs = [make_job(*j) for j in JOBS[:3]] + [make_job(19, 2, 1, 1000, 9), make_job(20, 1, 2, 7, 5)]
    with oco.Context(0) as c:
        c.set_physics(2e4, False, 0.4)
        with oco.Server(c, 1 * GiB, max_nz=1024, mode=oco.MODE_MIXED) as srv:
            outs = [srv.wait(srv.submit(*j)) for j in jobs]
        for (rho, T, dz, waves), (em, r) in zip(jobs, outs):
            assert r.status == 0
            c.set_wavelengths(waves)
            U, V, A, d = c.solve_block(rho, T, dz, mode=oco.MODE_MIXED)
            blk = np.stack([V[..., 0], V[..., -1]], -1)
            check_emergent(em, rho, T, dz, waves, "served")
            check_emergent(blk, rho, T, dz, waves, "block")
            same = (em == blk).mean()
            rel = np.abs(em - blk) / np.abs(blk)
            print("case", rho.shape, waves.size, "bit-identical %.4f, max rel diff %.2e" % (same, rel.max()),
                  flush=True)
            assert rel.max() <= 2 * FLUX_TOL


def test_mixed_long_columns_within_bounds(ctx):
    """Mixed serving of columns beyond the block sweep (nz 1500 and 2048, 33
    rows per lane): emergent V within 1e-6 of the exact solution."""
    import ocotillo_amd as oco
    jobs = [make_job(21, 2, 1, 1500, 12), make_job(22, 1, 1, 2048, 7)]
    with oco.Server(ctx, 1 * GiB, max_nz=2048, mode=oco.MODE_MIXED) as srv:
        outs = [srv.wait(srv.submit(*j)) for j in jobs]
    for (rho, T, dz, waves), (em, r) in zip(jobs, outs):
        print("case", rho.shape, waves.size, flush=True)
        assert r.status == 0 and np.isfinite(em).all()
        check_emergent(em, rho, T, dz, waves, "served")


def test_admission_within_budget(ctx):
    """A budget that holds a few jobs at a time: all complete, the arena's
    high-water mark stays within the budget, later jobs queue."""
    import ocotillo_amd as oco
    # a mixed job: 16 columns x 256 cells -> inputs 2 x 32 KiB, cell records
    # in the serving layout (5 rows per lane: 5 x 64 x 22 doubles a column,
    # 880 KiB), emergent flux 16 KiB; the arena holds two of them beside the
    # table
    job_bytes = 2 * 32768 + 16 * 5 * 64 * 22 * 8 + 16 * 64 * 2 * 8
    budget = 2 * job_bytes + (64 << 10)
    jobs = [make_job(100 + i, 4, 4, 256, 64) for i in range(10)]
    with oco.Server(ctx, max(budget, 1 << 20), max_nz=256, mode=oco.MODE_MIXED) as srv:
        budget = srv.stats().hbm_budget
        tickets = [srv.submit(*j) for j in jobs]
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert all(r[1].status == 0 for r in res)
    assert job_bytes <= st.hbm_peak <= budget
    assert st.hbm_peak < 3 * job_bytes + (64 << 10)   # never three admitted at once
    print("  peak %.2f MiB of %.2f MiB" % (st.hbm_peak / 2 ** 20, budget / 2 ** 20), flush=True)


def test_oversize_job_refused(ctx):
    import ocotillo_amd as oco
    with oco.Server(ctx, 4 << 20, max_nz=256, mode=oco.MODE_MIXED) as srv:
        with pytest.raises(oco.OcoError) as e:
            srv.submit(*make_job(1, 16, 16, 256, 64))
        assert e.value.status == oco.E_OOM


def test_pivot_reported_per_job(ctx):
    """The reference stops at its first zero pivot; in FP64 mode the job
    reports that system, other jobs are unaffected."""
    import ocotillo_amd as oco
    g = golden("dump_pivot0.npz")
    ncol, nz = int(g["ncol"]), int(g["nz"])
    rho, T = g["rho"].reshape(1, ncol, nz), g["T"].reshape(1, ncol, nz)
    with oco.Server(ctx, 1 * GiB, max_nz=nz, mode=oco.MODE_FP64) as srv:
        t_ok = srv.submit(*make_job(*JOBS[0]))
        t_bad = srv.submit(rho, T, float(g["dz"]), g["waves"])
        _, r_ok = srv.wait(t_ok)
        _, r_bad = srv.wait(t_bad, raise_on_pivot=False)
    assert r_ok.status == 0
    assert r_bad.status == int(g["ref_stop_stage"])
    assert (r_bad.fail_ix, r_bad.fail_iy, r_bad.fail_iw) == (1, int(g["ref_stop_col"]), int(g["ref_stop_iw"]))


def test_tight_budget_refused_at_submit(ctx):
    """A job whose admission needs more than the budget is refused at
    submit, sized as admission sizes it: 64 columns of 64 cells in mixed mode
    take 3.4 MB of cell records in the serving layout (5 rows per lane), so a
    1 MiB budget refuses them (and does not queue a job that could never be
    admitted)."""
    import ocotillo_amd as oco
    with oco.Server(ctx, 1 << 20, max_nz=256, mode=oco.MODE_MIXED) as srv:
        with pytest.raises(oco.OcoError) as e:
            srv.submit(*make_job(31, 8, 8, 64, 16))
        assert e.value.status == oco.E_OOM


def test_two_grids_in_a_one_job_budget(ctx):
    """Two jobs on different wavelength grids in a budget that holds one job
    and one grid's table (2 x 2 columns, nz 256, nw 2048: 319 KB of buffers
    and a 377 KB table each, 1 MiB budget): the second waits for the first,
    whose table is then freed to make room for its own; both complete with
    the bits they have alone."""
    import ocotillo_amd as oco
    from ocotillo_amd.workloads import disk_columns, log_waves
    jobs = []
    for seed, w0 in ((41, 100.0), (42, 120.0)):
        rho, T, dz = disk_columns(2, 2, 256, seed=seed)
        jobs.append((rho, T, dz, log_waves(2048, w0, 1e5)))
    alone = []
    for j in jobs:
        with oco.Server(ctx, 1 << 30, max_nz=256, mode=oco.MODE_MIXED) as srv:
            em, r = srv.wait(srv.submit(*j))
            assert r.status == 0
            alone.append(em)
    with oco.Server(ctx, 1 << 20, max_nz=256, mode=oco.MODE_MIXED) as srv:
        tickets = [srv.submit(*j) for j in jobs]
        outs = [srv.wait(t) for t in tickets]
        st = srv.stats()
    for (em, r), a in zip(outs, alone):
        assert r.status == 0
        assert np.array_equal(em, a)
    assert st.hbm_peak <= st.hbm_budget
