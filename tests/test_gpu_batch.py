"""Batched uneven columns (oco_batch_*): every column bit-exact against the
oracle built with the kernels' elementary functions, on the autotuned
kernel, on every hipGraph replay (with eager work between replays), and a
zero pivot reported for the first failing column in column order."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def oracle_columns(nz, nw, waves, dz, rho, T, portable=True):
    from oracle.oracle import Oracle
    outs = []
    c0 = w0 = 0
    for n, m, d in zip(nz.tolist(), nw.tolist(), dz.tolist()):
        o = Oracle(waves[w0:w0 + m], portable=portable)
        U, V, A, dg = o.solve_block(rho[c0:c0 + n].reshape(1, 1, n), T[c0:c0 + n].reshape(1, 1, n), d)
        outs.append((U[:, 0, 0], V[:, 0, 0], A[:, 0, 0], dg.status))
        c0 += n
        w0 += m
    return outs


def small_batch():
    from ocotillo_amd.workloads import config3
    nz, nw, waves, dz, rho, T = config3(ncol=14, seed=77, nz_range=(4, 300), nw_range=(1, 140))
    return nz, nw, waves, dz, rho, T


def test_batch_bit_exact_and_graph_replays():
    import ocotillo_amd as oco
    nz, nw, waves, dz, rho, T = small_batch()
    print("batch nz", nz.tolist(), "nw", nw.tolist(), flush=True)
    ref = oracle_columns(nz, nw, waves, dz, rho, T)
    with oco.Context(0) as ctx:
        ctx.set_physics(2e4, False, 0.4)
        b = oco.Batch(ctx, nz, nw, waves, dz)
        other = oco.Context(0)
        other.set_wavelengths(waves[:5])
        for rep in range(4):
            U, V, A, d = b.run_host(rho, T)
            assert d.status == 0
            Us, Vs, As = b.split(U, V, A)
            for c, (rU, rV, rA, st) in enumerate(ref):
                assert st == 0
                assert np.array_equal(Us[c], rU) and np.array_equal(Vs[c], rV) and np.array_equal(As[c], rA), \
                    ("replay", rep, "column", c)
            # the program's own eager work between replays
            other.solve_block(rho[:nz[0]].reshape(1, 1, -1), T[:nz[0]].reshape(1, 1, -1), float(dz[0]))
        v, nb, tuned = b.tuning()
        assert tuned == 1 and nb > 0
        other.close()
        b.close()


def edge_batch():
    """Columns across the uneven-batch configuration's whole range (nz 32 ..
    1024, nw 16 .. 256): the mixed path's column classes' edges 72/73,
    144/145, 288/289, 512/513, 768/769 and 1024, and the 64-wavelength tile
    edges 63, 64, 65, 128, 255 and 256, plus the range's ends; disk profiles
    with the lognormal rho perturbation of the configuration."""
    from ocotillo_amd.workloads import RHO_FLOOR, disk_columns, log_waves
    rng = np.random.default_rng(2026)
    nzs = [32, 72, 73, 144, 145, 288, 289, 512, 513, 768, 769, 1024, 1024, 600]
    nws = [16, 63, 64, 65, 256, 128, 255, 17, 64, 63, 65, 256, 16, 100]
    rhos, Ts, dzs, waves = [], [], [], []
    for n, m in zip(nzs, nws):
        r, t, d = disk_columns(1, 1, n, rng=rng)
        rhos.append(np.maximum(r.reshape(-1) * 10.0 ** rng.normal(0.0, 0.5, n), RHO_FLOOR))
        Ts.append(t.reshape(-1))
        dzs.append(d)
        waves.append(log_waves(m))
    return (np.array(nzs, np.int32), np.array(nws, np.int32), np.concatenate(waves), np.array(dzs),
            np.concatenate(rhos), np.concatenate(Ts))


def test_batch_fp64_whole_range_bit_exact():
    """FP64 batch over the configuration's whole range of nz and nw: every
    column bit-exact against the oracle on the autotuned kernel, on three
    graph replays with eager work between them."""
    import ocotillo_amd as oco
    nz, nw, waves, dz, rho, T = edge_batch()
    print("batch nz", nz.tolist(), "nw", nw.tolist(), flush=True)
    ref = oracle_columns(nz, nw, waves, dz, rho, T)
    with oco.Context(0) as ctx:
        ctx.set_physics(2e4, False, 0.4)
        b = oco.Batch(ctx, nz, nw, waves, dz)
        try:
            for rep in range(3):
                U, V, A, d = b.run_host(rho, T)
                assert d.status == 0
                Us, Vs, As = b.split(U, V, A)
                for c, (rU, rV, rA, st) in enumerate(ref):
                    assert st == 0
                    assert np.array_equal(Us[c], rU) and np.array_equal(Vs[c], rV) and np.array_equal(As[c], rA), \
                        ("replay", rep, "column", c, "nz", int(nz[c]), "nw", int(nw[c]))
                ctx.set_wavelengths(waves[:nw[0]])
                ctx.solve_block(rho[:nz[0]].reshape(1, 1, -1), T[:nz[0]].reshape(1, 1, -1), float(dz[0]))
            v, nb, tuned = b.tuning()
            assert tuned == 1 and nb > 0
        finally:
            b.close()


def test_batch_zero_pivot_first_column():
    import ocotillo_amd as oco
    g = golden("dump_pivot0.npz")
    nz0 = int(g["nz"])
    good_nz, good_nw, good_w, good_dz, good_rho, good_T = small_batch()
    # column 0: a good one; column 1: the column the reference stopped on
    nz = np.array([good_nz[0], nz0, good_nz[1]], np.int32)
    nw = np.array([good_nw[0], int(g["nw"]), good_nw[1]], np.int32)
    waves = np.concatenate([good_w[:good_nw[0]], g["waves"], good_w[good_nw[0]:good_nw[0] + good_nw[1]]])
    dz = np.array([good_dz[0], float(g["dz"]), good_dz[1]])
    bad = int(g["ref_stop_col"]) - 1
    rho = np.concatenate([good_rho[:good_nz[0]], g["rho"][bad], good_rho[good_nz[0]:good_nz[0] + good_nz[1]]])
    T = np.concatenate([good_T[:good_nz[0]], g["T"][bad], good_T[good_nz[0]:good_nz[0] + good_nz[1]]])
    with oco.Context(0) as ctx:
        ctx.set_physics(2e4, False, 0.4)
        b = oco.Batch(ctx, nz, nw, waves, dz)
        U, V, A, d = b.run_host(rho, T, raise_on_pivot=False)
        assert d.status == 2 and d.fail_ix == 2 and d.fail_iw == int(g["ref_stop_iw"])
        ref = oracle_columns(nz, nw, waves, dz, rho, T)
        Us, Vs, As = b.split(U, V, A)
        for c in (0, 2):
            assert np.array_equal(Us[c], ref[c][0]) and np.array_equal(Vs[c], ref[c][1])
        b.close()


@pytest.mark.parametrize("mode_name", ["fp64", "mixed"])
def test_device_rows_padded(mode_name):
    """The device outputs hold each system's row padded to a multiple of 16
    doubles (one 128-byte line) at the column's offset (oco_batch_column):
    the same values as run_host's packed arrays, columns back to back."""
    import ctypes as C
    import ocotillo_amd as oco
    mode = oco.MODE_FP64 if mode_name == "fp64" else oco.MODE_MIXED
    nz, nw, waves, dz, rho, T = small_batch()
    with oco.Context(0) as ctx:
        ctx.set_physics(2e4, False, 0.4)
        b = oco.Batch(ctx, nz, nw, waves, dz)
        U, V, A, d = b.run_host(rho, T, mode=mode)
        assert d.status == 0
        Us, Vs, As = b.split(U, V, A)
        hip = C.CDLL("libamdhip64.so.7")   # the runtime the library runs on
        hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _, _, dU, dV, dA = b.buffers()

        def fetch(base, off, n):
            out = np.empty(n, dtype=np.float64)
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(base + 8 * off), 8 * n, 2) == 0
            return out

        next_u = next_v = 0
        for c in range(nz.size):
            uo, lu, vo, lv = b.column(c)
            n, m = int(nz[c]), int(nw[c])
            assert (uo, vo) == (next_u, next_v)
            assert lu % 16 == 0 and n + 6 <= lu < n + 22 and lv % 16 == 0 and n <= lv < n + 16
            for base, off, ld, rows, ref in ((dU, uo, lu, n + 6, Us[c]), (dV, vo, lv, n, Vs[c]), (dA, vo, lv, n, As[c])):
                got = fetch(base, off, ld * m).reshape(m, ld)[:, :rows]
                assert np.array_equal(got, ref), (mode_name, c)   # split gives (nw_c, rows)
            next_u, next_v = uo + lu * m, vo + lv * m
        with pytest.raises(oco.OcoError):
            b.column(nz.size)
        b.close()
