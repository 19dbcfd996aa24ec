"""Pins the CPU oracle to the real reference.

The fixtures under tests/golden/ were written by tests/golden/make_golden.py
from the reference compiled unchanged (oracle/ref/build_ref.sh): full FP64
alpha, U, V and every per-point intermediate.  The glibc build of the oracle
uses the same libm as the reference, so it must agree bit for bit; the
portable-math build (the GPU's elementary functions) must agree to 1e-12.
"""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, flux_err, golden, relerr
from oracle.oracle import Oracle

DUMPS = sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "dump_*.npz")))
RUNNABLE = [d for d in DUMPS if "ref_stop_stage" not in golden(d)]


def run_dump(g, portable):
    o = Oracle(g["waves"], portable=portable)
    ncol, nz = int(g["ncol"]), int(g["nz"])
    U, V, A, d = o.solve_block(g["rho"].reshape(1, ncol, nz), g["T"].reshape(1, ncol, nz), float(g["dz"]),
                               float(g["fully_ionized_T"]), bool(g["lgrey"]), float(g["sigma_grey"]))
    # oracle arrays are [iw, ix=0, col, z]; fixtures are [col, iw, z]
    return {k: v[:, 0].transpose(1, 0, 2) for k, v in (("U", U), ("V", V), ("alpha", A))}, d


@pytest.mark.parametrize("name", RUNNABLE)
def test_glibc_oracle_bit_exact(name):
    g = golden(name)
    out, d = run_dump(g, portable=False)
    assert d.status == 0
    for k in ("alpha", "U", "V"):
        assert np.array_equal(out[k], g[k]), (k, int((out[k] != g[k]).sum()))


@pytest.mark.parametrize("name", RUNNABLE)
def test_portable_oracle_within_1e12(name):
    g = golden(name)
    out, d = run_dump(g, portable=True)
    assert d.status == 0
    for k in ("alpha", "U", "V"):
        err = flux_err(out[k], g[k]) if k == "V" else relerr(out[k], g[k])
        assert err.max() <= 1e-12, (k, err.max())


@pytest.mark.parametrize("name", [d for d in RUNNABLE if "albedo" in golden(d)])
def test_intermediates_bit_exact(name):
    """Per-cell state and per-point albedo/S/a/b/c/d of the glibc oracle."""
    g = golden(name)
    o = Oracle(g["waves"])
    ncol, nz, nw = int(g["ncol"]), int(g["nz"]), int(g["nw"])
    fion = float(g["fully_ionized_T"])
    for col in range(ncol):
        cells = [o.cell(g["rho"][col, i], g["T"][col, i], fion) for i in range(nz)]
        for key, attr in (("x", "x"), ("fn", "fn"), ("e_scatter", "es"), ("hm_bf", "hmbf"),
                          ("theta", "theta"), ("lgtheta", "lgtheta")):
            assert np.array_equal([getattr(c, attr) for c in cells], g[key][col]), (col, key)
        for iw in range(nw):
            pts = [o.point(c, iw, bool(g["lgrey"]), float(g["sigma_grey"])) for c in cells]
            assert np.array_equal([p.albedo for p in pts], g["albedo"][col, iw])
            assert np.array_equal([p.S for p in pts], g["S"][col, iw])


@pytest.mark.parametrize("name", ["dump_pivot0.npz", "dump_pivot1.npz"])
def test_zero_pivot_matches_reference_stop(name):
    """The reference stops on an exactly-zero pivot, in stage 2 on the thin
    floor of dump_pivot0 (source/auxiliary.f90:36-39) and in stage 1 where
    alpha^2 underflows at the boundary row of dump_pivot1 (:27-30); the
    oracle reports the same stage and system."""
    g = golden(name)
    o = Oracle(g["waves"])
    ncol, nz = int(g["ncol"]), int(g["nz"])
    _, _, _, d = o.solve_block(g["rho"].reshape(1, ncol, nz), g["T"].reshape(1, ncol, nz), float(g["dz"]),
                               lgrey=bool(g["lgrey"]), sigma_grey=float(g["sigma_grey"]))
    if name == "dump_pivot1.npz":
        assert d.status == 1 and d.fail_j == 1
    assert d.status == int(g["ref_stop_stage"])
    assert (d.fail_ix, d.fail_iy, d.fail_iw) == (1, int(g["ref_stop_col"]), int(g["ref_stop_iw"]))
    assert "ERROR tridiag stage %d" % d.status in str(g["ref_stdout"])
