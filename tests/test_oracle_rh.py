"""The oracle's difference-form (Rybicki-Hummer) elimination, the reference
for the mixed mode's solve, pinned against exact rational arithmetic.

The reference's rows b = -((a + c) + zeta) lose zeta entirely when
zeta < eps (a + c) (optically thin cells); the difference form keeps the
row excess H = zeta separately and never subtracts, so it solves the same
discrete problem to ~n eps.  Checks:
  * oo_tridag_rh against an exact Fraction solve of random systems whose H
    spans 30 decades below A + C;
  * oo_rh_coefficients against the reference's rows (A = a, C = c, R = -d,
    -(A + C + H) = b) on real opacity values;
  * a thin-floor column: the difference form is exact to 1e-12 where the
    reference's elimination is off by more than 1e-6.
"""
import ctypes as C
from fractions import Fraction as F

import numpy as np
import pytest


def exact_rh(A, Cc, H, R):
    """Exact solution of -A x[i-1] + (A+C+H) x[i] - C x[i+1] = R."""
    n = len(H)
    Af, Cf, Hf, Rf = ([F(float(v)) for v in arr] for arr in (A, Cc, H, R))
    cp = [F(0)] * n
    dp = [F(0)] * n
    bet = Af[0] + Cf[0] + Hf[0]
    dp[0] = Rf[0] / bet
    for j in range(1, n):
        cp[j] = Cf[j - 1] / bet
        bet = (Af[j] + Cf[j] + Hf[j]) - Af[j] * cp[j]
        dp[j] = (Rf[j] + Af[j] * dp[j - 1]) / bet
    x = [F(0)] * n
    x[-1] = dp[-1]
    for j in range(n - 2, -1, -1):
        x[j] = dp[j] + cp[j + 1] * x[j + 1]
    return np.array([float(v) for v in x])


def lib():
    from oracle.oracle import lib as L
    return L(False)


@pytest.mark.parametrize("n", [4, 5, 17, 64, 200])
def test_rh_elimination_exact(n):
    rng = np.random.default_rng(n)
    A = 10.0 ** rng.uniform(-3, 3, n)
    Cc = 10.0 ** rng.uniform(-3, 3, n)
    H = (A + Cc) * 10.0 ** rng.uniform(-30, 1, n)
    R = 10.0 ** rng.uniform(-5, 5, n)
    A[0] = 0.0
    Cc[-1] = 0.0
    u = np.zeros(n)
    D = np.zeros(n)
    lib().oo_tridag_rh(n, A, Cc, H, R, u, D)
    x = exact_rh(A, Cc, H, R)
    assert (np.abs(u - x) / np.abs(x)).max() <= 4 * n * 2.0 ** -53


def column_points(nz=64, iw_list=(0, 60, 100, 127)):
    """Opacity values of one column of the mixed-precision snapshot (thin
    2000 K / 1e-15 floor at both ends, thick middle)."""
    from oracle.oracle import Oracle
    from ocotillo_amd.workloads import config4
    rho, T, dz, waves = config4(nx=8, ny=8, nz=nz)
    o = Oracle(waves)
    out = []
    for iw in iw_list:
        pts = [o.point(o.cell(rho[3, 4, z], T[3, 4, z]), iw) for z in range(nz)]
        al = np.array([p.alpha for p in pts])
        out.append((al, np.array([p.alpha1 for p in pts]), np.array([p.alpha2 for p in pts]),
                    np.array([p.albedo for p in pts]), np.array([p.S for p in pts])))
    return out, dz


def test_rh_rows_match_reference_rows():
    L = lib()
    pts, dz = column_points()
    for al, al1, al2, om, S in pts:
        n = al.size
        a, b, c, d = (np.zeros(n) for _ in range(4))
        A, Cc, H, R = (np.zeros(n) for _ in range(4))
        L.oo_tridag_coefficients(n, al, al1, al2, om, S, dz, dz * dz, a, b, c, d)
        L.oo_rh_coefficients(n, al, om, S, dz, dz * dz, A, Cc, H, R)
        assert np.array_equal(A, a) and np.array_equal(Cc, c)
        assert np.array_equal(R[1:-1], -d[1:-1]) and np.array_equal(-((A + Cc) + H)[1:-1], b[1:-1])
        for i in (0, n - 1):   # boundary rows: alpha^2 (1/alpha) vs alpha, one rounding apart
            assert abs(R[i] + d[i]) <= 1e-15 * abs(d[i])
            assert abs((A[i] + Cc[i] + H[i]) + b[i]) <= 4e-16 * abs(b[i])


def test_rh_exact_on_thin_column():
    L = lib()
    pts, dz = column_points()
    worst_ref = 0.0
    for al, al1, al2, om, S in pts:
        n = al.size
        A, Cc, H, R = (np.zeros(n) for _ in range(4))
        L.oo_rh_coefficients(n, al, om, S, dz, dz * dz, A, Cc, H, R)
        x = exact_rh(A, Cc, H, R)
        u = np.zeros(n)
        D = np.zeros(n)
        L.oo_tridag_rh(n, A, Cc, H, R, u, D)
        assert (np.abs(u - x) / np.abs(x)).max() <= 1e-12
        a, b, c, d, gam = (np.zeros(n) for _ in range(5))
        L.oo_tridag_coefficients(n, al, al1, al2, om, S, dz, dz * dz, a, b, c, d)
        ur = np.zeros(n)
        jf = C.c_int(0)
        assert L.oo_tridag(n, a, b, c, d, ur, gam, C.byref(jf)) == 0
        worst_ref = max(worst_ref, (np.abs(ur - x) / np.abs(x)).max())
    assert worst_ref > 1e-6   # the reference's elimination loses zeta on the floor
