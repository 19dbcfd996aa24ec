"""The oracle's exact solve (oo_exact_system: the reference's rows, Thomas,
ghosts, der and flux in binary128) pinned against exact rational arithmetic
and against the independent FP64 difference-form elimination.

This oracle is the comparator of the accuracy statements: on optically thin
cells the reference's FP64 arithmetic loses zeta in b = -((a + c) + zeta)
(source/auxiliary.f90:162-163) and forms the emergent flux from U values that
agree to ~13 digits (source/auxiliary.f90:74-76, 107), so its own answer
there is off from the exact solution of its discrete problem by up to ~1e-2.
Checks:
  * random systems (alpha dz from 1e-15 to 1e3, albedo in [0, 1)) solved
    with the reference's formulas in Fractions: U and V equal to the
    binary128 result rounded to FP64, within 1 ulp;
  * a thin-floor column of the mixed-precision snapshot in Fractions;
  * U against oo_tridag_rh (FP64, difference form, itself pinned to
    Fractions in test_oracle_rh.py) on snapshot columns: 1e-12;
  * the FP64 reference against it: equal to 1e-12 on grey columns with
    alpha dz from 0.16 to 300, where nothing cancels.
"""
from fractions import Fraction as F

import numpy as np
import pytest


def exact_fraction(al, om, S, dz, dz1, dz2):
    """source/auxiliary.f90:111-201 (rows), :14-47 (Thomas), :49-62 (ghosts),
    :64-78 (der), :96-109 (flux) in exact rational arithmetic on the FP64
    inputs; returns U (nz + 6) and V (nz) rounded to FP64."""
    nz = len(al)
    a_ = [F(float(v)) for v in al]
    w_ = [F(float(v)) for v in om]
    s_ = [F(float(v)) for v in S]
    dz_, dz1_, dz2_ = F(dz), F(dz1), F(dz2)
    a = [F(0)] * nz
    b = [F(0)] * nz
    c = [F(0)] * nz
    d = [F(0)] * nz
    for i in range(1, nz - 1):
        kp = (a_[i + 1] + a_[i]) / 2
        km = (a_[i] + a_[i - 1]) / 2
        a[i] = a_[i] ** 2 / km
        c[i] = a_[i] ** 2 / kp
        zeta = dz2_ * a_[i] ** 3 * (1 - w_[i])
        b[i] = -(a[i] + c[i] + zeta)
        d[i] = -s_[i] * zeta
    for i, top in ((0, False), (nz - 1, True)):
        zeta = a_[i] * dz2_ * (1 - w_[i]) / 4
        a[i] = a_[i] if top else F(0)
        c[i] = F(0) if top else a_[i]
        b[i] = -(1 / a_[i] + dz_ + zeta) * a_[i] ** 2
        d[i] = -s_[i] * zeta * a_[i] ** 2
    gam = [F(0)] * nz
    u = [F(0)] * nz
    bet = b[0]
    u[0] = d[0] / bet
    for j in range(1, nz):
        gam[j] = c[j - 1] / bet
        bet = b[j] - a[j] * gam[j]
        u[j] = (d[j] - a[j] * u[j - 1]) / bet
    for j in range(nz - 2, -1, -1):
        u[j] = u[j] - gam[j + 1] * u[j + 1]
    ext = [F(0)] * 3 + u + [F(0)] * 3
    for i in range(1, 4):
        ext[3 - i] = 2 * ext[3] - ext[3 + i]
        ext[nz + 2 + i] = 2 * ext[nz + 2] - ext[nz + 2 - i]
    V = []
    for i in range(nz):
        f = lambda k: ext[3 + i + k]
        df = (45 * (f(1) - f(-1)) - 9 * (f(2) - f(-2)) + (f(3) - f(-3))) / 60
        V.append(float(-(df / a_[i]) * dz1_))
    return np.array([float(v) for v in ext]), np.array(V)


def lib():
    from oracle.oracle import lib as L
    return L(False)


def exact_oracle(al, om, S, dz, dz1, dz2):
    nz = len(al)
    U = np.zeros(nz + 6)
    V = np.zeros(nz)
    lib().oo_exact_system(nz, np.ascontiguousarray(al), np.ascontiguousarray(om), np.ascontiguousarray(S),
                          dz, dz1, dz2, U, V)
    return U, V


def ulps(x, y):
    x = np.asarray(x)
    y = np.asarray(y)
    return np.where(x == y, 0.0, np.abs(x - y) / np.spacing(np.abs(y)))


def check_V(V, Vf):
    """Emergent V(1), V(nz) within 1 ulp; interior V within 1 ulp of the
    system's max |V| (V crosses zero inside a column, where a point value has
    no scale)."""
    assert ulps(V[[0, -1]], Vf[[0, -1]]).max() <= 1.0
    assert np.abs(V - Vf).max() <= np.spacing(np.abs(Vf).max())


@pytest.mark.parametrize("nz,seed", [(4, 1), (5, 2), (8, 3), (12, 4), (12, 5), (20, 6)])
def test_exact_equals_fractions_random(nz, seed):
    rng = np.random.default_rng(seed)
    dz = 1.9569471624266145e13
    # alpha dz from 1e-15 (thin floor) to 1e3 (thick), albedo up to 1 - 1e-12
    al = 10.0 ** rng.uniform(-15, 3, nz) / dz
    om = np.minimum(10.0 ** rng.uniform(-20, 0, nz), 1 - 1e-12)
    S = 10.0 ** rng.uniform(-60, 10, nz)
    dz1, dz2 = 1.0 / dz, dz * dz
    U, V = exact_oracle(al, om, S, dz, dz1, dz2)
    Uf, Vf = exact_fraction(al, om, S, dz, dz1, dz2)
    assert ulps(U, Uf).max() <= 1.0
    check_V(V, Vf)


def snapshot_column(nz=64, ix=3, iy=4, iw_list=(0, 40, 70, 100, 127)):
    """alpha, albedo, S of one column of the mixed-precision snapshot (thin
    2000 K / 1e-15 floor at both ends, thick middle) from the glibc oracle."""
    from oracle.oracle import Oracle
    from ocotillo_amd.workloads import config4
    rho, T, dz, waves = config4(nx=8, ny=8, nz=nz)
    o = Oracle(waves)
    cells = [o.cell(rho[ix, iy, z], T[ix, iy, z]) for z in range(nz)]
    out = []
    for iw in iw_list:
        pts = [o.point(c, iw) for c in cells]
        out.append((np.array([p.alpha for p in pts]), np.array([p.albedo for p in pts]),
                    np.array([p.S for p in pts])))
    return out, dz


def test_exact_equals_fractions_thin_floor():
    pts, dz = snapshot_column()
    dz1, dz2 = 1.0 / dz, dz * dz
    thin = 0
    for al, om, S in pts:
        thin += int((al * dz < 1e-8).sum())
        U, V = exact_oracle(al, om, S, dz, dz1, dz2)
        Uf, Vf = exact_fraction(al, om, S, dz, dz1, dz2)
        assert ulps(U, Uf).max() <= 1.0
        check_V(V, Vf)
    assert thin > 0   # the case reaches the thin floor


def test_exact_U_equals_difference_form():
    """Two independent formulations of the same discrete problem: binary128
    Thomas on the reference's rows, and FP64 elimination in difference form."""
    from oracle.oracle import Oracle
    from ocotillo_amd.workloads import config4
    rho, T, dz, waves = config4(nx=8, ny=8, nz=128, nw=32)
    r = np.ascontiguousarray(rho[2:4, 5:7])
    t = np.ascontiguousarray(T[2:4, 5:7])
    o = Oracle(waves)
    Ue, Ve, Ae, _ = o.solve_block(r, t, dz, exact=True)
    Ur, Vr, Ar, _ = o.solve_block(r, t, dz, rh=True)
    assert np.array_equal(Ae, Ar)
    assert (np.abs(Ur - Ue) / np.abs(Ue)).max() <= 1e-12


@pytest.mark.parametrize("rho0,dz", [(1e-9, 1e11), (1e-10, 1e11), (1e-10, 1e10)])
def test_reference_equals_exact_when_resolved(rho0, dz):
    """Grey columns with a temperature gradient and alpha dz between 0.16 and
    300 (neither thin enough to lose zeta nor thick enough for U to equal S
    to the last digit): the FP64 reference loses nothing there, so it equals
    the exact solution to 1e-12 (measured 1e-16 .. 4e-15)."""
    from oracle.oracle import Oracle
    nz = 64
    z = np.linspace(-1.0, 1.0, nz)
    rho = (rho0 * np.exp(-z * z)).reshape(1, 1, nz)
    T = (6000.0 + 2000.0 * z).reshape(1, 1, nz)
    o = Oracle(np.array([5000.0]))
    U0, V0, A0, d0 = o.solve_block(rho, T, dz, lgrey=True, sigma_grey=0.4)
    Ue, Ve, Ae, _ = o.solve_block(rho, T, dz, lgrey=True, sigma_grey=0.4, exact=True)
    assert d0.status == 0
    assert 0.1 < (A0 * dz).min() and (A0 * dz).max() < 1e3
    assert (np.abs(U0 - Ue) / np.abs(Ue)).max() <= 1e-12
    assert (np.abs(V0 - Ve) / np.abs(Ve).max()).max() <= 1e-12
    assert (np.abs(V0 - Ve) / np.abs(Ve))[..., [0, -1]].max() <= 1e-12


def test_reference_far_from_exact_on_thin_floor():
    """The reason for this oracle: on the snapshot's thin floor the reference's
    own emergent flux is off from the exact solution by far more than 1e-6
    (bounded by 1e-1 here, 8.6e-3 measured on 4x4-column slices)."""
    from oracle.oracle import Oracle
    from ocotillo_amd.workloads import config4
    rho, T, dz, waves = config4(nx=8, ny=8, nz=256, nw=32)
    r = np.ascontiguousarray(rho[0:2, 0:2])
    t = np.ascontiguousarray(T[0:2, 0:2])
    o = Oracle(waves)
    U0, V0, _, _ = o.solve_block(r, t, dz)
    Ue, Ve, _, _ = o.solve_block(r, t, dz, exact=True)
    e = np.abs(V0[..., [0, -1]] - Ve[..., [0, -1]]) / np.abs(Ve[..., [0, -1]])
    assert e.max() > 1e-6
    assert e.max() < 1e-1
