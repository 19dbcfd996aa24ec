"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/lib/liboco_oracle.so (glibc libm, as the reference) or
liboco_oracle_pm.so (the product's portable elementary functions, which the
GPU kernels use).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module; nothing in the product does.

Layouts follow the reference (source/flux_feautrier.f90:36-42): Fortran
order, i.e. numpy arrays indexed [iw, ix, iy, z] in C order.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
_libs = {}

DP = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Tables(C.Structure):
    _fields_ = ([("nw", C.c_int)]
                + [(n, C.POINTER(C.c_double)) for n in
                   ("wa", "wcm1", "nu", "chi", "chi1", "wa3", "w1cm5", "gff", "s_coeff", "f1", "f2", "f3", "g")]
                + [("n1", C.c_double * 6), ("chi_n", C.c_double * 6), ("chi_m", C.c_double),
                   ("damping_factor_const", C.c_double), ("source_function_const", C.c_double)])


class Cell(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("rho", "rho1", "T", "T1", "x", "n", "ninv", "nHI", "nHII", "ne", "fn", "Pe", "es",
                 "theta", "theta1", "lgtheta", "lgtheta2", "hmbf")]


class Point(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("alpha", "albedo", "alpha1", "alpha2", "S", "stim")]


class Diag(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("status", "fail_ix", "fail_iy", "fail_iw", "fail_j")]
                + [(n, C.c_double) for n in ("sum_a", "sum_b", "sum_c", "sum_d",
                                             "minU", "maxU", "minV", "maxV", "minA", "maxA")])


def build():
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)


def lib(portable=False):
    name = {False: "liboco_oracle.so", True: "liboco_oracle_pm.so", "env": "liboco_oracle_env.so"}[portable]
    if name not in _libs:
        path = os.path.join(LIBDIR, name)
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.oo_tables_init.argtypes = [C.POINTER(Tables), C.c_int, DP]
        L.oo_tables_free.argtypes = [C.POINTER(Tables)]
        L.oo_cell_state.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(Cell)]
        L.oo_opacity_point.argtypes = [C.POINTER(Tables), C.POINTER(Cell), C.c_int, C.POINTER(Point)]
        L.oo_grey_point.argtypes = [C.POINTER(Cell), C.c_double, C.POINTER(Point)]
        L.oo_solve_block.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(Tables),
                                     C.c_double, C.c_double, C.c_double, DP, DP,
                                     C.c_double, C.c_int, C.c_double, DP, DP, DP, C.POINTER(Diag), C.c_int]
        L.oo_solve_block_ex.argtypes = L.oo_solve_block.argtypes + [C.c_int]
        L.oo_rh_coefficients.argtypes = [C.c_int, DP, DP, DP, C.c_double, C.c_double, DP, DP, DP, DP]
        L.oo_tridag_rh.argtypes = [C.c_int, DP, DP, DP, DP, DP, DP]
        L.oo_exact_system.argtypes = [C.c_int, DP, DP, DP, C.c_double, C.c_double, C.c_double, DP, DP]
        L.oo_tridag.argtypes = [C.c_int, DP, DP, DP, DP, DP, DP, C.POINTER(C.c_int)]
        L.oo_tridag_coefficients.argtypes = [C.c_int, DP, DP, DP, DP, DP, C.c_double, C.c_double,
                                             DP, DP, DP, DP]
        L.oo_ghosts_flux.argtypes = [C.c_int, DP, DP, C.c_double, DP]
        L.oo_calc_wavelength.argtypes = [C.c_double, C.c_double, C.c_int, DP]
        L.oo_calc_grid.argtypes = [C.c_int, C.c_double, C.c_double, DP, C.POINTER(C.c_double)]
        L.oo_calc_density.argtypes = [C.c_int, DP, C.c_double, C.c_double, C.c_double, DP]
        L.oo_calc_temperature.argtypes = [C.c_int, DP, C.c_double, C.c_double, C.c_double, C.c_double,
                                          C.c_void_p, C.c_int, DP]
        L.oo_athena_units.argtypes = [C.c_double] * 5 + [C.POINTER(C.c_double)] * 3
        L.oo_athena_slab_raw.argtypes = [C.c_int] * 5 + [DP] * 5 + [C.c_int, DP, DP]
        L.oo_athena_postprocess.argtypes = ([C.c_int] * 3 + [C.c_double] * 4 + [DP, C.c_double, C.c_double]
                                            + [DP, DP])
        if portable == "env":
            L.oo_set_perturbation.argtypes = [C.c_ulonglong]
        _libs[name] = L
    return _libs[name]


class Oracle:
    """One oracle instance bound to a wavelength grid (pre_calc_opacity_quantities)."""

    def __init__(self, waves, portable=False):
        self.L = lib(portable)
        self.waves = np.ascontiguousarray(waves, dtype=np.float64)
        self.t = Tables()
        if self.L.oo_tables_init(C.byref(self.t), self.waves.size, self.waves) != 0:
            raise MemoryError
        self.nw = self.waves.size

    def __del__(self):
        try:
            self.L.oo_tables_free(C.byref(self.t))
        except Exception:
            pass

    def tables(self):
        nw = self.nw
        out = {}
        for n in ("wa", "wcm1", "nu", "chi", "chi1", "wa3", "w1cm5", "gff", "s_coeff", "f1", "f2", "f3"):
            out[n] = np.ctypeslib.as_array(getattr(self.t, n), (nw,)).copy()
        out["g"] = np.ctypeslib.as_array(self.t.g, (nw * 6,)).reshape(nw, 6).copy()
        out["n1"] = np.array(self.t.n1[:])
        out["chi_n"] = np.array(self.t.chi_n[:])
        out["chi_m"] = self.t.chi_m
        out["damping_factor_const"] = self.t.damping_factor_const
        out["source_function_const"] = self.t.source_function_const
        return out

    def cell(self, rho, T, fully_ionized_T=2e4):
        c = Cell()
        self.L.oo_cell_state(rho, T, fully_ionized_T, C.byref(c))
        return c

    def point(self, cell, iw, lgrey=False, sigma_grey=0.4):
        p = Point()
        if lgrey:
            self.L.oo_grey_point(C.byref(cell), sigma_grey, C.byref(p))
        else:
            self.L.oo_opacity_point(C.byref(self.t), C.byref(cell), iw, C.byref(p))
        return p

    def solve_block(self, rho3d, T3d, dz, fully_ionized_T=2e4, lgrey=False, sigma_grey=0.4,
                    dz1=None, dz2=None, nthreads=0, rh=False, exact=False):
        """rho3d, T3d: arrays [ix, iy, z] (C order == Fortran (z,iy,ix)).
        Returns U [nw,ix,iy,mz], V and alpha [nw,ix,iy,nz], Diag.
        rh=True solves the rows in difference form (the accurate solution of
        the same discrete problem; not the reference's elimination).
        exact=True evaluates the reference's rows, Thomas, ghosts, der and
        flux in binary128 on this oracle's FP64 opacity (oo_exact_system): the
        exact solution of the reference's discrete problem, rounded to FP64."""
        rho3d = np.ascontiguousarray(rho3d, dtype=np.float64)
        T3d = np.ascontiguousarray(T3d, dtype=np.float64)
        nxloc, nyloc, nz = rho3d.shape
        nw = self.nw
        U = np.zeros((nw, nxloc, nyloc, nz + 6))
        V = np.zeros((nw, nxloc, nyloc, nz))
        A = np.zeros((nw, nxloc, nyloc, nz))
        d = Diag()
        dz1 = 1.0 / dz if dz1 is None else dz1
        dz2 = dz * dz if dz2 is None else dz2
        self.L.oo_solve_block_ex(nz, nyloc, nxloc, C.byref(self.t), dz, dz1, dz2,
                                 rho3d.reshape(-1), T3d.reshape(-1), fully_ionized_T, int(bool(lgrey)),
                                 sigma_grey, U.reshape(-1), V.reshape(-1), A.reshape(-1), C.byref(d), nthreads,
                                 2 if exact else int(bool(rh)))
        return U, V, A, d


def calc_wavelength(w1, w0, nw, portable=False):
    wa = np.zeros(nw)
    lib(portable).oo_calc_wavelength(w1, w0, nw, wa)
    return wa


def envelope(rho3d, T3d, dz, waves, seeds=(1, 2, 3), **kw):
    """Conditioning envelope of the reference's answer: the largest change of
    U, V, alpha over runs whose libm results are moved by +-1 ulp.
    Returns (U0, V0, A0, EU, EV, EA) with *0 the unperturbed glibc result."""
    o = Oracle(waves, portable="env")
    o.L.oo_set_perturbation(0)
    U0, V0, A0, d0 = o.solve_block(rho3d, T3d, dz, **kw)
    EU = np.zeros_like(U0)
    EV = np.zeros_like(V0)
    EA = np.zeros_like(A0)
    for s in seeds:
        o.L.oo_set_perturbation(s)
        U, V, A, d = o.solve_block(rho3d, T3d, dz, **kw)
        np.maximum(EU, np.abs(U - U0), out=EU)
        np.maximum(EV, np.abs(V - V0), out=EV)
        np.maximum(EA, np.abs(A - A0), out=EA)
    o.L.oo_set_perturbation(0)
    return U0, V0, A0, EU, EV, EA
