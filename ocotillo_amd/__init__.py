"""ocotillo_amd -- MI355X implementation of ocotillo's per-column,
per-wavelength hot path (Saha state, hydrogen continuous opacity, Feautrier
solve, 6th-order flux), behind the C ABI in include/ocotillo.h.

This module is the ctypes binding of that ABI (the binding the reference's
Python tools would use).  It never computes on the CPU: if the HIP library
is missing or no GPU is present, every solve raises.

Layouts are the reference's Fortran ones (source/flux_feautrier.f90:36-42);
as C-ordered numpy arrays they are indexed
    rho3d, T3d : [ix, iy, z]        (Fortran (nz, nyloc, nxloc))
    U          : [iw, ix, iy, z+6]  (Fortran (mz, nyloc, nxloc, nw))
    V, absorp  : [iw, ix, iy, z]
"""
import ctypes as C
import os

import numpy as np

__all__ = ["OcoError", "Context", "Batch", "Server", "Diag", "AthenaUnits", "load_library", "LIB_PATH",
           "MODE_FP64", "MODE_MIXED", "MODE_ACCURATE", "EXPORTS"]

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libocotillo.so")

OK, E_PIVOT0_STAGE1, E_PIVOT0_STAGE2, E_ARG, E_OOM, E_HIP, E_STATE, E_NODEVICE, E_BUSY, E_IO, E_NONFINITE = range(11)
MODE_FP64, MODE_MIXED, MODE_ACCURATE = 0, 1, 2
CHECK_FINITE = 1

# every symbol include/ocotillo.h declares
EXPORTS = ["oco_create", "oco_destroy", "oco_set_wavelengths", "oco_set_physics", "oco_set_checks", "oco_solve_block",
           "oco_solve_block_device", "oco_solve_block_device_noalloc", "oco_strerror", "oco_version", "oco_athena_slab", "oco_athena_slab_device",
           "oco_solve_block_files", "oco_write_record", "oco_opacity_points",
           "oco_batch_create", "oco_batch_destroy", "oco_batch_layout", "oco_batch_buffers", "oco_batch_column", "oco_batch_run",
           "oco_batch_run_host", "oco_batch_tuning", "oco_batch_tune_times",
           "oco_server_create", "oco_server_submit", "oco_server_wait", "oco_server_poll", "oco_server_stats",
           "oco_server_destroy"]


class Diag(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("status", "fail_ix", "fail_iy", "fail_iw", "fail_j")]
                + [(n, C.c_double) for n in ("sum_a", "sum_b", "sum_c", "sum_d", "minU", "maxU",
                                             "minV", "maxV", "minA", "maxA", "cells_ms", "sweep_ms",
                                             "total_ms")])


class AthenaUnits(C.Structure):
    _fields_ = ([(n, C.c_double) for n in ("unit_density", "unit_temperature", "gamma1", "rho_floor",
                                             "switch_temp", "floor_temp")]
                + [("avoid_momentum_underflow", C.c_int)])


class Job(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nw", C.c_int), ("dz", C.c_double),
                ("rho3d", C.c_void_p), ("T3d", C.c_void_p), ("waves", C.c_void_p), ("emergent", C.c_void_p)]


class JobResult(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("status", "fail_ix", "fail_iy", "fail_iw", "fail_j")]
                + [(n, C.c_double) for n in ("queue_ms", "latency_ms", "points")])


class ServeStats(C.Structure):
    _fields_ = ([(n, C.c_longlong) for n in ("submitted", "completed", "queued", "active")]
                + [(n, C.c_double) for n in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "points", "wall_s",
                                             "points_per_s")]
                + [("iterations", C.c_longlong), ("hbm_budget", C.c_size_t), ("hbm_peak", C.c_size_t)])


class OcoError(RuntimeError):
    def __init__(self, status, msg, diag=None):
        super().__init__(msg)
        self.status = status
        self.diag = diag


_lib = None
_DP = C.POINTER(C.c_double)


def load_library(path=None):
    """Load libocotillo.so (raises if it has not been built).  OCO_LIBRARY in
    the environment selects another build of the same library."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("OCO_LIBRARY") or LIB_PATH
    if not os.path.exists(p):
        raise OcoError(E_STATE, f"HIP library {p} is not built (run python -m ocotillo_amd.build)")
    L = C.CDLL(p)
    L.oco_create.argtypes = [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]
    L.oco_destroy.argtypes = [C.c_void_p]
    L.oco_set_wavelengths.argtypes = [C.c_void_p, C.c_int, _DP]
    L.oco_set_physics.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double]
    L.oco_set_checks.argtypes = [C.c_void_p, C.c_int]
    L.oco_solve_block.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  _DP, _DP, _DP, _DP, _DP, C.c_int, C.POINTER(Diag)]
    L.oco_solve_block_device.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_int, C.POINTER(Diag)]
    L.oco_solve_block_device_noalloc.argtypes = L.oco_solve_block_device.argtypes
    _SZ = C.POINTER(C.c_size_t)
    _PP = C.POINTER(C.c_void_p)
    L.oco_solve_block_files.argtypes = ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                         _DP, _DP, C.c_int, C.c_char_p, C.c_char_p, C.c_size_t, _DP, _DP,
                                         C.POINTER(Diag)])
    L.oco_write_record.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t]
    L.oco_opacity_points.argtypes = [C.c_void_p, C.c_int, _DP, _DP, C.c_int, _DP, _DP, _DP]
    L.oco_athena_slab.argtypes = ([C.c_void_p] + [C.c_int] * 5 + [_DP] * 6 + [C.POINTER(AthenaUnits)]
                                  + [_DP] * 2)
    L.oco_athena_slab_device.argtypes = ([C.c_void_p] + [C.c_int] * 5 + [C.c_void_p] * 6
                                         + [C.POINTER(AthenaUnits)] + [C.c_void_p] * 2)
    L.oco_batch_create.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _DP, _DP,
                                   C.POINTER(C.c_void_p)]
    L.oco_batch_destroy.argtypes = [C.c_void_p]
    L.oco_batch_layout.argtypes = [C.c_void_p, _SZ, _SZ, _SZ, _SZ]
    L.oco_batch_buffers.argtypes = [C.c_void_p, _PP, _PP, _PP, _PP, _PP]
    L.oco_batch_column.argtypes = [C.c_void_p, C.c_int, _SZ, C.POINTER(C.c_int), _SZ, C.POINTER(C.c_int)]
    L.oco_batch_run.argtypes = [C.c_void_p, C.c_int, C.POINTER(Diag)]
    L.oco_batch_run_host.argtypes = [C.c_void_p, _DP, _DP, _DP, _DP, _DP, C.c_int, C.POINTER(Diag)]
    L.oco_batch_tuning.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int)]
    L.oco_batch_tune_times.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_int]
    L.oco_server_create.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.oco_server_submit.argtypes = [C.c_void_p, C.POINTER(Job), C.POINTER(C.c_longlong)]
    L.oco_server_wait.argtypes = [C.c_void_p, C.c_longlong, C.POINTER(JobResult)]
    L.oco_server_poll.argtypes = [C.c_void_p, C.c_longlong, C.POINTER(JobResult)]
    L.oco_server_stats.argtypes = [C.c_void_p, C.POINTER(ServeStats)]
    L.oco_server_destroy.argtypes = [C.c_void_p]
    L.oco_strerror.restype = C.c_char_p
    L.oco_strerror.argtypes = [C.c_int]
    L.oco_version.restype = C.c_char_p
    if path is None:
        _lib = L
    return L


def strerror(status):
    return load_library().oco_strerror(status).decode()


def write_record(path, array):
    """One array as a one-record Fortran unformatted sequential file in
    gfortran's layout (oco_write_record; no GPU involved)."""
    a = np.ascontiguousarray(array)
    st = load_library().oco_write_record(os.fsencode(path), a.ctypes.data_as(C.c_void_p), a.nbytes)
    if st != OK:
        raise OcoError(st, "oco_write_record: " + strerror(st))


def _ptr(a):
    return a.ctypes.data_as(_DP)


class Context:
    """One device context: resident wavelength tables and work buffers.

    Mirrors the reference's run set-up: set_wavelengths replaces
    pre_calc_opacity_quantities, set_physics the &gas_state_input /
    lgrey / sigma_grey values, solve_block the column loop of one
    processor block."""

    def __init__(self, device=0, hbm_budget=0):
        self.L = load_library()
        h = C.c_void_p()
        st = self.L.oco_create(device, hbm_budget, C.byref(h))
        if st != OK:
            raise OcoError(st, "oco_create: " + strerror(st))
        self.h = h
        self.nw = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.oco_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_wavelengths(self, waves_angstrom):
        w = np.ascontiguousarray(waves_angstrom, dtype=np.float64)
        st = self.L.oco_set_wavelengths(self.h, w.size, _ptr(w))
        if st != OK:
            self.nw = 0   # the library holds no grid after a failed call
            raise OcoError(st, "oco_set_wavelengths: " + strerror(st))
        self.nw = w.size

    def set_physics(self, fully_ionized_T=2e4, lgrey=False, sigma_grey=0.4):
        st = self.L.oco_set_physics(self.h, fully_ionized_T, int(bool(lgrey)), sigma_grey)
        if st != OK:
            raise OcoError(st, "oco_set_physics: " + strerror(st))

    def set_checks(self, finite=False):
        """finite=True: every block solve checks its outputs and returns
        E_NONFINITE (OcoError) at the first NaN or infinity (OCO_CHECK_FINITE)."""
        st = self.L.oco_set_checks(self.h, CHECK_FINITE if finite else 0)
        if st != OK:
            raise OcoError(st, "oco_set_checks: " + strerror(st))

    def solve_block(self, rho3d, T3d, dz, dz1=None, dz2=None, mode=MODE_FP64, out=None, raise_on_pivot=True):
        """Host arrays in, host arrays out.  Returns (U, V, absorp, Diag)."""
        rho3d = np.ascontiguousarray(rho3d, dtype=np.float64)
        T3d = np.ascontiguousarray(T3d, dtype=np.float64)
        if rho3d.ndim != 3 or rho3d.shape != T3d.shape:
            raise OcoError(E_ARG, "rho3d and T3d must both be [nxloc, nyloc, nz]")
        nxloc, nyloc, nz = rho3d.shape
        nw = self.nw
        if out is None:
            U = np.empty((nw, nxloc, nyloc, nz + 6))
            V = np.empty((nw, nxloc, nyloc, nz))
            A = np.empty((nw, nxloc, nyloc, nz))
        else:
            U, V, A = out
        d = Diag()
        dz1 = 1.0 / dz if dz1 is None else dz1
        dz2 = dz * dz if dz2 is None else dz2
        st = self.L.oco_solve_block(self.h, nz, nyloc, nxloc, dz, dz1, dz2, _ptr(rho3d), _ptr(T3d), _ptr(U),
                                    _ptr(V), _ptr(A), mode, C.byref(d))
        if st != OK and (raise_on_pivot or st not in (E_PIVOT0_STAGE1, E_PIVOT0_STAGE2)):
            raise OcoError(st, "oco_solve_block: " + strerror(st), d)
        return U, V, A, d

    def solve_block_device(self, nz, nyloc, nxloc, dz, d_rho, d_T, d_U, d_V, d_A, dz1=None, dz2=None,
                           mode=MODE_FP64, raise_on_pivot=True, noalloc=False):
        """Device pointers (ints, e.g. torch tensor.data_ptr()) in and out.
        noalloc=True: oco_solve_block_device_noalloc (no device allocation;
        difference-form modes, blocks within the static scratch)."""
        d = Diag()
        dz1 = 1.0 / dz if dz1 is None else dz1
        dz2 = dz * dz if dz2 is None else dz2
        name = "oco_solve_block_device_noalloc" if noalloc else "oco_solve_block_device"
        st = getattr(self.L, name)(self.h, nz, nyloc, nxloc, dz, dz1, dz2, C.c_void_p(d_rho), C.c_void_p(d_T),
                                   C.c_void_p(d_U), C.c_void_p(d_V), C.c_void_p(d_A), mode, C.byref(d))
        if st != OK and (raise_on_pivot or st not in (E_PIVOT0_STAGE1, E_PIVOT0_STAGE2)):
            raise OcoError(st, name + ": " + strerror(st), d)
        return d


    def opacity(self, rho, T, mode=MODE_FP64):
        """Absorption coefficient (1/cm), albedo and source function of (rho,
        T) points at the context's wavelengths (oco_opacity_points): arrays
        [npts, nw]."""
        rho = np.ascontiguousarray(np.atleast_1d(rho), dtype=np.float64).reshape(-1)
        T = np.ascontiguousarray(np.atleast_1d(T), dtype=np.float64).reshape(-1)
        if rho.shape != T.shape:
            raise OcoError(E_ARG, "rho and T must have the same number of points")
        out = [np.empty((rho.size, self.nw)) for _ in range(3)]
        st = self.L.oco_opacity_points(self.h, rho.size, _ptr(rho), _ptr(T), mode, *[_ptr(o) for o in out])
        if st != OK:
            raise OcoError(st, "oco_opacity_points: " + strerror(st))
        return tuple(out)

    def solve_block_files(self, rho3d, T3d, dz, outdir, snapshot="0000", dz1=None, dz2=None, mode=MODE_FP64,
                          stage_bytes=0, raise_on_pivot=True):
        """Solve a block and write its three output files into `outdir` (which
        must exist) as the reference's output_binary does.  Returns
        (U_col [nw, nz+6], A_col [nw, nz], Diag): the column (nxloc, nyloc)
        the reference prints to diagnostics.txt."""
        rho3d = np.ascontiguousarray(rho3d, dtype=np.float64)
        T3d = np.ascontiguousarray(T3d, dtype=np.float64)
        if rho3d.ndim != 3 or rho3d.shape != T3d.shape:
            raise OcoError(E_ARG, "rho3d and T3d must both be [nxloc, nyloc, nz]")
        nxloc, nyloc, nz = rho3d.shape
        U_col = np.zeros((self.nw, nz + 6))
        A_col = np.zeros((self.nw, nz))
        d = Diag()
        dz1 = 1.0 / dz if dz1 is None else dz1
        dz2 = dz * dz if dz2 is None else dz2
        st = self.L.oco_solve_block_files(self.h, nz, nyloc, nxloc, dz, dz1, dz2, _ptr(rho3d), _ptr(T3d), mode,
                                          os.fsencode(outdir), snapshot.encode(), stage_bytes, _ptr(U_col),
                                          _ptr(A_col), C.byref(d))
        if st != OK and (raise_on_pivot or st not in (E_PIVOT0_STAGE1, E_PIVOT0_STAGE2)):
            raise OcoError(st, "oco_solve_block_files: " + strerror(st), d)
        return U_col, A_col, d

    def athena_slab(self, slab, tgauss, units, iz0, rho3d, T3d):
        """Raw Athena fields of one z slab -> the slab's z range of rho3d, T3d
        ([nxloc, nyloc, nz], written in place; replaces the arithmetic of
        source/read_athena.f90:163-258 and source/disk.f90:96-104).
        `slab` is a sequence (rho, Mx, My, Mz, E), each a C-ordered
        [nzl, nyloc, nxloc] array (the file's Fortran (nxloc, nyloc, nzl))."""
        f = [np.ascontiguousarray(a, dtype=np.float64) for a in slab]
        if len(f) != 5 or any(a.ndim != 3 or a.shape != f[0].shape for a in f):
            raise OcoError(E_ARG, "slab must be five [nzl, nyloc, nxloc] arrays")
        nzl, nyloc, nxloc = f[0].shape
        for a in (rho3d, T3d):
            if a.dtype != np.float64 or not a.flags.c_contiguous or a.shape[:2] != (nxloc, nyloc):
                raise OcoError(E_ARG, "rho3d and T3d must be C-contiguous float64 [nxloc, nyloc, nz]")
        nz = rho3d.shape[2]
        tg = np.ascontiguousarray(tgauss, dtype=np.float64)
        if tg.size != nz or T3d.shape != rho3d.shape:
            raise OcoError(E_ARG, "tgauss must hold nz values")
        st = self.L.oco_athena_slab(self.h, nz, nyloc, nxloc, nzl, iz0, *[_ptr(a) for a in f], _ptr(tg),
                                    C.byref(units), _ptr(rho3d), _ptr(T3d))
        if st != OK:
            raise OcoError(st, "oco_athena_slab: " + strerror(st))

    def athena_slab_device(self, nz, nyloc, nxloc, nzl, iz0, d_fields, d_tgauss, units, d_rho3d, d_T3d):
        """Device pointers: d_fields = (rho, Mx, My, Mz, E) of the slab."""
        st = self.L.oco_athena_slab_device(self.h, nz, nyloc, nxloc, nzl, iz0,
                                           *[C.c_void_p(p) for p in d_fields], C.c_void_p(d_tgauss),
                                           C.byref(units), C.c_void_p(d_rho3d), C.c_void_p(d_T3d))
        if st != OK:
            raise OcoError(st, "oco_athena_slab_device: " + strerror(st))


class Batch:
    """A plan for many independent columns of uneven shape (oco_batch_*).

    nz, nw: per-column cell and wavelength counts; waves: all columns'
    wavelength grids concatenated (Angstrom); dz: per-column grid step.
    Inputs rho, T are the columns concatenated (sum(nz) values); outputs are
    per column U (nw_c, nz_c+6), V and absorp (nw_c, nz_c) -- the Fortran
    (z, iw) layout -- concatenated.  The first run autotunes and captures a
    hipGraph; later runs replay it."""

    def __init__(self, ctx, nz, nw, waves, dz):
        self.ctx = ctx
        self.L = ctx.L
        self.nz = np.ascontiguousarray(nz, dtype=np.int32)
        self.nw = np.ascontiguousarray(nw, dtype=np.int32)
        w = np.ascontiguousarray(waves, dtype=np.float64)
        d = np.ascontiguousarray(dz, dtype=np.float64)
        if w.size != int(self.nw.sum()) or d.size != self.nz.size or self.nw.size != self.nz.size:
            raise OcoError(E_ARG, "Batch: nz, nw, dz must have one entry per column and waves sum(nw)")
        h = C.c_void_p()
        st = self.L.oco_batch_create(ctx.h, self.nz.size, self.nz.ctypes.data_as(C.POINTER(C.c_int)),
                                     self.nw.ctypes.data_as(C.POINTER(C.c_int)), _ptr(w), _ptr(d), C.byref(h))
        if st != OK:
            raise OcoError(st, "oco_batch_create: " + strerror(st))
        self.h = h
        sizes = [C.c_size_t() for _ in range(4)]
        self.L.oco_batch_layout(h, *[C.byref(x) for x in sizes])
        self.ncells, self.nsys, self.nU, self.nVA = [x.value for x in sizes]

    def close(self):
        if getattr(self, "h", None):
            self.L.oco_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def buffers(self):
        """Device pointers (rho, T, U, V, absorp) of the plan's resident arrays
        (outputs with padded rows: `column`)."""
        p = [C.c_void_p() for _ in range(5)]
        self.L.oco_batch_buffers(self.h, *[C.byref(x) for x in p])
        return [x.value for x in p]

    def column(self, col):
        """(u_off, ldu, va_off, ldva): where column `col` sits in the device
        outputs, offsets in doubles and leading dimensions (rows padded to 16
        doubles)."""
        uo, vo, lu, lv = C.c_size_t(), C.c_size_t(), C.c_int(), C.c_int()
        st = self.L.oco_batch_column(self.h, int(col), C.byref(uo), C.byref(lu), C.byref(vo), C.byref(lv))
        if st != OK:
            raise OcoError(st, "oco_batch_column: " + strerror(st))
        return uo.value, lu.value, vo.value, lv.value

    def tuning(self, mode=MODE_FP64):
        """(kernel variant, workgroups, tuned?) of the autotuner for `mode`."""
        v, nb, t = C.c_int(), C.c_int(), C.c_int()
        self.L.oco_batch_tuning(self.h, mode, C.byref(v), C.byref(nb), C.byref(t))
        return v.value, nb.value, t.value

    def tune_times(self, mode=MODE_FP64):
        """The autotuner's median timings (ms) for `mode`, in the order it
        tried its candidates (oco_batch_tune_times); empty before the first
        run in the mode."""
        buf = (C.c_float * 32)()
        n = self.L.oco_batch_tune_times(self.h, mode, buf, 32)
        if n < 0:
            raise OcoError(-n, "oco_batch_tune_times: " + strerror(-n))
        return [round(buf[i], 4) for i in range(min(n, 32))]

    def run(self, mode=MODE_FP64, raise_on_pivot=True):
        d = Diag()
        st = self.L.oco_batch_run(self.h, mode, C.byref(d))
        if st != OK and (raise_on_pivot or st not in (E_PIVOT0_STAGE1, E_PIVOT0_STAGE2)):
            raise OcoError(st, "oco_batch_run: " + strerror(st), d)
        return d

    def run_host(self, rho, T, mode=MODE_FP64, raise_on_pivot=True):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        if rho.size != self.ncells or T.size != self.ncells:
            raise OcoError(E_ARG, "Batch.run_host: rho and T need sum(nz) values")
        U = np.empty(self.nU)
        V = np.empty(self.nVA)
        A = np.empty(self.nVA)
        d = Diag()
        st = self.L.oco_batch_run_host(self.h, _ptr(rho), _ptr(T), _ptr(U), _ptr(V), _ptr(A), mode, C.byref(d))
        if st != OK and (raise_on_pivot or st not in (E_PIVOT0_STAGE1, E_PIVOT0_STAGE2)):
            raise OcoError(st, "oco_batch_run_host: " + strerror(st), d)
        return U, V, A, d

    def split(self, U, V, A):
        """Per-column views: U[c] is (nw_c, nz_c+6), V[c], A[c] are (nw_c, nz_c)."""
        outU, outV, outA = [], [], []
        u = v = 0
        for nz, nw in zip(self.nz.tolist(), self.nw.tolist()):
            outU.append(U[u:u + (nz + 6) * nw].reshape(nw, nz + 6))
            outV.append(V[v:v + nz * nw].reshape(nw, nz))
            outA.append(A[v:v + nz * nw].reshape(nw, nz))
            u += (nz + 6) * nw
            v += nz * nw
        return outU, outV, outA


class Server:
    """The serving loop (oco_server_*): submit jobs as they arrive, each a
    block of columns with its own shape and wavelength grid; a job returns the
    emergent flux V(1), V(nz) of every system as an array [iw, ix, iy, 2]
    (Fortran (2, ny, nx, nw)).  Inputs and output arrays are held here until
    the job is collected."""

    def __init__(self, ctx, hbm_budget, max_nz=2048, mode=MODE_MIXED):
        self.ctx = ctx
        self.L = ctx.L
        h = C.c_void_p()
        st = self.L.oco_server_create(ctx.h, int(hbm_budget), int(max_nz), int(mode), C.byref(h))
        if st != OK:
            raise OcoError(st, "oco_server_create: " + strerror(st))
        self.h = h
        self.jobs = {}

    def submit(self, rho3d, T3d, dz, waves, out=None):
        """rho3d, T3d: [ix, iy, z]; returns a ticket.  `out`, if given, is the
        emergent-flux array [iw, ix, iy, 2] (C-ordered float64) the job writes,
        e.g. in pinned host memory, where the download runs at full rate; by
        default a new (pageable) array."""
        rho3d = np.ascontiguousarray(rho3d, dtype=np.float64)
        T3d = np.ascontiguousarray(T3d, dtype=np.float64)
        waves = np.ascontiguousarray(waves, dtype=np.float64)
        if rho3d.ndim != 3 or rho3d.shape != T3d.shape:
            raise OcoError(E_ARG, "rho3d and T3d must both be [nxloc, nyloc, nz]")
        nx, ny, nz = rho3d.shape
        if out is None:
            out = np.empty((waves.size, nx, ny, 2))
        elif (out.dtype != np.float64 or out.shape != (waves.size, nx, ny, 2) or not out.flags.c_contiguous
              or not out.flags.writeable):
            raise OcoError(E_ARG, "out must be a writeable C-ordered float64 array of shape (nw, nx, ny, 2)")
        job = Job(nx, ny, nz, waves.size, float(dz), rho3d.ctypes.data, T3d.ctypes.data, waves.ctypes.data,
                  out.ctypes.data)
        t = C.c_longlong()
        st = self.L.oco_server_submit(self.h, C.byref(job), C.byref(t))
        if st != OK:
            raise OcoError(st, "oco_server_submit: " + strerror(st))
        self.jobs[t.value] = (rho3d, T3d, waves, out)
        return t.value

    def wait(self, ticket, raise_on_pivot=True):
        """Block until the job completes; returns (emergent [iw, ix, iy, 2], JobResult)."""
        r = JobResult()
        st = self.L.oco_server_wait(self.h, ticket, C.byref(r))
        out = self.jobs.pop(ticket)[3]
        if st != OK and (raise_on_pivot or st not in (E_PIVOT0_STAGE1, E_PIVOT0_STAGE2)):
            raise OcoError(st, "oco_server_wait: " + strerror(st))
        return out, r

    def poll(self, ticket):
        """(emergent, JobResult) if the job has completed, else None."""
        r = JobResult()
        st = self.L.oco_server_poll(self.h, ticket, C.byref(r))
        if st == E_BUSY:
            return None
        out = self.jobs.pop(ticket)[3]
        if st not in (OK, E_PIVOT0_STAGE1, E_PIVOT0_STAGE2):
            raise OcoError(st, "oco_server_poll: " + strerror(st))
        return out, r

    def stats(self):
        s = ServeStats()
        self.L.oco_server_stats(self.h, C.byref(s))
        return s

    def close(self):
        if getattr(self, "h", None):
            self.L.oco_server_destroy(self.h)
            self.h = None
            self.jobs.clear()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
