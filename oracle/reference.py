"""The compiled reference as a CPU baseline -- TEST INFRASTRUCTURE ONLY.

Runs oracle/_ref/time_nz<nz>_nw<nw>/ref_time.x (the reference's own modules,
compiled unchanged at -O3 by oracle/ref/build_ref.sh, driven by the column
loop of source/flux_feautrier.f90:160-228 in oracle/ref/ref_time.f90) on a
block's columns: P processes at once on disjoint column slices, the way the
reference's processor blocks would run side by side ("MPI over all host
cores": the reference has no MPI code and its columns never interact).  Only
bench.py's cpu_baseline leg and tests use this module.
"""
import os
import subprocess
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REFDIR = os.path.join(HERE, "_ref")


def binary(nz, nw):
    """Path of the timing build for these compile-time shapes, or None."""
    p = os.path.join(REFDIR, "time_nz%d_nw%d" % (nz, nw), "ref_time.x")
    return p if os.path.exists(p) else None


def build(nz, nw):
    """Compile the timing build (needs the reference sources)."""
    subprocess.check_call(["sh", os.path.join(HERE, "ref", "build_ref.sh"), "time", str(nz), str(nw)],
                          stdout=subprocess.DEVNULL)
    return binary(nz, nw)


def write_input(path, rho_cols, T_cols, dz, waves, lgrey=False, sigma_grey=0.4):
    """rho_cols, T_cols: [ncol, nz] arrays (one column per row)."""
    rho_cols = np.ascontiguousarray(rho_cols, dtype="<f8")
    T_cols = np.ascontiguousarray(T_cols, dtype="<f8")
    with open(path, "wb") as f:
        f.write(np.array([rho_cols.shape[0], int(bool(lgrey))], "<i4").tobytes())
        f.write(np.array([dz, sigma_grey], "<f8").tobytes())
        f.write(np.asarray(waves, "<f8").tobytes())
        f.write(rho_cols.tobytes())
        f.write(T_cols.tobytes())


def run(rho3d, T3d, dz, waves, nprocs, fully_ionized_T=2e4, dump_cols=0):
    """The block's columns (rho3d, T3d [nx, ny, nz]) split over nprocs
    processes of the reference started together.  Returns (wall seconds from
    the first start to the last exit, list of per-process result lines:
    ncol, nz, nw, loop wall s, loop cpu s, sum U, sum V, sum absorp).  With
    dump_cols > 0, the first process also writes its first dump_cols
    columns' outputs (after its timed loop), and a third item is returned:
    (U [nw, dump_cols, nz+6], V [nw, dump_cols, nz], absorp [nw, dump_cols,
    nz]), the block's columns in (ix, iy) order, iy fastest."""
    nx, ny, nz = rho3d.shape
    exe = binary(nz, len(waves))
    if exe is None:
        raise FileNotFoundError("no reference timing build for nz=%d nw=%d" % (nz, len(waves)))
    cols_rho = np.ascontiguousarray(rho3d).reshape(-1, nz)
    cols_T = np.ascontiguousarray(T3d).reshape(-1, nz)
    slices = [s for s in np.array_split(np.arange(cols_rho.shape[0]), nprocs) if s.size]
    with tempfile.TemporaryDirectory() as td:
        nml = os.path.join(td, "gas.nml")
        with open(nml, "w") as f:
            f.write("&gas_state_input\n  fully_ionized_T = %r\n/\n" % float(fully_ionized_T))
        files = []
        for i, s in enumerate(slices):
            fi = os.path.join(td, "in%d.bin" % i)
            write_input(fi, cols_rho[s], cols_T[s], dz, waves)
            files.append(fi)
        nd = min(int(dump_cols), len(slices[0]))
        dump = os.path.join(td, "dump.bin")
        cmds = [[exe, fi, nml] + ([dump, str(nd)] if i == 0 and nd > 0 else []) for i, fi in enumerate(files)]
        t0 = time.perf_counter()
        procs = [subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for c in cmds]
        outs = [p.communicate() for p in procs]
        wall = time.perf_counter() - t0
        for p, (o, e) in zip(procs, outs):
            if p.returncode != 0:
                raise RuntimeError("reference run failed: %s %s" % (o[-500:], e[-500:]))
        lines = [[float(x) for x in o.split()] for o, _ in outs]
        if nd <= 0:
            return wall, lines
        nw = len(waves)
        raw = np.fromfile(dump, dtype="<f8")
        nu, nv = nw * nd * (nz + 6), nw * nd * nz
        if raw.size != nu + 2 * nv:
            raise RuntimeError("reference dump: %d doubles, expected %d" % (raw.size, nu + 2 * nv))
        U = raw[:nu].reshape(nw, nd, nz + 6)
        V = raw[nu:nu + nv].reshape(nw, nd, nz)
        A = raw[nu + nv:].reshape(nw, nd, nz)
    return wall, lines, (U, V, A)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """Threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set (the share of a shared host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)
