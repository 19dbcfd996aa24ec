"""Snapshot job source: the snapshots of a run directory, block by block,
through the GPU path -- to the reference's output files, or as jobs of the
serving loop.

The reference runs one snapshot per process, taken from argv
(source/flux_feautrier.f90:69-70). samples/multiInTest/snapshot_looper.sh
meant to run snapshots 0000-0015, each into output_snapshot<tag>/, but it
edits a namelist line no code reads, so every pass reruns 0000. Here one
process takes any list of snapshots:

    python -m ocotillo_amd.snapshots RUN_DIR [0:15 | 0000 0003 ...] [--mixed] [--emergent]

With files (the default) each snapshot's directory holds what a separate
run of the Fortran host (ocotillo_amd/fortran/flux_feautrier_gpu.x <tag>)
writes there: per block mean_intensity_, absorption_coefficients_ and
flux_<tag>.bin (oco_solve_block_files), diagnostics.txt, zgrid.bin and
wavegrid.bin. The list-directed zgrid.txt and wavegrid.txt, whose layout
the Fortran standard leaves to the compiler, are left to the Fortran host.
With --emergent the blocks are submitted to one server (oco_server_*) as
they are built, and each block's emergent flux V(1), V(nz) per (column,
wavelength) is saved to emergent_<tag>_procx<i>_procy<j>.npy instead.

The run's namelists are the reference's: &input (w0, w1, lgrey, sigma_grey,
lread_athena), &gas_state_input, &athena_input or, for the analytic disk,
&temperature_input and &density_input; &oco_input (lwave_log, oco_mode)
as the Fortran host reads it.
"""
import argparse
import math
import os
import re
import sys

import numpy as np

from .athena import NG, AthenaRun, read_namelist, read_resolution
from .fortio import diagnostics_text


def parse_snapshots(args):
    """Snapshot tags from command-line words: a tag, or a range a:b of
    snapshot numbers (four digits, zero-padded).  None or empty: 0000."""
    tags = []
    for a in args or []:
        m = re.fullmatch(r"(\d+):(\d+)", a)
        if m:
            lo, hi = int(m.group(1)), int(m.group(2))
            if hi < lo:
                raise ValueError("bad snapshot range " + a)
            tags += ["%04d" % k if k < 10000 else str(k) for k in range(lo, hi + 1)]
        else:
            tags.append(a)
    return tags or ["0000"]


class Run:
    """The run-wide settings of a run directory (input.in, resolution.in)."""

    def __init__(self, run_dir):
        self.dir = run_dir
        inp = os.path.join(run_dir, "input.in")
        self.shapes = read_resolution(os.path.join(run_dir, "resolution.in"))
        i = read_namelist(inp, "input")
        self.w0, self.w1 = float(i.get("w0", 3000.0)), float(i.get("w1", 5000.0))
        self.z0, self.z1 = float(i.get("z0", 0.0)), float(i.get("z1", 0.0))
        self.lgrey = bool(i.get("lgrey", False))
        self.sigma_grey = float(i.get("sigma_grey", 0.0))
        self.lread_athena = bool(i.get("lread_athena", True))
        self.fully_ionized_T = float(read_namelist(inp, "gas_state_input")["fully_ionized_t"])
        try:
            o = read_namelist(inp, "oco_input")
        except KeyError:
            o = {}
        self.lwave_log = bool(o.get("lwave_log", False))
        self.mode = int(o.get("oco_mode", 0))
        self.athena = AthenaRun(run_dir) if self.lread_athena else None
        if not self.lread_athena:
            t = read_namelist(inp, "temperature_input")
            d = read_namelist(inp, "density_input")
            self.mid_temp, self.sigma, self.floor_temp = float(t["midtemp"]), float(t["sigma"]), float(t["floortemp"])
            self.switch_temp = (1 + 2.0 ** -52) * float(t["isotemp"])
            self.rho0, self.rho_floor, self.H = float(d["rho0"]), float(d["rho_floor"]), float(d["h"])

    def waves(self):
        """calc_wavelength (source/continuous_opacity.f90:37-62), or the
        host's log-spaced grid under lwave_log."""
        nw, w0, w1 = self.shapes["nw"], self.w0, self.w1
        if nw == 1:
            return np.array([w0])
        if self.lwave_log:
            return np.array([math.exp(math.log(w0) + (i * (math.log(w1) - math.log(w0))) / (nw - 1))
                             for i in range(nw)])
        dw = (w1 - w0) / (nw - 1)
        return np.array([w0 + i * dw for i in range(nw)])

    def disk_block(self):
        """The analytic disk (source/disk.f90:42-129): z grid, one pillar
        copied to every column of the block."""
        s = self.shapes
        nz = s["nz"]
        dz = (self.z1 - self.z0) / (nz - 1)
        z = np.zeros(nz + 2 * NG)
        for i in range(nz):
            z[NG + i] = self.z0 + i * dz
        for i in range(1, NG + 1):
            z[NG - i] = z[NG] - i * dz
            z[NG + nz - 1 + i] = z[NG + nz - 1] + i * dz
        zz = z[NG:NG + nz]
        rho = np.array([self.rho0 * math.exp(-((0.5 * (v * v)) / (self.H * self.H))) for v in zz])
        rho = np.where(rho < self.rho_floor, self.rho_floor, rho)
        s2 = self.sigma * self.sigma
        tg = np.array([self.mid_temp * math.exp(-((0.5 * (v * v)) / s2)) for v in zz])
        T = np.full(nz, self.mid_temp)
        below = T < self.switch_temp
        T = np.where(below, np.where(tg > self.floor_temp, tg, self.floor_temp), T)
        shape = (s["nxloc"], s["nyloc"], nz)
        return np.broadcast_to(rho, shape).copy(), np.broadcast_to(T, shape).copy(), z, dz

    def block(self, ctx, iprocx, iprocy, tag):
        if self.athena is not None:
            return self.athena.load_block(ctx, iprocx, iprocy, tag)
        return self.disk_block()

    def blocks(self):
        """(iprocx, iprocy) in the reference's order (source/flux_feautrier.f90:132-133)."""
        return [(px, py) for px in range(self.shapes["nprocx"]) for py in range(self.shapes["nprocy"])]


def out_root(run, tags, tag):
    return os.path.join(run.dir, "output" if len(tags) == 1 else "output_snapshot" + tag)


def write_snapshots(ctx, run, tags, mode=None, stage_bytes=0, log=print):
    """Each snapshot's outputs into its directory, as separate runs of the
    Fortran host would write them (binary files and diagnostics.txt).
    Returns {tag: output directory}."""
    from . import write_record
    mode = run.mode if mode is None else mode
    waves = run.waves()
    ctx.set_wavelengths(waves)
    ctx.set_physics(run.fully_ionized_T, run.lgrey, run.sigma_grey)
    done = {}
    for tag in tags:
        if run.athena is not None:
            run.athena._tgauss = None            # each snapshot as its own run
        root = out_root(run, tags, tag)
        dz1 = dz2 = None
        for px, py in run.blocks():
            rho, T, z, dz = run.block(ctx, px, py, tag)
            if dz1 is None:                          # root block (source/flux_feautrier.f90:151-154)
                dz1, dz2 = 1.0 / dz, dz * dz
            d = os.path.join(root, "procx%d_procy%d" % (px, py))
            os.makedirs(d, exist_ok=True)
            Ucol, Acol, diag = ctx.solve_block_files(rho, T, dz, d, tag, dz1=dz1, dz2=dz2, mode=mode,
                                                     stage_bytes=stage_bytes)
            if (px, py) == (0, 0):
                write_record(os.path.join(root, "zgrid.bin"), z)
                write_record(os.path.join(root, "wavegrid.bin"), waves)
                with open(os.path.join(root, "diagnostics.txt"), "w") as f:
                    f.write(diagnostics_text(Ucol, Acol))
        log("snapshot %s done: %s" % (tag, root))
        done[tag] = root
    return done


def serve_snapshots(ctx, run, tags, mode=None, hbm_budget=64 << 30, log=print):
    """Every block of every snapshot as a job of one server, submitted as it
    is built (device input construction overlaps the server's launches);
    returns {(tag, iprocx, iprocy): emergent [nw, nxloc, nyloc, 2]}."""
    from . import Server
    mode = run.mode if mode is None else mode
    waves = run.waves()
    ctx.set_physics(run.fully_ionized_T, run.lgrey, run.sigma_grey)
    tickets = {}
    with Server(ctx, hbm_budget, max_nz=run.shapes["nz"], mode=mode) as srv:
        for tag in tags:
            if run.athena is not None:
                run.athena._tgauss = None
            for px, py in run.blocks():
                rho, T, z, dz = run.block(ctx, px, py, tag)
                tickets[(tag, px, py)] = srv.submit(rho, T, dz, waves)
        out = {k: srv.wait(t)[0] for k, t in tickets.items()}
        st = srv.stats()
    log("%d jobs, p50 %.1f ms, p99 %.1f ms, %.3g points/s" % (st.completed, st.p50_ms, st.p99_ms, st.points_per_s))
    return out


def main(argv=None):
    import ocotillo_amd as oco
    p = argparse.ArgumentParser(prog="python -m ocotillo_amd.snapshots", description=__doc__.split("\n\n")[0])
    p.add_argument("run_dir")
    p.add_argument("snapshots", nargs="*", help="tags or ranges a:b (default: every snapshot present)")
    p.add_argument("--mixed", action="store_true", help="mixed precision (default: &oco_input oco_mode)")
    p.add_argument("--emergent", action="store_true", help="serve the blocks, save only the emergent flux")
    p.add_argument("--device", type=int, default=0)
    args = p.parse_args(argv)
    run = Run(args.run_dir)
    tags = parse_snapshots(args.snapshots) if args.snapshots else (
        run.athena.snapshots() if run.athena is not None else ["0000"])
    mode = oco.MODE_MIXED if args.mixed else None
    try:
        with oco.Context(args.device) as ctx:
            if args.emergent:
                out = serve_snapshots(ctx, run, tags, mode)
                for (tag, px, py), em in out.items():
                    np.save(os.path.join(run.dir, "emergent_%s_procx%d_procy%d.npy" % (tag, px, py)), em)
            else:
                write_snapshots(ctx, run, tags, mode)
    except oco.OcoError as e:
        print("ocotillo GPU library: %s" % e, file=sys.stderr)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main())
