"""Writes the golden fixtures under tests/golden/ from the real ocotillo reference.

Runs only where /root/reference and amdflang exist (the build container).  It
  1. copies the reference's own sample fixtures (input.in, resolution.in,
     reference.out, Athena .bin inputs) into tests/golden/samples/,
  2. builds and runs the reference executable on each sample
     (oracle/ref/build_ref.sh sample ...) and keeps its full-precision binary
     outputs (U, alpha, V; source/io.f90:76-101) for a subset of columns plus
     whole-array sums,
  3. builds the golden-vector driver (oracle/ref/ref_dump.f90) for several
     (nz, nw) shapes and runs it on seeded synthetic columns, keeping every
     per-point intermediate (alpha, albedo, S, a, b, c, d, U, V).
Inputs are stored inside every .npz, so the fixtures do not depend on this
script's generators.
"""
import os
import shutil
import subprocess
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("OCO_REFERENCE", "/root/reference")
BUILD = os.path.join(REPO, "oracle", "ref", "build_ref.sh")
REFOUT = os.path.join(REPO, "oracle", "_ref")
SEED0 = 20261203
SAMPLES = ["GreyIsothermal", "NongreyAGN", "ReadFromAthena", "multiInTest"]
AU = 1.496e14


def read_records(path):
    """Fortran unformatted sequential file -> list of raw record bytes."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        n = int(np.frombuffer(data, "<i4", 1, pos)[0])
        out.append(data[pos + 4: pos + 4 + n])
        pos += 8 + n
    return out


def parse_resolution(path):
    env = {}
    for line in open(path):
        line = line.split("!")[0]
        if "::" not in line:
            continue
        for item in line.split("::")[1].split(","):
            k, v = item.split("=")
            env[k.strip()] = int(eval(v.strip(), {}, dict(env)) // 1)
    return env


def copy_samples():
    for s in SAMPLES:
        src = os.path.join(REF, "samples", s)
        dst = os.path.join(HERE, "samples", s)
        if os.path.exists(dst):
            shutil.rmtree(dst)
        os.makedirs(dst)
        for f in ("input.in", "resolution.in", "reference.out"):
            shutil.copy(os.path.join(src, f), dst)
        if os.path.isdir(os.path.join(src, "input_athena")):
            shutil.copytree(os.path.join(src, "input_athena"), os.path.join(dst, "input_athena"))
            for root, dirs, files in os.walk(os.path.join(dst, "input_athena")):
                for d in dirs:
                    os.chmod(os.path.join(root, d), 0o755)
                for f in files:
                    os.chmod(os.path.join(root, f), 0o644)


def run_samples():
    rng = np.random.default_rng(SEED0)
    for s in SAMPLES:
        case = os.path.join(HERE, "samples", s)
        subprocess.check_call([BUILD, "sample", s, case])
        wd = os.path.join(REFOUT, s)
        subprocess.check_call(["./source/flux_feautrier.x"], cwd=wd, stdout=subprocess.DEVNULL)
        res = parse_resolution(os.path.join(case, "resolution.in"))
        nz, nw, nyloc, nxloc = res["nz"], res["nw"], res["nyloc"], res["nxloc"]
        mz = nz + 6
        out = {"nz": nz, "nw": nw, "nyloc": nyloc, "nxloc": nxloc,
               "nprocx": res["nprocx"], "nprocy": res["nprocy"]}
        diag = open(os.path.join(wd, "output", "diagnostics.txt")).read()
        assert diag == open(os.path.join(case, "reference.out")).read(), s
        z = np.frombuffer(read_records(os.path.join(wd, "output", "zgrid.bin"))[0], "<f8")
        w = np.frombuffer(read_records(os.path.join(wd, "output", "wavegrid.bin"))[0], "<f8")
        out["zgrid"] = z
        out["wavegrid"] = w
        ncols = nxloc * nyloc
        # keep every column for small blocks, a seeded subset otherwise
        if ncols <= 32:
            pick = np.arange(ncols)
        else:
            pick = np.sort(rng.choice(ncols, 24, replace=False))
            pick = np.unique(np.concatenate([pick, [ncols - 1, 0]]))
        out["cols"] = pick
        for px in range(res["nprocx"]):
            for py in range(res["nprocy"]):
                d = os.path.join(wd, "output", f"procx{px}_procy{py}")
                U = np.frombuffer(read_records(os.path.join(d, "mean_intensity_0000.bin"))[0], "<f8")
                A = np.frombuffer(read_records(os.path.join(d, "absorption_coefficients_0000.bin"))[0], "<f8")
                V = np.frombuffer(read_records(os.path.join(d, "flux_0000.bin"))[0], "<f8")
                U = U.reshape(nw, nxloc, nyloc, mz)
                A = A.reshape(nw, nxloc, nyloc, nz)
                V = V.reshape(nw, nxloc, nyloc, nz)
                tag = f"p{px}_{py}"
                # column index c = ix*nyloc + iy (Fortran (iy, ix) order flattened)
                out[f"U_{tag}"] = U.reshape(nw, ncols, mz)[:, pick]
                out[f"A_{tag}"] = A.reshape(nw, ncols, nz)[:, pick]
                out[f"V_{tag}"] = V.reshape(nw, ncols, nz)[:, pick]
                for nm, arr in (("U", U), ("A", A), ("V", V)):
                    out[f"sum{nm}_{tag}"] = np.array([arr.sum(), np.abs(arr).sum()])
        np.savez_compressed(os.path.join(HERE, f"sample_{s}.npz"), **out)
        print("sample", s, "ok")


def column_profiles(rng, ncol, z, kind, rho_floor=1e-15):
    """Seeded (rho, T) columns, shape (ncol, nz)."""
    nz = z.size
    if kind == "c2":
        Tm = rng.uniform(2e4, 1e5, ncol)[:, None]
        rho0 = 10 ** rng.uniform(-11, -8, ncol)[:, None]
        H = rng.uniform(0.5, 2.0, ncol)[:, None] * AU
        T = np.maximum(Tm * np.exp(-z ** 2 / (2 * 2.54e14 ** 2)), 2000.0)
        rho = np.maximum(rho0 * np.exp(-z ** 2 / (2 * H ** 2)), rho_floor)
    elif kind == "mix":
        T = 10 ** rng.uniform(np.log10(2000), np.log10(2e5), (ncol, nz))
        rho = 10 ** rng.uniform(-20, -8, (ncol, nz))
    elif kind == "saha":
        T = rng.uniform(5000, 2.0e4, (ncol, nz))
        rho = 10 ** rng.uniform(-14, -8, (ncol, nz))
    elif kind == "grey":
        T = 10 ** rng.uniform(np.log10(2000), 5, (ncol, nz))
        rho = 10 ** rng.uniform(-24, -9, (ncol, nz))
    else:
        raise ValueError(kind)
    return rho, T


def run_dump(name, nz, nw, ncol, kind, waves, lgrey=0, full=True, fion=2e4, sigma_grey=0.4,
             rho_floor=1e-15):
    rng = np.random.default_rng(SEED0 + zlib.crc32(name.encode()) % 1000)
    z0, z1 = -5e15, 5e15
    dz = (z1 - z0) / (nz - 1)
    z = z0 + np.arange(nz) * dz
    if kind == "c1":
        rho = np.maximum(1e-9 * np.exp(-z ** 2 / (2 * AU ** 2)), 1e-24)[None, :].repeat(ncol, 0)
        T = np.full((ncol, nz), 2000.0)
    elif kind == "pivot1":
        # grey: the second column's bottom cell at rho = 1e-170 g/cm^3 has
        # alpha = 4e-171 /cm, alpha^2 underflows to 0, so b(1) = 0
        rho = np.maximum(1e-9 * np.exp(-z ** 2 / (2 * AU ** 2)), 1e-24)[None, :].repeat(ncol, 0)
        rho[1, 0] = 1e-170
        T = np.full((ncol, nz), 2000.0)
    else:
        rho, T = column_profiles(rng, ncol, z, kind, rho_floor)
    subprocess.check_call([BUILD, "dump", str(nz), str(nw)], stdout=subprocess.DEVNULL)
    wd = os.path.join(REFOUT, f"dump_nz{nz}_nw{nw}")
    fin = os.path.join(wd, f"{name}.in.bin")
    fnml = os.path.join(wd, f"{name}.nml")
    fout = os.path.join(wd, f"{name}.out.bin")
    with open(fin, "wb") as f:
        f.write(np.array([ncol, lgrey], "<i4").tobytes())
        f.write(np.array([dz, sigma_grey], "<f8").tobytes())
        f.write(np.asarray(waves, "<f8").tobytes())
        f.write(np.ascontiguousarray(rho, "<f8").tobytes())
        f.write(np.ascontiguousarray(T, "<f8").tobytes())
    with open(fnml, "w") as f:
        f.write("&gas_state_input\n  fully_ionized_T = %r\n/\n" % fion)
    run = subprocess.run([os.path.join(wd, "ref_dump.x"), fin, fnml, fout], capture_output=True, text=True)
    raw = np.fromfile(fout, "<f8")
    consts, raw = raw[:32], raw[32:]
    mz = nz + 6
    per_col = 6 * nz + nw * (7 * nz + mz + nz)
    if "ERROR tridiag stage" in run.stdout:
        # the reference stopped on an exactly-zero pivot (source/auxiliary.f90:27-39);
        # record which system it was working on and its message
        stage = 1 if "stage 1" in run.stdout else 2
        done = raw.size
        icol = done // per_col
        iw = (done - icol * per_col - 6 * nz) // (7 * nz + mz + nz)
        np.savez_compressed(os.path.join(HERE, f"dump_{name}.npz"), nz=nz, nw=nw, ncol=ncol, dz=dz, z=z,
                            waves=np.asarray(waves, float), rho=rho, T=T, lgrey=lgrey, sigma_grey=sigma_grey,
                            fully_ionized_T=fion, consts=consts, ref_stop_stage=stage,
                            ref_stop_col=icol + 1, ref_stop_iw=iw + 1, ref_stdout=run.stdout)
        print("dump", name, "reference stopped: stage", stage, "col", icol + 1, "iw", iw + 1)
        return
    assert run.returncode == 0, run.stdout + run.stderr
    raw = raw.reshape(ncol, per_col)
    cell = raw[:, :6 * nz].reshape(ncol, 6, nz)
    pts = raw[:, 6 * nz:].reshape(ncol, nw, 7 * nz + mz + nz)
    out = dict(nz=nz, nw=nw, ncol=ncol, dz=dz, z=z, waves=np.asarray(waves, float), rho=rho, T=T,
               lgrey=lgrey, sigma_grey=sigma_grey, fully_ionized_T=fion, consts=consts,
               alpha=pts[:, :, 0:nz], U=pts[:, :, 7 * nz:7 * nz + mz], V=pts[:, :, 7 * nz + mz:])
    if full:
        out.update(x=cell[:, 0], fn=cell[:, 1], e_scatter=cell[:, 2], hm_bf=cell[:, 3],
                   theta=cell[:, 4], lgtheta=cell[:, 5],
                   albedo=pts[:, :, nz:2 * nz], S=pts[:, :, 2 * nz:3 * nz],
                   a=pts[:, :, 3 * nz:4 * nz], b=pts[:, :, 4 * nz:5 * nz],
                   c=pts[:, :, 5 * nz:6 * nz], d=pts[:, :, 6 * nz:7 * nz])
    np.savez_compressed(os.path.join(HERE, f"dump_{name}.npz"), **out)
    print("dump", name, "ok")


def logwaves(nw, w0=100.0, w1=1e5):
    return np.exp(np.linspace(np.log(w0), np.log(w1), nw))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "pivot1":
        return run_pivot1()
    if len(sys.argv) > 1 and sys.argv[1] == "tiny":
        return run_tiny()
    copy_samples()
    run_samples()
    # C1 exactly (grey isothermal), plus random grey columns
    run_dump("c1_grey", 10, 1, 1, "c1", [3000.0], lgrey=1)
    run_dump("grey_rand", 16, 1, 8, "grey", [3000.0], lgrey=1)
    # every opacity branch: Lyman (<912 A), H- bf cut (>=16000 A), S clamp, neutral/ionised
    w = np.concatenate([logwaves(13), [911.0, 15999.0, 16000.0]])
    run_dump("mix", 32, 16, 16, "mix", w)
    run_dump("saha", 32, 16, 16, "saha", logwaves(16, 500.0, 2e4))
    run_dump("saha_lowfion", 32, 8, 8, "saha", logwaves(8, 800.0, 9000.0), fion=1.2e4)
    # reference-style linear grid and the C2 column profile
    run_dump("c2cols", 256, 32, 2, "c2", logwaves(32), full=False)
    # optically thin floor (rho 1e-24 at 2000 K): dz*alpha < 2^-52 makes the
    # reference's last pivot exactly zero; it prints the error and stops
    run_dump("pivot0", 256, 32, 2, "c2", logwaves(32), full=False, rho_floor=1e-24)
    run_dump("uneven", 257, 3, 4, "c2", [3000.0, 5000.0, 1e4], full=False)
    run_dump("long", 1024, 4, 2, "c2", logwaves(4, 500.0, 2e4), full=False)
    run_pivot1()
    run_tiny()


def run_tiny():
    """Columns of 2 and 3 cells: both rows are boundary rows (nz = 2) or one
    interior row (nz = 3), and update_ghosts (source/auxiliary.f90:49-62)
    reads ghosts it set earlier in the same loop (f(n1+i) beyond n2)."""
    run_dump("tiny2", 2, 8, 8, "mix", logwaves(8))
    run_dump("tiny3", 3, 8, 8, "mix", logwaves(8))
    run_dump("tiny3_grey", 3, 1, 4, "grey", [3000.0], lgrey=1)


def run_pivot1():
    """alpha^2 underflowing at the boundary row: the reference's first pivot
    b(1) is exactly zero and it stops in stage 1 (source/auxiliary.f90:27-30)."""
    run_dump("pivot1", 16, 1, 3, "pivot1", [3000.0], lgrey=1, full=False)


if __name__ == "__main__":
    sys.exit(main())
