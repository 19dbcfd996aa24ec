"""Snapshot job source on the GPU.

A run directory with three snapshots (multiInTest's 0000 as shipped, 0001
and 0002 made from it by scaling the density, momenta and energy) and:
  * the Fortran host given 0:2 writes output_snapshot<tag>/ for each, every
    file identical to a separate one-snapshot run of that tag (output/), and
    0000's diagnostics.txt is the reference's reference.out;
  * python -m ocotillo_amd.snapshots writes the same binary files and
    diagnostics.txt;
  * served as jobs (--emergent), each block's emergent flux equals V(1),
    V(nz) of the written flux files.
"""
import filecmp
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, REPO

pytestmark = pytest.mark.gpu
EXE = os.path.join(REPO, "ocotillo_amd", "fortran", "flux_feautrier_gpu.x")
SRC = os.path.join(GOLDEN, "samples", "multiInTest")
TAGS = ["0000", "0001", "0002"]


def make_run(d):
    from ocotillo_amd.athena import SlabHeader, write_slab
    os.makedirs(d)
    for f in ("input.in", "resolution.in"):
        shutil.copy(os.path.join(SRC, f), os.path.join(d, f))
    base = os.path.join(SRC, "input_athena", "distributed")
    for idd in sorted(os.listdir(base)):
        os.makedirs(os.path.join(d, "input_athena", "distributed", idd))
        for f in os.listdir(os.path.join(base, idd)):
            src = os.path.join(base, idd, f)
            h = SlabHeader(src)
            ints = tuple(int(v) for v in np.fromfile(src, "<i4", 8)[4:])
            for k, tag in enumerate(TAGS):
                out = os.path.join(d, "input_athena", "distributed", idd, f.replace(".0000.", ".%s." % tag))
                if k == 0:
                    shutil.copy(src, out)
                    continue
                rho, mx, my, mz, e = (np.asarray(a) for a in h.fields())
                fields = [rho * (1 + 0.25 * k), mx * (1 - 0.1 * k), my, mz * (1 + 0.1 * k), e * (1 + 0.5 * k)]
                write_slab(out, fields, h.x, h.y, h.z, h.gamma1, h.coordsys, ints, h.cs, h.t, h.dt)
    return d


def run_host(d, args):
    print("host", d, args, flush=True)
    r = subprocess.run([EXE] + args, cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    return r.stdout


def tree(d):
    return sorted(os.path.relpath(os.path.join(p, f), d) for p, _, fs in os.walk(d) for f in fs)


def same_files(a, b, names):
    for n in names:
        assert filecmp.cmp(os.path.join(a, n), os.path.join(b, n), shallow=False), n


@pytest.fixture(scope="module")
def host_multi(tmp_path_factory):
    d = make_run(str(tmp_path_factory.mktemp("multi") / "run"))
    out = run_host(d, ["0:2"])
    for tag in TAGS:
        assert "snapshot %s done" % tag in " ".join(out.split())
    return d


def test_host_multi_equals_single_runs(host_multi, tmp_path):
    single = make_run(str(tmp_path / "single"))
    for tag in TAGS:
        run_host(single, [tag])
        a = os.path.join(host_multi, "output_snapshot" + tag)
        b = os.path.join(single, "output")
        assert tree(a) == tree(b)
        same_files(a, b, tree(a))
        shutil.rmtree(b)
    ref = open(os.path.join(SRC, "reference.out")).read()
    assert open(os.path.join(host_multi, "output_snapshot0000", "diagnostics.txt")).read() == ref
    d1 = open(os.path.join(host_multi, "output_snapshot0001", "diagnostics.txt")).read()
    assert d1 != ref and len(d1) == len(ref)


def test_python_source_equals_host(host_multi, tmp_path):
    import ocotillo_amd as oco
    from ocotillo_amd.snapshots import Run, write_snapshots
    d = make_run(str(tmp_path / "py"))
    with oco.Context(0) as ctx:
        roots = write_snapshots(ctx, Run(d), TAGS, stage_bytes=4 << 20)
    for tag in TAGS:
        a = os.path.join(host_multi, "output_snapshot" + tag)
        b = roots[tag]
        names = [n for n in tree(a) if not n.endswith(("zgrid.txt", "wavegrid.txt"))]
        assert names == tree(b)
        same_files(a, b, names)


def test_served_emergent_flux(host_multi, tmp_path):
    import ocotillo_amd as oco
    from ocotillo_amd.fortio import read_block
    from ocotillo_amd.snapshots import Run, serve_snapshots
    d = make_run(str(tmp_path / "srv"))
    run = Run(d)
    s = run.shapes
    with oco.Context(0) as ctx:
        out = serve_snapshots(ctx, run, TAGS, mode=oco.MODE_FP64, hbm_budget=1 << 30)
    assert sorted(out) == sorted((t, px, py) for t in TAGS for px, py in run.blocks())
    for (tag, px, py), em in out.items():
        _, _, V = read_block(os.path.join(host_multi, "output_snapshot" + tag), px, py, tag, s["nw"], s["nxloc"],
                             s["nyloc"], s["nz"])
        assert np.array_equal(em[..., 0], V[..., 0]) and np.array_equal(em[..., 1], V[..., -1])
