"""Athena-4.2 snapshot input for the GPU path.

The host side of source/read_athena.f90: the run directory's namelists
(input.in, resolution.in), the per-process .bin files and the code units of
read_athena_input (:33-62).  The arithmetic that turns the raw fields into
the block's density and temperature (:163-268 and the temperature rule of
source/disk.f90:96-104) runs on the device, one launch per z slab
(oco_athena_slab); this module only reads bytes and sets up scalars.

    run = AthenaRun("samples/ReadFromAthena")
    with Context(0) as ctx:
        rho3d, T3d, z, dz = run.load_block(ctx, iprocx=0, iprocy=0, snapshot="0000")

rho3d and T3d come back C-ordered [nxloc, nyloc, nz] (Fortran
(nz, nyloc, nxloc)), ready for Context.solve_block.
"""
import ast
import math
import os
import re

import numpy as np

# source/common.f90:26, 33, 49-51, 57, 76-77 (8-byte reals)
EPSI = 2.0 ** -52
AMU = 1.66053906660e-24
K1_CGS = 1.0 / 1.380649e-16
C_LIGHT_CGS = 2.99792458e10
G_NEWTON_CGS = 6.6742e-8
SOLAR_MASS = 1.988409870698051e+33
NG = 3


# ---------------------------------------------------------------------------
# input files

_TOKEN = re.compile(r"""'([^']*)'|"([^"]*)"|(![^\n]*)|(=)|(/)|([^\s,=/!'"]+)""")


def _value(tok):
    low = tok.lower()
    if low in ("t", ".t.", ".true.", "true"):
        return True
    if low in ("f", ".f.", ".false.", "false"):
        return False
    try:
        return int(tok)
    except ValueError:
        return float(low.replace("d", "e"))


def read_namelist(path, group):
    """One namelist group of a Fortran namelist file as {lower-case name: value}
    (scalars only).  Strings keep their text, logicals become bool, numbers
    int or float (a d exponent is read as e).  KeyError if the group is absent."""
    text = open(path).read()
    m = re.search(r"&" + re.escape(group) + r"\b", text, re.I)
    if not m:
        raise KeyError("namelist group &%s not found in %s" % (group, path))
    toks = []
    for t in _TOKEN.finditer(text, m.end()):
        sq, dq, comment, eq, slash, word = t.groups()
        if comment is not None:
            continue
        if slash is not None:
            break
        toks.append(("=", None) if eq is not None else ("s", sq if sq is not None else dq) if word is None
                    else ("w", word))
    out, name, i = {}, None, 0
    while i < len(toks):
        kind, v = toks[i]
        if kind == "w" and i + 1 < len(toks) and toks[i + 1][0] == "=":
            name, i = v.lower(), i + 2
            continue
        if name is not None and name not in out:
            out[name] = v if kind == "s" else _value(v)
        i += 1
    return out


def _int_eval(node, names):
    if isinstance(node, ast.Expression):
        return _int_eval(node.body, names)
    if isinstance(node, ast.Constant) and isinstance(node.value, int):
        return node.value
    if isinstance(node, ast.Name):
        return names[node.id.lower()]
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _int_eval(node.operand, names)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.BinOp):
        a, b = _int_eval(node.left, names), _int_eval(node.right, names)
        if isinstance(node.op, ast.Add):
            return a + b
        if isinstance(node.op, ast.Sub):
            return a - b
        if isinstance(node.op, ast.Mult):
            return a * b
        if isinstance(node.op, ast.Div):   # Fortran integer division truncates toward zero
            q = abs(a) // abs(b)
            return q if (a >= 0) == (b >= 0) else -q
    raise ValueError("unsupported expression in resolution.in")


def read_resolution(path):
    """resolution.in (source/common.f90:9 includes it): lines
    'integer, parameter :: a=expr, b=expr ! comment' -> {name: int}."""
    names = {}
    for line in open(path):
        line = line.split("!", 1)[0]
        if "::" not in line:
            continue
        for item in line.split("::", 1)[1].split(","):
            if "=" not in item:
                continue
            k, e = item.split("=", 1)
            names[k.strip().lower()] = _int_eval(ast.parse(e.strip(), mode="eval"), names)
    names.setdefault("nprocz", 1)
    return names


def nint(x):
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


# ---------------------------------------------------------------------------
# slab files

class SlabHeader:
    """Header of one Athena .bin file (source/read_athena.f90:123-150):
    coordsys, nx, ny, nz, nvar, nscalars, selfgrav, particles (int32), gamma1,
    cs, t, dt (float64), x, y, z coordinates; the five fields follow."""

    def __init__(self, path):
        self.path = path
        with open(path, "rb") as f:
            ints = np.frombuffer(f.read(32), "<i4")
            self.coordsys, self.nxloc, self.nyloc, self.nzl = (int(v) for v in ints[:4])
            self.gamma1, self.cs, self.t, self.dt = np.frombuffer(f.read(32), "<f8")
            self.x = np.frombuffer(f.read(8 * self.nxloc), "<f8")
            self.y = np.frombuffer(f.read(8 * self.nyloc), "<f8")
            self.z = np.frombuffer(f.read(8 * self.nzl), "<f8")
        self.offset = 64 + 8 * (self.nxloc + self.nyloc + self.nzl)
        n = self.nxloc * self.nyloc * self.nzl
        need = self.offset + 5 * 8 * n
        if os.path.getsize(path) < need:
            raise ValueError("%s: %d bytes, the header promises %d" % (path, os.path.getsize(path), need))

    def fields(self):
        """(rho, Mx, My, Mz, E) as C-ordered [nzl, nyloc, nxloc] (the file's
        Fortran (nxloc, nyloc, nzl)), memory-mapped."""
        n = self.nxloc * self.nyloc * self.nzl
        mm = np.memmap(self.path, "<f8", "r", offset=self.offset, shape=(5, self.nzl, self.nyloc, self.nxloc))
        assert mm[0].size == n
        return [mm[i] for i in range(5)]


def write_slab(path, fields, x, y, z, gamma1, coordsys=-1, header_ints=(5, 0, 0, 0), cs=0.0, t=0.0, dt=0.0):
    """An Athena .bin file in the layout SlabHeader reads: `fields` are the
    five C-ordered [nzl, nyloc, nxloc] arrays (rho, Mx, My, Mz, E)."""
    f0 = np.asarray(fields[0])
    nzl, nyloc, nxloc = f0.shape
    if len(fields) != 5 or any(np.shape(a) != f0.shape for a in fields):
        raise ValueError("five fields of one shape are needed")
    if (len(x), len(y), len(z)) != (nxloc, nyloc, nzl):
        raise ValueError("coordinates do not match the fields")
    with open(path, "wb") as f:
        np.array([coordsys, nxloc, nyloc, nzl] + list(header_ints), "<i4").tofile(f)
        np.array([gamma1, cs, t, dt], "<f8").tofile(f)
        for a in (x, y, z):
            np.asarray(a, "<f8").tofile(f)
        for a in fields:
            np.ascontiguousarray(a, "<f8").tofile(f)


class AthenaRun:
    """A run directory laid out as the reference expects (input.in,
    resolution.in, <datadir>/id<iproc>/<RunName>[-id<iproc>].<snap>.bin)."""

    def __init__(self, run_dir, inputfile="input.in", resolution="resolution.in"):
        self.dir = run_dir
        inp = os.path.join(run_dir, inputfile)
        self.shapes = read_resolution(os.path.join(run_dir, resolution))
        a = read_namelist(inp, "athena_input")
        t = read_namelist(inp, "temperature_input")
        self.run_name = a["runname"]
        self.datadir = a.get("datadir", "./input_athena/distributed")
        self.avoid_momentum_underflow = bool(a.get("lavoid_momentum_underflow", True))
        if "rho_floor" not in a or a["rho_floor"] != a["rho_floor"]:
            raise ValueError("Must set `rho_floor` (cgs) value in namelist: `input.in/athena_input`.")
        self.rho_floor = float(a["rho_floor"])
        self.unit_length, self.unit_density, self.unit_temperature = code_units(
            float(a["mbh_solarmasses"]), float(a["r0ref_rg"]), float(a["aspect_ratio"]),
            float(a["mean_molecular_weight"]), float(a["rho0"]))
        self.mid_temp, self.sigma = float(t["midtemp"]), float(t["sigma"])
        self.floor_temp = float(t["floortemp"])
        self.switch_temp = (1 + EPSI) * float(t["isotemp"])   # source/disk.f90:28
        self._tgauss = None

    def snapshots(self):
        """Tags of the snapshots present for process 0, sorted."""
        d = os.path.join(self.dir, self.datadir, "id0")
        pre, suf = self.run_name + ".", ".bin"
        if not os.path.isdir(d):
            return []
        return sorted(f[len(pre):-len(suf)] for f in os.listdir(d) if f.startswith(pre) and f.endswith(suf))

    def slab_path(self, iprocx, iprocy, iprocz, snapshot="0000"):
        """source/read_athena.f90:102-117."""
        s = self.shapes
        npx, npy, npz = s["nprocx"], s["nprocy"], s["nprocz"]
        iproc = (iprocz % npz) * npx * npy + (iprocy % npy) * npx + (iprocx % npx)
        head = os.path.join(self.dir, self.datadir, "id%d" % iproc, self.run_name)
        return head + ("" if iproc == 0 else "-id%d" % iproc) + "." + snapshot + ".bin"

    def block_layout(self, iprocx, iprocy, snapshot="0000"):
        """Headers of the block's z slabs, each slab's 0-based z offset, the
        physical z grid (mz points, ghosts included), dz and gamma1 (the last
        slab's, as the reference uses it): source/read_athena.f90:100-160 and
        :245-256."""
        s = self.shapes
        nz, nxloc, nyloc = s["nz"], s["nxloc"], s["nyloc"]
        heads, offs = [], []
        zn = np.zeros(nz)
        z0 = dzc = 0.0
        for iprocz in range(s["nprocz"]):
            p = self.slab_path(iprocx, iprocy, iprocz, snapshot)
            if not os.path.exists(p):
                raise FileNotFoundError("File %s does not exist" % p)
            h = SlabHeader(p)
            if (h.nxloc, h.nyloc) != (nxloc, nyloc):
                raise ValueError("%s: Athena block %dx%d, resolution.in %dx%d"
                                 % (p, h.nxloc, h.nyloc, nxloc, nyloc))
            if iprocz == 0:
                z0 = float(h.z[0])
                dzc = float(h.z[1]) - float(h.z[0])
            iz0 = nint((float(h.z[0]) - z0) / dzc)
            if iz0 < 0 or iz0 + h.nzl > nz:
                raise ValueError("%s: slab z range outside nz=%d" % (p, nz))
            zn[iz0:iz0 + h.nzl] = h.z
            heads.append(h)
            offs.append(iz0)
        zn = zn * self.unit_length
        dz = (zn[nz - 1] - zn[0]) / (nz - 1)
        z = np.zeros(nz + 2 * NG)
        z[NG:NG + nz] = zn
        for i in range(1, NG + 1):
            z[NG - i] = z[NG] - i * dz
            z[NG + nz - 1 + i] = z[NG + nz - 1] + i * dz
        return heads, offs, z, dz, float(heads[-1].gamma1)

    def tgauss(self, z):
        """The Gaussian temperature profile midTemp exp(-z^2/(2 sigma^2)) of
        the first calc_temperature call, kept for later blocks
        (source/disk.f90:87-90); libm exp, as the reference's."""
        if self._tgauss is None:
            zz = z[NG:len(z) - NG]
            s2 = self.sigma * self.sigma
            self._tgauss = np.array([self.mid_temp * math.exp(-((0.5 * (v * v)) / s2)) for v in zz])
        return self._tgauss

    def units(self, gamma1):
        from . import AthenaUnits
        return AthenaUnits(self.unit_density, self.unit_temperature, gamma1, self.rho_floor, self.switch_temp,
                           self.floor_temp, int(self.avoid_momentum_underflow))

    def load_block(self, ctx, iprocx, iprocy, snapshot="0000"):
        """rho3d, T3d ([nxloc, nyloc, nz]), z (mz) and dz of one processor
        block; the fields are converted on the device, slab by slab."""
        heads, offs, z, dz, gamma1 = self.block_layout(iprocx, iprocy, snapshot)
        s = self.shapes
        rho3d = np.zeros((s["nxloc"], s["nyloc"], s["nz"]))
        T3d = np.zeros_like(rho3d)
        units = self.units(gamma1)
        tg = self.tgauss(z)
        for h, iz0 in zip(heads, offs):
            ctx.athena_slab(h.fields(), tg, units, iz0, rho3d, T3d)
        return rho3d, T3d, z, dz


def code_units(Mbh_SolarMasses, r0ref_rg, aspect_ratio, mean_molecular_weight, rho0):
    """read_athena_input's code units (source/read_athena.f90:42-56) in its
    operation order: (unit_length, unit_density, unit_temperature)."""
    Mbh = Mbh_SolarMasses * SOLAR_MASS
    rg = (2 * (G_NEWTON_CGS / (C_LIGHT_CGS * C_LIGHT_CGS))) * Mbh
    rr = r0ref_rg * rg
    g0 = G_NEWTON_CGS * Mbh
    Omega = math.sqrt(g0) * math.pow(rr, -1.5)
    H = aspect_ratio * rr
    unit_time = 1.0 / Omega
    unit_velocity = H / unit_time
    unit_temperature = ((mean_molecular_weight * AMU) * K1_CGS) * (unit_velocity * unit_velocity)
    return H, rho0, unit_temperature
