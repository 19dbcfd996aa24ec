"""Readers for the drop-in's outputs (the reference's formats, source/io.f90:31-101).

Unformatted sequential files hold one record per array, framed by 4-byte
little-endian length markers; records longer than 2**31-1 bytes are split
into sub-records whose markers are negated for all but the last, as
gfortran writes them.  Arrays come back in C order: U as [iw, ix, iy, mz],
alpha and V as [iw, ix, iy, nz] (Fortran (z, iy, ix, iw)).
"""
import numpy as np


def read_records(path):
    """All records of an unformatted sequential file, as bytes objects."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        parts = []
        while True:
            n = int(np.frombuffer(data, "<i4", 1, pos)[0])
            size = abs(n)
            parts.append(data[pos + 4: pos + 4 + size])
            pos += 8 + size
            if n >= 0:
                break
        out.append(b"".join(parts))
    return out


def read_array(path, shape_c):
    """The single FP64 record of `path`, reshaped to the C-order shape."""
    rec = read_records(path)[0]
    return np.frombuffer(rec, "<f8").reshape(shape_c)


def read_block(outdir, iprocx, iprocy, snapshot, nw, nxloc, nyloc, nz):
    """(U, alpha, V) of one processor block written by output_binary."""
    d = "%s/procx%d_procy%d" % (outdir, iprocx, iprocy)
    U = read_array("%s/mean_intensity_%s.bin" % (d, snapshot), (nw, nxloc, nyloc, nz + 6))
    A = read_array("%s/absorption_coefficients_%s.bin" % (d, snapshot), (nw, nxloc, nyloc, nz))
    V = read_array("%s/flux_%s.bin" % (d, snapshot), (nw, nxloc, nyloc, nz))
    return U, A, V


def format_e(x, w=15, d=5):
    """Fortran Ew.d output of one value: 0.ddddd scaled, exponent E+xx, or
    +xxx (no E) when it needs three digits, right-justified in w columns."""
    x = float(x)
    if x == 0.0:
        digits, e = "0" * d, 0
    else:
        s = "%.*e" % (d - 1, abs(x))            # d.ddddde+XX, correctly rounded
        mant, ex = s.split("e")
        digits = mant.replace(".", "")
        e = int(ex) + 1
    body = "0." + digits + ("E%+03d" % e if -99 <= e <= 99 else "%+04d" % e)
    if x < 0 or (x == 0.0 and np.signbit(x)):
        body = "-" + body
    return body.rjust(w)


def diagnostics_text(U_col, A_col):
    """diagnostics.txt (source/io.f90:31-46): for each cell i, I6 then the nw
    mean intensities U(n1+i-1, :) and the nw absorption coefficients, each
    E15.5.  U_col is [nw, nz+6], A_col [nw, nz]."""
    nw, nz = A_col.shape
    lines = []
    for i in range(nz):
        vals = [U_col[iw, 3 + i] for iw in range(nw)] + [A_col[iw, i] for iw in range(nw)]
        lines.append("%6d" % (i + 1) + "".join(format_e(v) for v in vals))
    return "\n".join(lines) + "\n"
