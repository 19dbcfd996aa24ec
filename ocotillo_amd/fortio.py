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
