"""Fortran unformatted sequential records as the reference's compiler
(gfortran, source/Makefile.source:4) writes them (source/io.f90:57-99), via
oco_write_record -- host code, no GPU.

The layout is the GNU Fortran manual's: a record is one or more subrecords
of at most 2147483639 bytes, each framed by 4-byte little-endian lengths;
the leading length is negated on all subrecords but the last, the trailing
one on all but the first.  Checked byte for byte on small records, on an
empty one, and on a record 40 bytes over one subrecord (a real 2 GiB file);
the package's reader (ocotillo_amd.fortio) reads each back.  amdflang, which
builds the Fortran host here, truncates the length of such a record to 32
bits instead, which is why the outputs are written by this library.
"""
import os

import numpy as np
import pytest

SUBREC = 2147483639


def marker(path, pos):
    with open(path, "rb") as f:
        f.seek(pos)
        return int(np.frombuffer(f.read(4), "<i4")[0])


@pytest.mark.parametrize("n", [0, 1, 5, 1000])
def test_small_records(tmp_path, n):
    from ocotillo_amd import write_record
    from ocotillo_amd.fortio import read_records
    a = np.arange(n, dtype=np.float64) * 1.5 - 3
    p = tmp_path / "r.bin"
    write_record(p, a)
    raw = p.read_bytes()
    assert raw == np.int32(8 * n).tobytes() + a.tobytes() + np.int32(8 * n).tobytes()
    assert read_records(str(p)) == [a.tobytes()]
    assert not (tmp_path / "r.bin.partial").exists()


def test_replaces_existing_file(tmp_path):
    from ocotillo_amd import write_record
    p = tmp_path / "r.bin"
    p.write_bytes(b"x" * 100)
    write_record(p, np.ones(3))
    assert p.stat().st_size == 8 + 24


def test_unwritable_path_reports_io_error(tmp_path):
    import ocotillo_amd as oco
    with pytest.raises(oco.OcoError) as e:
        oco.write_record(tmp_path / "missing_dir" / "r.bin", np.ones(3))
    assert e.value.status == oco.E_IO


def test_record_over_one_subrecord(tmp_path):
    """2147483639 + 41 bytes: the data split after byte 2147483639, markers
    [-L][..L..][L] [41][..41..][-41]."""
    from ocotillo_amd import write_record
    n = (SUBREC + 41) // 8 + 1          # doubles: 2^28 + 5 -> 2147483688 bytes
    nbytes = 8 * n
    rest = nbytes - SUBREC
    a = np.arange(n, dtype=np.float64)
    p = str(tmp_path / "big.bin")
    write_record(p, a)
    assert os.path.getsize(p) == nbytes + 16
    assert marker(p, 0) == -SUBREC
    assert marker(p, 4 + SUBREC) == SUBREC
    assert marker(p, SUBREC + 8) == rest
    assert marker(p, SUBREC + 12 + rest) == -rest
    mm = np.memmap(p, np.uint8, "r")
    src = a.view(np.uint8)
    assert np.array_equal(mm[4:4 + SUBREC], src[:SUBREC])
    assert np.array_equal(mm[SUBREC + 12:SUBREC + 12 + rest], src[SUBREC:])
    del mm
    from ocotillo_amd.fortio import read_records
    recs = read_records(p)
    assert len(recs) == 1 and len(recs[0]) == nbytes
    assert np.array_equal(np.frombuffer(recs[0], np.float64), a)
