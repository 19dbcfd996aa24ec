/*
 * oco_output.cpp -- the reference's binary output files
 * (source/io.f90:76-101): one Fortran unformatted sequential record per
 * file, in the layout the reference's compiler (gfortran,
 * source/Makefile.source:4) writes.
 *
 * A record of N bytes is stored as subrecords of at most 2147483639 bytes
 * (2^31 - 9), each framed by a leading and a trailing 4-byte little-endian
 * length.  The leading length is negated on every subrecord but the last
 * (more follow); the trailing length is negated on every subrecord but the
 * first (it continues a previous one).  A record that fits one subrecord is
 * the familiar  [N][data][N].  This is the format described in the GNU
 * Fortran manual ("File format of unformatted sequential files").
 *
 * Files are written under <path>.partial and renamed when complete, so a
 * failed block leaves earlier files with the same name untouched.
 */
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <new>
#include <string>
#include <sys/stat.h>
#include <unistd.h>

#include "../../include/ocotillo.h"
#include "oco_internal.h"

static const uint64_t SUBREC = 2147483639ULL;

struct oco_record_file {
    int fd = -1;
    uint64_t n = 0;
    std::string path, tmp;
};

static int pwrite_all(int fd, const void* data, size_t n, uint64_t off)
{
    const char* p = (const char*)data;
    while (n > 0) {
        ssize_t w = pwrite(fd, p, n, (off_t)off);
        if (w < 0) {
            if (errno == EINTR) continue;
            return OCO_E_IO;
        }
        p += w;
        n -= (size_t)w;
        off += (uint64_t)w;
    }
    return OCO_OK;
}

/* File offset of data byte d of the record. */
static uint64_t file_offset(uint64_t d) { return 4 + d + 8 * (d / SUBREC); }

int oco_record_open(const char* path, unsigned long long nbytes, oco_record_file** out)
{
    *out = nullptr;
    oco_record_file* f = new (std::nothrow) oco_record_file;
    if (!f) return OCO_E_OOM;
    f->n = nbytes;
    f->path = path;
    f->tmp = f->path + ".partial";
    f->fd = open(f->tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
    if (f->fd < 0) {
        fprintf(stderr, "ocotillo: cannot create %s: %s\n", f->tmp.c_str(), strerror(errno));
        delete f;
        return OCO_E_IO;
    }
    /* the markers of every subrecord; the data fills the gaps */
    const uint64_t nsub = nbytes == 0 ? 1 : (nbytes + SUBREC - 1) / SUBREC;
    for (uint64_t s = 0; s < nsub; s++) {
        const uint64_t len = s + 1 < nsub ? SUBREC : nbytes - s * SUBREC;
        const int32_t lead = (int32_t)(s + 1 < nsub ? -(int64_t)len : (int64_t)len);
        const int32_t trail = (int32_t)(s > 0 ? -(int64_t)len : (int64_t)len);
        const uint64_t at = s * (SUBREC + 8);
        if (pwrite_all(f->fd, &lead, 4, at) || pwrite_all(f->fd, &trail, 4, at + 4 + len)) {
            oco_record_close(f, 0);
            return OCO_E_IO;
        }
    }
    *out = f;
    return OCO_OK;
}

int oco_record_write(oco_record_file* f, unsigned long long off, const void* data, size_t n)
{
    if (off + n > f->n) return OCO_E_ARG;
    const char* p = (const char*)data;
    while (n > 0) {   /* split at subrecord boundaries */
        const uint64_t room = SUBREC - off % SUBREC;
        const size_t k = n < room ? n : (size_t)room;
        if (int st = pwrite_all(f->fd, p, k, file_offset(off))) return st;
        p += k;
        off += k;
        n -= k;
    }
    return OCO_OK;
}

int oco_record_close(oco_record_file* f, int keep)
{
    if (!f) return OCO_OK;
    int st = OCO_OK;
    if (f->fd >= 0 && close(f->fd) != 0) st = OCO_E_IO;
    if (keep && st == OCO_OK) {
        if (rename(f->tmp.c_str(), f->path.c_str()) != 0) st = OCO_E_IO;
    }
    if (!keep || st != OCO_OK) unlink(f->tmp.c_str());
    delete f;
    return st;
}

extern "C" int oco_write_record(const char* path, const void* data, size_t nbytes)
{
    if (!path || (!data && nbytes)) return OCO_E_ARG;
    oco_record_file* f = nullptr;
    int st = oco_record_open(path, nbytes, &f);
    if (st) return st;
    st = oco_record_write(f, 0, data, nbytes);
    int st2 = oco_record_close(f, st == OCO_OK);
    return st ? st : st2;
}
