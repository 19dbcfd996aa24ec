#!/bin/sh
# Recipe that compiles the ocotillo reference (Fortran 90) from its sources
# where they lie under /root/reference/source, unchanged, with amdflang and
# the reference's FP64 promotion (source/Makefile.source:8 minus the gfortran-only
# debug flags, which flang does not accept).  Test infrastructure only: every
# output goes under oracle/_ref/, which is never committed.
#
#   build_ref.sh sample <name> <case_dir>
#       case_dir holds input.in + resolution.in (+ input_athena/); builds
#       oracle/_ref/<name>/source/flux_feautrier.x, the reference executable.
#   build_ref.sh dump <nz> <nw>
#       builds oracle/_ref/dump_nz<nz>_nw<nw>/ref_dump.x: the reference modules
#       plus the golden-vector driver oracle/ref/ref_dump.f90.
#   build_ref.sh time <nz> <nw>
#       builds oracle/_ref/time_nz<nz>_nw<nw>/ref_time.x: the reference modules
#       at -O3 (source/Makefile.source:8's optimisation level) plus the timing
#       driver oracle/ref/ref_time.f90 (the CPU baseline beside the GPU).
set -e
REF=${OCO_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$(cd "$HERE/.." && pwd)/_ref
FC=${FC:-amdflang}
FFLAGS="-O2 -fdefault-real-8 -fdefault-double-8"
mkdir -p "$OUT"

case "$1" in
sample)
  name=$2; case_dir=$(cd "$3" && pwd)
  dst="$OUT/$name"
  rm -rf "$dst"; mkdir -p "$dst/source" "$dst/output"
  cp "$case_dir/input.in" "$case_dir/resolution.in" "$dst/"
  [ -d "$case_dir/input_athena" ] && ln -s "$case_dir/input_athena" "$dst/input_athena"
  cd "$dst/source"
  for f in common disk auxiliary continuous_opacity gas_state io read_athena flux_feautrier; do
    $FC $FFLAGS -I. -c "$REF/source/$f.f90" -o $f.o
  done
  $FC common.o auxiliary.o disk.o continuous_opacity.o gas_state.o read_athena.o io.o flux_feautrier.o -o flux_feautrier.x
  ;;
dump)
  nz=$2; nw=$3
  dst="$OUT/dump_nz${nz}_nw${nw}"
  mkdir -p "$dst/obj"
  printf 'integer, parameter :: nw=%s\ninteger, parameter :: nx=1,ny=1,nz=%s\ninteger, parameter :: nprocx=1,nprocy=1,nprocz=1\ninteger, parameter :: nyloc=ny/nprocy,nxloc=nx/nprocx\n' "$nw" "$nz" > "$dst/resolution.in"
  cd "$dst/obj"
  for f in common continuous_opacity gas_state auxiliary; do
    $FC $FFLAGS -I. -c "$REF/source/$f.f90" -o $f.o
  done
  $FC $FFLAGS -I. -c "$HERE/ref_dump.f90" -o ref_dump.o
  $FC common.o continuous_opacity.o gas_state.o auxiliary.o ref_dump.o -o ../ref_dump.x
  ;;
time)
  nz=$2; nw=$3
  dst="$OUT/time_nz${nz}_nw${nw}"
  mkdir -p "$dst/obj"
  printf 'integer, parameter :: nw=%s\ninteger, parameter :: nx=1,ny=1,nz=%s\ninteger, parameter :: nprocx=1,nprocy=1,nprocz=1\ninteger, parameter :: nyloc=ny/nprocy,nxloc=nx/nprocx\n' "$nw" "$nz" > "$dst/resolution.in"
  cd "$dst/obj"
  TFLAGS="-O3 -fdefault-real-8 -fdefault-double-8"
  for f in common continuous_opacity gas_state auxiliary; do
    $FC $TFLAGS -I. -c "$REF/source/$f.f90" -o $f.o
  done
  $FC $TFLAGS -I. -c "$HERE/ref_time.f90" -o ref_time.o
  $FC common.o continuous_opacity.o gas_state.o auxiliary.o ref_time.o -o ../ref_time.x
  rm -f ./*.o ./*.mod
  ;;
*)
  echo "usage: $0 sample <name> <case_dir> | dump <nz> <nw> | time <nz> <nw>" >&2; exit 2;;
esac
