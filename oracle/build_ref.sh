#!/usr/bin/env bash
# Builds the reference's serial solver (dgswem_serial) from its Fortran sources where they lie
# under /root/reference/src, into oracle/_ref/.  Nothing is copied or patched: the file list is the
# reference's own (src/meson.build:3-59 minus its commented-out entries, plus c_interface_src at
# src/meson.build:62), the flags are the ones its build sets for a double-precision serial build
# (meson.build:6-9, 28-34).  read_fort_dg.F90 is compiled at -O0: FORT_DG_SETUP copies and compares
# an uninitialised CHARACTER(15) local (src/read_fort_dg.F90:428,438,489,510), which flang -O2 may
# fold to "unequal"; at -O0 the copy and the compare see the same bytes.
# Fortran module dependencies are resolved by retrying the files that failed in the previous pass.
set -euo pipefail
REF=${DGSWEM_REFERENCE:-/root/reference}
S=$REF/src
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
OBJ=$OUT/obj
FC=${FC:-flang}
FF="-cpp -O2 -ffixed-line-length-132 -fdefault-real-8 -fdefault-double-8 -DRKSSP -DLINUX -DSLOPE5"
[ -d "$S" ] || { echo "build_ref: $S not found; nothing to build" >&2; exit 0; }
mkdir -p "$OBJ"
LIST="c_interface.F90 $(sed -n "/dgswem_src = files/,/\])/p" "$S/meson.build" | grep -v "^ *#" | grep -o "^'[^']*'" | tr -d "'")"
cd "$OBJ"
todo="$LIST"
for pass in 1 2 3 4 5 6; do
  next=""
  for f in $todo; do
    o=${f%.*}.o
    [ -f "$o" ] && [ "$o" -nt "$S/$f" ] && continue
    flags="$FF"
    [ "$f" = read_fort_dg.F90 ] && flags="${FF/-O2/-O0}"
    $FC $flags -c "$S/$f" -o "$o" >"$o.log" 2>&1 || next="$next $f"
  done
  todo="$next"
  [ -z "$todo" ] && break
done
if [ -n "$todo" ]; then echo "build_ref: could not compile:$todo" >&2; exit 1; fi
g++ -O2 -c "$S/count_edges.cpp" -o count_edges.o
$FC -o "$OUT/dgswem_serial" *.o -lstdc++
echo "build_ref: built $OUT/dgswem_serial"
