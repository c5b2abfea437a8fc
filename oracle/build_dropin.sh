#!/usr/bin/env bash
# Links the reference's serial solver with the GPU step as a drop-in: every object file that
# oracle/build_ref.sh compiled from /root/reference/src except DG_timestep.o, plus
# dgswem_amd/fortran/{gpu_interface,dg_timestep_gpu}.F90 and dgswem_amd/libdgswem_hip.so.
# Output: oracle/_ref/dgswem_gpu (reads fort.14/15/dg/13 and writes fort.6x like dgswem_serial).
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(dirname "$HERE")
OBJ=$HERE/_ref/obj
OUT=$HERE/_ref/dropin
FC=${FC:-flang}
FF="-cpp -O2 -ffixed-line-length-132 -fdefault-real-8 -fdefault-double-8 -DRKSSP -DLINUX -DSLOPE5"
[ -f "$OBJ/dgswem.o" ] || { echo "build_dropin: run oracle/build_ref.sh first" >&2; exit 1; }
[ -f "$ROOT/dgswem_amd/libdgswem_hip.so" ] || { echo "build_dropin: build the HIP library first" >&2; exit 1; }
mkdir -p "$OUT"
cd "$OUT"
$FC $FF -I "$OBJ" -c "$ROOT/dgswem_amd/fortran/gpu_interface.F90" -o gpu_interface.o
$FC $FF -I "$OBJ" -I "$OUT" -c "$ROOT/dgswem_amd/fortran/dg_timestep_gpu.F90" -o dg_timestep_gpu.o
objs=$(ls "$OBJ"/*.o | grep -v '/DG_timestep.o$')
$FC -o "$HERE/_ref/dgswem_gpu" $objs gpu_interface.o dg_timestep_gpu.o \
    -L"$ROOT/dgswem_amd" -ldgswem_hip -Wl,-rpath,'$ORIGIN/../../dgswem_amd' -lstdc++
echo "build_dropin: built $HERE/_ref/dgswem_gpu"
