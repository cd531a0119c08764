#!/bin/bash
# Generate, compile and sweep one stencil, with the reference's argv
# (hip/RunBenchmark.sh:24-47):
#   RunBenchmark.sh Box|Star 2D|3D range Jacobi|Variable Row|Column None|Read|Write CLsize xC yC [zC]
# Grid sizes and the time per block shape default to the reference's
# (~1 GB per grid, 1 s); override with X2D/Y2D, X3D/Y3D/Z3D, TIME.  Unlike the
# reference, every failed step stops the script with a non-zero status.
set -euo pipefail
err() { >&2 echo "$@"; exit 1; }
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
x2D=${X2D:-11500}; y2D=${Y2D:-11500}
x3D=${X3D:-500}; y3D=${Y3D:-500}; z3D=${Z3D:-500}
time=${TIME:-1.0}
if [ $# != 9 ] && [ $# != 10 ]; then
    err "Usage: neighbourhood dimension range coefficients orientation alignment CLsize xC yC [zC]"
fi
nb=$1; dim=$2; r=$3; coeff=$4; orient=$5; align=$6; cl=$7; xc=$8; yc=$9
if [ "$dim" == "2D" ] && [ $# == 9 ]; then :
elif [ "$dim" == "3D" ] && [ $# == 10 ]; then zc=${10}
else err "Dimensions and amount of arguments mismatch"; fi
[[ "$nb" == "Box" || "$nb" == "Star" ]] || err "Unknown neighbourhood"
[[ "$r" =~ ^[0-9]+$ && "$r" -ge 1 ]] || err "range has to be an integer >= 1"
[[ "$coeff" == "Jacobi" || "$coeff" == "Variable" ]] || err "Unknown coefficients"
[[ "$orient" == "Row" || "$orient" == "Column" ]] || err "orientation must be 'Column' or 'Row'"
[[ "$align" == "None" || "$align" == "Read" || "$align" == "Write" ]] || err "Unknown alignment"
[[ "$cl" =~ ^[0-9]+$ ]] || err "alignment size doesn't seem to be an integer"
if [ "$align" == "None" ] && [ "$cl" != 0 ]; then err "alignment size must be 0 if alignment is \"None\""; fi
genargs=()
[ "$align" != "None" ] && genargs=("$align" "$cl")
if [ "$dim" == "2D" ]; then
    python3 "$HERE/generate2DStencil.py" "$nb" "$orient" "$coeff" "$r" "$xc" "$yc" "${genargs[@]}" ${TCSG_GEN_OPTS:-} -o kernels.co
    TCSG_KERNEL=kernels.co "$HERE/harness/bench" "$nb" "$dim" "$r" "$coeff" "$orient" "$align" "$cl" "$time" "$xc" "$yc" "$x2D" "$y2D"
else
    python3 "$HERE/generate3DStencil.py" "$nb" "$orient" "$coeff" "$r" "$xc" "$yc" "$zc" "${genargs[@]}" ${TCSG_GEN_OPTS:-} -o kernels.co
    TCSG_KERNEL=kernels.co "$HERE/harness/bench" "$nb" "$dim" "$r" "$coeff" "$orient" "$align" "$cl" "$time" "$xc" "$yc" "$zc" "$x3D" "$y3D" "$z3D"
fi
