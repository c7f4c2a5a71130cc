#!/bin/bash
# Builds the reference Fortran solver (read from its own sources under the
# reference checkout, never copied into this repository) into oracle/_ref/:
#
#   oracle/_ref/libref_<pde>_n<N>.so   reference routines + the ref_harness_<pde>.f90
#                                      bind(C) harness (golden vectors, CPU baseline)
#   oracle/_ref/aderdg_<pde>_n<N>      the reference's own program (main.f90)
#   oracle/_ref/refmain_euler_n<N>     (mode `refmain`, after the HIP library is built) the
#                                      reference's Euler main.f90 with the maintainer's edit of
#                                      oracle/refmain_edit.py, linked with the reference routines,
#                                      ader-dg_amd/fortran/aderdg_hip.f90 + aderdg_refglue.f90 and
#                                      libaderdg_amd_exact.so: the drop-in run of the reference
#                                      program on the GPU
#   oracle/_ref/percell_euler_n<N>     (same mode) the reference's Euler main.f90 as it is, linked
#                                      with ader-dg_amd/fortran/aderdg_percell.f90 in place of the
#                                      reference's space_time_predictor, volume_integral,
#                                      surface_integral, element_update and DMP: the per-cell tier of
#                                      the drop-in (the reference's DMP symbol in limiter.o is made
#                                      weak, so the shim's DMP serves every caller)
#
# TEST INFRASTRUCTURE ONLY: the product never links or loads these files.
#
# N is a compile-time parameter of the reference (Euler/ader_dg.f90:17), so one
# build is made per degree.  The sources are copied to a scratch directory under
# /tmp, the `N = ...` parameter line of ader_dg.f90 is rewritten there, and
# everything is compiled with the reference Makefile's flags
# (Euler/Makefile:1-2: -O3 -fdefault-real-8) using LLVM flang, minus the
# -Wall/-Wextra that flang ignores.  No -march: x86-64 baseline, no FMA.
set -euo pipefail
REF=${ADERDG_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-flang}
FFLAGS="-O3 -fdefault-real-8 -fPIC"
if [ ! -d "$REF" ]; then
    echo "build_ref.sh: no reference checkout at $REF; nothing built" >&2
    exit 0
fi
mkdir -p "$OUT"
WORK=$(mktemp -d /tmp/aderdg_ref.XXXXXX)
trap 'rm -rf "$WORK"' EXIT

# prepare <work dir> <reference dir> <N> [sed expressions "file|expr" for the fixes]:
# scratch copy of the reference sources with the degree set and the fixes applied
prepare() {
    local w=$1 dir=$2 n=$3
    shift 3
    mkdir -p "$w"
    cp "$REF/$dir"/*.f90 "$w/"
    sed -i -E "s/^( *integer, parameter *:: *N *= *)[0-9]+/\1${n}/" "$w/ader_dg.f90"
    grep -q -E "^ *integer, parameter *:: *N *= *${n}( |$)" "$w/ader_dg.f90"
    for fix in "$@"; do
        local file=${fix%%|*} expr=${fix#*|}
        cp "$w/$file" "$w/$file.orig"
        sed -i -E "$expr" "$w/$file"
        if cmp -s "$w/$file" "$w/$file.orig"; then
            echo "build_ref.sh: patch did not apply: $expr" >&2
            exit 1
        fi
        rm "$w/$file.orig"
    done
}

# the reference routines of a prepared directory (everything but the module and main.f90)
routines() {
    local srcs=""
    for f in "$1"/*.f90; do
        case $(basename "$f") in ader_dg.f90|main.f90|main_gpu.f90) ;; *) srcs="$srcs $f";; esac
    done
    echo "$srcs"
}

# build_variant <pde tag> <reference dir> <N> <harness file> [fixes]
build_variant() {
    local tag=$1 dir=$2 n=$3 harness=$4
    shift 4
    local w="$WORK/${tag}_n${n}"
    prepare "$w" "$dir" "$n" "$@"
    local srcs
    srcs=$(routines "$w")
    (cd "$w" && $FC $FFLAGS -c ader_dg.f90 && \
        for f in $srcs; do $FC $FFLAGS -c "$f"; done && \
        $FC $FFLAGS -c "$HERE/$harness" && \
        $FC $FFLAGS -c main.f90)
    local objs=""
    for f in $srcs; do objs="$objs $w/$(basename "${f%.f90}").o"; done
    local hobj="$w/$(basename "${harness%.f90}").o"
    $FC $FFLAGS -shared -o "$OUT/libref_${tag}_n${n}.so" "$w/ader_dg.o" $objs "$hobj" \
        -Wl,-Bsymbolic -Wl,--exclude-libs,ALL
    $FC $FFLAGS -o "$OUT/aderdg_${tag}_n${n}" "$w/ader_dg.o" $objs "$w/main.o"
    echo "built $OUT/libref_${tag}_n${n}.so"
}

# build_refmain <N> [fixes]: the reference Euler program with the maintainer's edit, on the GPU
build_refmain() {
    local n=$1
    shift
    local w="$WORK/refmain_n${n}" pkg="$HERE/../ader-dg_amd"
    if [ ! -e "$pkg/lib/libaderdg_amd_exact.so" ]; then
        echo "build_ref.sh: $pkg/lib/libaderdg_amd_exact.so missing; build the HIP library first" >&2
        exit 1
    fi
    prepare "$w" Euler "$n" "$@"
    python3 "$HERE/refmain_edit.py" < "$w/main.f90" > "$w/main_gpu.f90"
    local srcs
    srcs=$(routines "$w")
    (cd "$w" && $FC $FFLAGS -c ader_dg.f90 && \
        for f in $srcs; do $FC $FFLAGS -c "$f"; done && \
        $FC $FFLAGS -c "$pkg/fortran/aderdg_hip.f90" && \
        $FC $FFLAGS -c "$pkg/fortran/aderdg_refglue.f90" && \
        $FC $FFLAGS -c main_gpu.f90)
    local objs=""
    for f in $srcs; do objs="$objs $w/$(basename "${f%.f90}").o"; done
    $FC $FFLAGS -o "$OUT/refmain_euler_n${n}" "$w/ader_dg.o" $objs "$w/aderdg_hip.o" "$w/aderdg_refglue.o" \
        "$w/main_gpu.o" -L"$pkg/lib" -laderdg_amd_exact -Wl,-rpath,'$ORIGIN/../../ader-dg_amd/lib'
    echo "built $OUT/refmain_euler_n${n}"
}

# build_percell <N> [fixes]: the unedited reference Euler program over the per-cell shim
build_percell() {
    local n=$1
    shift
    local w="$WORK/percell_n${n}" pkg="$HERE/../ader-dg_amd"
    prepare "$w" Euler "$n" "$@"
    local srcs=""
    for f in $(routines "$w"); do
        case $(basename "$f") in
            space_time_predictor.f90|volume_integral.f90|surface_integral.f90|element_update.f90) ;;
            *) srcs="$srcs $f";;
        esac
    done
    (cd "$w" && $FC $FFLAGS -c ader_dg.f90 && \
        for f in $srcs; do $FC $FFLAGS -c "$f"; done && \
        $FC $FFLAGS -c "$pkg/fortran/aderdg_hip.f90" && \
        $FC $FFLAGS -c "$pkg/fortran/aderdg_refglue.f90" && \
        $FC $FFLAGS -c "$pkg/fortran/aderdg_percell.f90" && \
        $FC $FFLAGS -c main.f90)
    objcopy --weaken-symbol=dmp_ "$w/limiter.o"
    local objs=""
    for f in $srcs; do objs="$objs $w/$(basename "${f%.f90}").o"; done
    $FC $FFLAGS -o "$OUT/percell_euler_n${n}" "$w/ader_dg.o" $objs "$w/aderdg_hip.o" "$w/aderdg_refglue.o" \
        "$w/aderdg_percell.o" "$w/main.o" -L"$pkg/lib" -laderdg_amd_exact \
        -Wl,-rpath,'$ORIGIN/../../ader-dg_amd/lib'
    echo "built $OUT/percell_euler_n${n}"
}

# The Euler subcell recompute applies Double-Mach-Reflection ghost states on the
# left, bottom and top boundaries unconditionally (Euler/limiter.f90:491-522).
# Guarding them with ICType == 11 (DMR) leaves DMR runs unchanged and lets the
# harness drive the other boundary modes through the same routine.  The
# troubled-cell seeding at t=0 uses the uninitialised x0 (Euler/initialize.f90:297);
# it gets the cell corner, as in the Baer-Nunziato build below.
EULER_FIXES=(
  'limiter.f90|s/^( *)if \(i \.eq\. 1\) then ! Left/\1if (i .eq. 1 .and. ICType .eq. 11) then ! Left/'
  'limiter.f90|s/^( *)if \(j \.eq\. 1\) then ! Bottom/\1if (j .eq. 1 .and. ICType .eq. 11) then ! Bottom/'
  'limiter.f90|s/^( *)if \(j \.eq\. JMAX\) then ! Top/\1if (j .eq. JMAX .and. ICType .eq. 11) then ! Top/'
  'initialize.f90|s#xGP = x0 \+ \(/ xilimbary#xGP = corner + (/ xilimbary#'
)
if [ "${1:-all}" = refmain ]; then
    for n in 2 3; do
        build_refmain "$n" "${EULER_FIXES[@]}"
        build_percell "$n" "${EULER_FIXES[@]}"
    done
    exit 0
fi
for n in 1 2 3 4; do
    build_variant euler Euler "$n" ref_harness_euler.f90 "${EULER_FIXES[@]}"
done

# Baer-Nunziato: two memory-safety fixes.  Gm is allocated with the x-face shape
# (Baer-Nunziato/initialize.f90:136) but written up to JMAX+1 in its last index
# (Baer-Nunziato/riemann_problem.f90:72), a heap overflow; it gets the y-face shape of
# Gp.  Every element the reference reads keeps its value.  The troubled-cell seeding at
# t=0 uses the uninitialised x0 (:288); it gets the cell corner.
BN_FIXES=(
  'initialize.f90|s/allocate \( Gm\(nVar, nDOF\(2\), IMAX\+1 , JMAX \) \)/allocate ( Gm(nVar, nDOF(1), IMAX , JMAX+1 ) )/'
  'initialize.f90|s#xGP = x0 \+ \(/ xilimbary#xGP = corner + (/ xilimbary#'
)
for n in 1 2 3 4; do
    build_variant bn Baer-Nunziato "$n" ref_harness_bn.f90 "${BN_FIXES[@]}"
done

# Advection-diffusion: the reference as it is (its predictor evaluates only the
# viscous flux, Advection-Diffusion/space_time_predictor.f90:78), and a second
# build whose predictor calls the fused advective + viscous flux
# (PDEFusedViscConvFlux, Advection-Diffusion/pde.f90:115-129) at that line.
AD_FIXED=(
  'space_time_predictor.f90|s/call PDEViscFlux\(lFh/call PDEFusedViscConvFlux(lFh/'
)
for n in 1 2 3 4; do
    build_variant ad Advection-Diffusion "$n" ref_harness_ad.f90
    build_variant adfix Advection-Diffusion "$n" ref_harness_ad.f90 "${AD_FIXED[@]}"
done
