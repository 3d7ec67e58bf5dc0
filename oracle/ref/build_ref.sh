#!/bin/bash
# Builds the reference CPU physics from the reference's own Fortran sources, where they lie,
# plus the test driver ref_harness.f95, into oracle/_ref/<tag>/ (never committed).
#   build_ref.sh ref150|ref300 -> grid from the reference's params_common_sn_{150x150,300x300}x90.f95
#   build_ref.sh IPxJPxKP      -> grid chosen through a params_common_sn module with those sizes
#                                 (the reference selects its grid the same way: one module of three
#                                 integer parameters, params_common_sn_*.f95)
#   build_ref.sh <tag>cfg      -> the same grid with press.f95's SOR parameters (press.f95:44-46)
#                                 set to the benchmark's: omega 1.5, pjuge 0 (no early exit: the
#                                 residual is never below 0), nmaxp 50 (unchanged).  Only those two
#                                 parameter values differ; the edited press.f95 is written to a
#                                 temporary directory outside the repository, compiled from there
#                                 and deleted.
# Flags: the reference's cpp switch set for the CPU path with an external wind profile
# (SConstruct:86-106), at -O2 with FP contraction off instead of -Ofast (SConstruct:117), so
# that each FP32 operation is one IEEE rounding.  init.f95 is built with IFBF=0 so that it
# skips the GIS file read (feedbfm.f95:33); the harness supplies the masks instead.
set -euo pipefail
TAG=${1:?usage: build_ref.sh ref150|ref300|IPxJPxKP[cfg]}
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${LES_REFERENCE_SRC:-/root/reference/src/Fortran}
FC=${FC:-/opt/rocm/lib/llvm/bin/amdflang}
OUT="$HERE/../_ref/$TAG"
mkdir -p "$OUT"
cd "$OUT"
GRIDTAG=${TAG%cfg}
CFG=0
[ "$GRIDTAG" != "$TAG" ] && CFG=1
if [ "$GRIDTAG" = ref150 ]; then
  PARAMS="$REF/params_common_sn_150x150x90.f95"
elif [ "$GRIDTAG" = ref300 ]; then
  PARAMS="$REF/params_common_sn_300x300x90.f95"
else
  IFS=x read -r IP JP KP <<< "$GRIDTAG"
  PARAMS="$OUT/params_common_sn.f95"
  cat > "$PARAMS" <<PEOF
module params_common_sn
    integer, parameter  :: ipmax = $IP;       integer, parameter  :: jpmax = $JP
    integer, parameter  :: ip = $IP;       integer, parameter  :: jp = $JP;       integer, parameter  :: kp = $KP
end module params_common_sn
PEOF
fi
FFLAGS="-cpp -O2 -ffp-contract=off -DICAL=0 -DIFBF=1 -DIANIME=0 -DIADAM=0 -DNO_IO -DNO_FILE_IO -DEXTERNAL_WIND_PROFILE -module-dir $OUT"
OBJS=()
comp() {  # file [extra flags]
  local src=$1; shift
  local obj="$OUT/$(basename "${src%.*}").o"
  "$FC" $FFLAGS "$@" -c "$src" -o "$obj"
  OBJS+=("$obj")
}
comp "$PARAMS"
# bondv1.f95 and ifdata.f95 are left out: bondv1.f95's fixed inflow levels (k=1..78, u(:,:,77))
# do not compile for kp < 77, and the harness restates ifdata.f95's call sequence itself.
for f in common_sn boundsm vel2 velFG feedbf feedbfm les bondv1_ext_wind_profile boundp \
         bondFG velnw adam; do
  comp "$REF/$f.f95"
done
if [ $CFG = 1 ]; then
  TMPD=$(mktemp -d "${TMPDIR:-/tmp}/les_ref_cfg.XXXXXX")
  trap 'rm -rf "$TMPD"' EXIT
  # press.f95:44  real, parameter  :: pjuge = 0.0001   -> 0.0
  # press.f95:46  real, parameter  :: omega = 1.       -> 1.5
  sed -e '44s/:: pjuge = 0\.0001$/:: pjuge = 0.0/' -e '46s/:: omega = 1\.$/:: omega = 1.5/' \
      "$REF/press.f95" > "$TMPD/press.f95"
  sed -n '44p;46p' "$TMPD/press.f95" | grep -q 'pjuge = 0.0$' || { echo "press.f95:44 not as expected" >&2; exit 1; }
  sed -n '46p' "$TMPD/press.f95" | grep -q 'omega = 1.5$' || { echo "press.f95:46 not as expected" >&2; exit 1; }
  [ "$(diff "$REF/press.f95" "$TMPD/press.f95" | grep -c '^[<>]')" = 4 ] || { echo "unexpected press.f95 edit" >&2; exit 1; }
  comp "$TMPD/press.f95"
else
  comp "$REF/press.f95"
fi
for f in aveflow anime; do
  comp "$REF/$f.f95"
done
comp "$REF/init.f95" -UIFBF -DIFBF=0
comp "$HERE/ref_harness.f95"
"$FC" -O2 -o "$OUT/ref_harness" "${OBJS[@]}"
echo "built $OUT/ref_harness"
