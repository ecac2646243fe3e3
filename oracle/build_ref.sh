#!/usr/bin/env bash
# Build oracle/_ref/librrtmg_ref.so: the reference RRTMG_SW Fortran, compiled in place from the
# reference tree, plus this repository's C-callable wrapper oracle/ref_capi.F90.
#
# TEST INFRASTRUCTURE ONLY (the compiled reference pins the C oracle and writes golden fixtures).
# Outputs go to oracle/_ref/ only; the reference's sources are read where they lie, never copied
# into the repository. Objects are kept in oracle/_ref/obj so a rebuild only recompiles what changed
# (the 64.5 k-line data file src/rrtmg_sw_k_g.f90 takes about 4 minutes at -O0).
#
# The GCM-mode sub-column generator src/mcica_subcol_gen_sw.f90 does not compile as shipped. A
# patched copy is written to a temporary directory, compiled, and deleted:
#   P1 :147            append the missing THEN to `if (decorr_len(i) .ge. 0.0_rb)`
#   P2 :585,596,607,625,636,647   undeclared `nlayers` -> `nlay`
#   P3 :583-592, :623-632 (KISS branch of icld 4/5): drop the inner `do i = 1, ncol` loop and
#      assign CDF(isubcol,:,ilev) / CDF2(isubcol,:,ilev), the order of the compiling single-column
#      generator src/mcica_subcol_gen_sw.1col.f90:593-600 and of cases 1-3 (:499-503).
# The patch is tied to the file's SHA-256; a different file stops the build.
#
# Usage: oracle/build_ref.sh [REFERENCE_ROOT]      (default /root/reference)
set -euo pipefail
REF=${1:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
OBJ="$OUT/obj"
FC=${FC:-amdflang}
FFLAGS="-O2 -ffp-contract=off -fdefault-real-8 -fPIC"
EXPECT_SHA=b73fbb68a9a36b78e890895dfd9668dc3732371478a30181522e7030a20e867b

[ -d "$REF/src" ] || { echo "build_ref: no reference tree at $REF" >&2; exit 2; }
command -v "$FC" >/dev/null || { echo "build_ref: $FC not found" >&2; exit 2; }
mkdir -p "$OBJ"

compile() { # src obj [flags]
  local src=$1 obj=$2 fl=${3:-$FFLAGS}
  if [ ! -f "$obj" ] || [ "$src" -nt "$obj" ]; then
    echo "  FC $(basename "$src")"
    (cd "$OBJ" && $FC $fl -c "$src" -o "$obj")
  fi
}

MODS="parkind parrrsw rrsw_aer rrsw_cld rrsw_con rrsw_kg16 rrsw_kg17 rrsw_kg18 rrsw_kg19 rrsw_kg20
      rrsw_kg21 rrsw_kg22 rrsw_kg23 rrsw_kg24 rrsw_kg25 rrsw_kg26 rrsw_kg27 rrsw_kg28 rrsw_kg29
      rrsw_ref rrsw_tbl rrsw_vsn rrsw_wvn"
OBJS=()
for m in $MODS; do
  compile "$REF/modules/$m.f90" "$OBJ/$m.o"
  OBJS+=("$OBJ/$m.o")
done
# data only: -O0 compiles in minutes instead of ~11 at -O1 and does not change any value
compile "$REF/src/rrtmg_sw_k_g.f90" "$OBJ/rrtmg_sw_k_g.o" "-O0 -fdefault-real-8 -fPIC"
OBJS+=("$OBJ/rrtmg_sw_k_g.o")
for f in rrtmg_sw_setcoef rrtmg_sw_init mcica_random_numbers; do
  compile "$REF/src/$f.f90" "$OBJ/$f.o"
  OBJS+=("$OBJ/$f.o")
done

# patched sub-column generator
MC="$REF/src/mcica_subcol_gen_sw.f90"
if [ ! -f "$OBJ/mcica_subcol_gen_sw.o" ] || [ "$MC" -nt "$OBJ/mcica_subcol_gen_sw.o" ]; then
  got=$(sha256sum "$MC" | cut -d' ' -f1)
  [ "$got" = "$EXPECT_SHA" ] || { echo "build_ref: $MC changed ($got); patch not applicable" >&2; exit 3; }
  TMP=$(mktemp -d)
  trap 'rm -rf "$TMP"' EXIT
  sed -e '147s/$/ then/' \
      -e '585s/nlayers/nlay/;596s/nlayers/nlay/;607s/nlayers/nlay/' \
      -e '625s/nlayers/nlay/;636s/nlayers/nlay/;647s/nlayers/nlay/' \
      -e '587s/CDF(isubcol,i,ilev)/CDF(isubcol,:,ilev)/;589s/CDF2(isubcol,i,ilev)/CDF2(isubcol,:,ilev)/' \
      -e '627s/CDF(isubcol,i,ilev)/CDF(isubcol,:,ilev)/;629s/CDF2(isubcol,i,ilev)/CDF2(isubcol,:,ilev)/' \
      -e '584d;591d;624d;631d' "$MC" > "$TMP/mcica_subcol_gen_sw.f90"
  echo "  FC mcica_subcol_gen_sw.f90 (patched P1-P3)"
  (cd "$OBJ" && $FC $FFLAGS -c "$TMP/mcica_subcol_gen_sw.f90" -o "$OBJ/mcica_subcol_gen_sw.o")
  rm -rf "$TMP"
  trap - EXIT
fi
OBJS+=("$OBJ/mcica_subcol_gen_sw.o")
for f in rrtmg_sw_cldprmc rrtmg_sw_reftra rrtmg_sw_taumol rrtmg_sw_vrtqdr rrtmg_sw_spcvmc rrtmg_sw_rad; do
  compile "$REF/src/$f.f90" "$OBJ/$f.o"
  OBJS+=("$OBJ/$f.o")
done
compile "$HERE/ref_capi.F90" "$OBJ/ref_capi.o"
OBJS+=("$OBJ/ref_capi.o")
echo "  LD librrtmg_ref.so"
$FC -shared -o "$OUT/librrtmg_ref.so.tmp" "${OBJS[@]}"
mv "$OUT/librrtmg_ref.so.tmp" "$OUT/librrtmg_ref.so"
echo "build_ref: $OUT/librrtmg_ref.so"

# Non-McICA driver: src/rrtmg_sw_rad.nomcica.f90 declares module rrtmg_sw_rad like the McICA
# driver, so it and its wrapper are compiled with their module files in obj/nomcica and linked
# into a library of their own with the shared modules, tables and routines.
NM="$OBJ/nomcica"
mkdir -p "$NM"
compile "$REF/src/rrtmg_sw_cldprop.f90" "$OBJ/rrtmg_sw_cldprop.o"
compile "$REF/src/rrtmg_sw_spcvrt.f90" "$OBJ/rrtmg_sw_spcvrt.o"
compile_nm() { # src obj
  if [ ! -f "$2" ] || [ "$1" -nt "$2" ]; then
    echo "  FC $(basename "$1") (non-McICA)"
    (cd "$NM" && $FC $FFLAGS -module-dir "$NM" -I "$NM" -I "$OBJ" -c "$1" -o "$2")
  fi
}
compile_nm "$REF/src/rrtmg_sw_rad.nomcica.f90" "$NM/rrtmg_sw_rad_nomcica.o"
compile_nm "$HERE/ref_capi_nomcica.F90" "$NM/ref_capi_nomcica.o"
NMOBJS=()
for o in "${OBJS[@]}"; do
  case "$(basename "$o")" in
    rrtmg_sw_rad.o|ref_capi.o|mcica_subcol_gen_sw.o|rrtmg_sw_spcvmc.o|rrtmg_sw_cldprmc.o) ;;
    *) NMOBJS+=("$o") ;;
  esac
done
NMOBJS+=("$OBJ/rrtmg_sw_cldprop.o" "$OBJ/rrtmg_sw_spcvrt.o" "$NM/rrtmg_sw_rad_nomcica.o" "$NM/ref_capi_nomcica.o")
echo "  LD librrtmg_ref_nomcica.so"
$FC -shared -Wl,-Bsymbolic -o "$OUT/librrtmg_ref_nomcica.so.tmp" "${NMOBJS[@]}"
mv "$OUT/librrtmg_ref_nomcica.so.tmp" "$OUT/librrtmg_ref_nomcica.so"
echo "build_ref: $OUT/librrtmg_ref_nomcica.so"
