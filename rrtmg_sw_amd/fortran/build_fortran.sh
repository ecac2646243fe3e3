#!/usr/bin/env bash
# Build the Fortran drop-in modules (rrtmg_sw_init, mcica_subcol_gen_sw, rrtmg_sw_rad) into
# rrtmg_sw_amd/fortran/librrtmgsw_fortran.so (+ .mod files in rrtmg_sw_amd/fortran/mod), linked
# against rrtmg_sw_amd/librrtmgsw_hip.so, and the example host program example_gcm; then the
# non-McICA driver's modules (rrtmg_sw_init, rrtmg_sw_rad) into librrtmgsw_fortran_nomcica.so
# (+ mod_nomcica) and example_gcm_nomcica.
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
PKG=$(dirname "$HERE")
FC=${FC:-amdflang}
mkdir -p "$HERE/mod"
$FC -O2 -fPIC -shared -J "$HERE/mod" -o "$HERE/librrtmgsw_fortran.so" "$HERE/rrtmgsw_lead.f90" \
    "$HERE/rrtmg_sw_hip.f90" \
    -L"$PKG" -lrrtmgsw_hip -Wl,-rpath,'$ORIGIN/..'
$FC -O2 -J "$HERE/mod" -o "$HERE/example_gcm" "$HERE/example_pad.f90" "$HERE/example_gcm.f90" \
    -L"$HERE" -lrrtmgsw_fortran -L"$PKG" -lrrtmgsw_hip -Wl,-rpath,'$ORIGIN' -Wl,-rpath,'$ORIGIN/..'
mkdir -p "$HERE/mod_nomcica"
$FC -O2 -fPIC -shared -J "$HERE/mod_nomcica" -o "$HERE/librrtmgsw_fortran_nomcica.so" \
    "$HERE/rrtmgsw_lead.f90" "$HERE/rrtmg_sw_hip_nomcica.f90" -L"$PKG" -lrrtmgsw_hip -Wl,-rpath,'$ORIGIN/..'
$FC -O2 -J "$HERE/mod_nomcica" -o "$HERE/example_gcm_nomcica" "$HERE/example_pad.f90" \
    "$HERE/example_gcm_nomcica.f90" \
    -L"$HERE" -lrrtmgsw_fortran_nomcica -L"$PKG" -lrrtmgsw_hip -Wl,-rpath,'$ORIGIN' \
    -Wl,-rpath,'$ORIGIN/..'
