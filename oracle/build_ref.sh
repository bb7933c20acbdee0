#!/usr/bin/env bash
# Builds the reference FMT2D executables (fmt2d, invert2d) from the Fortran
# sources under /root/reference/src into oracle/_ref/, plus the field-dump
# driver oracle/ref_fields.f90 linked against the reference's own modules.
#
# The sources are compiled where they lie, except src/rays2d.f90: it relies on
# short-circuit evaluation of `.and.` at four sites (rays2d.f90:471-472,
# 540-541, 604-605, 645-646: `associated(raysec%istart) .and.
# raysec%istart%id == ...`), which the Fortran standard does not guarantee and
# amdflang does not provide (the unpatched binary faults on the first Pg ray).
# A temporary copy of that one file is rewritten so that the id is read once,
# guarded by `associated`, before the loop; this is a no-op wherever the
# compiler short-circuits.  The copy lives in a temp dir and is deleted.
#
# Compiler: gfortran -O3 when present, otherwise amdflang -O3.
set -euo pipefail
REF=${FMT2D_REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ]; then
  echo "reference sources not found at $REF/src; skipping" >&2
  exit 0
fi
if command -v gfortran >/dev/null 2>&1; then FC=gfortran; else FC=amdflang; fi
FFLAGS="-O3 -cpp"
mkdir -p "$OUT"
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
cp "$REF/src/rays2d.f90" "$TMP/rays2d.f90"
chmod u+w "$TMP/rays2d.f90"
# declare and set the guarded id (after the last declaration / the pt allocation)
sed -i 's/^\(  real(kind=dp) *:: seed_dist, frac_seed\)$/\1\n  integer :: istart_id_sc/' "$TMP/rays2d.f90"
sed -i 's/^\(  allocate(pt(2, maxpoints))\)$/\1\n  istart_id_sc = -huge(1)\n  if (associated(raysec%istart)) istart_id_sc = raysec%istart%id/' "$TMP/rays2d.f90"
# the four two-line conditions
perl -0pi -e 's/associated\(raysec%istart\) \.and\. &\n(\s*)raysec%istart%id == (iftop|ifbot)%id\) then/.true. .and. &\n$1istart_id_sc == $2%id) then/g' "$TMP/rays2d.f90"
n=$(grep -c "istart_id_sc ==" "$TMP/rays2d.f90")
if [ "$n" != 4 ]; then echo "patch applied at $n sites, expected 4" >&2; exit 1; fi
cd "$TMP"
S="$REF/src"
$FC $FFLAGS -c "$S/mod_fmt2d.f90"
$FC $FFLAGS -c "$S/propagate2d.f90"
$FC $FFLAGS -c rays2d.f90
$FC $FFLAGS -c "$S/read_tomo2d.f90"
$FC $FFLAGS -c "$S/frechet2d.f90"
$FC $FFLAGS -c "$S/inversion2d.f90"
$FC $FFLAGS -c "$S/fmt2d_main.f90"
$FC $FFLAGS -c "$S/invert2d_main.f90"
$FC $FFLAGS -o "$OUT/fmt2d" mod_fmt2d.o propagate2d.o rays2d.o read_tomo2d.o fmt2d_main.o
$FC $FFLAGS -o "$OUT/invert2d" mod_fmt2d.o propagate2d.o rays2d.o read_tomo2d.o frechet2d.o inversion2d.o invert2d_main.o
# checker build with the ray / Frechet-row dump (oracle/raydump.inc): a temp
# copy of inversion2d.f90 includes the dump routines before
# compute_frechet_for_ray2d and calls them after every ray of run_forward2d,
# after every stage field of run_forward2d, at the end of run_lsqr and before
# the iteration's update is accumulated
cp "$S/inversion2d.f90" "$TMP/inversion2d_dump.f90"
chmod u+w "$TMP/inversion2d_dump.f90"
cp "$HERE/raydump.inc" "$TMP/raydump.inc"
perl -0pi -e 's/\n(subroutine compute_frechet_for_ray2d\(ray\))/\ninclude "raydump.inc"\n\n$1/' "$TMP/inversion2d_dump.f90"
perl -0pi -e 's/(          t_fr_acc = t_fr_acc \+ real\(t_c3 - t_c2, kind=dp\) \/ real\(t_rate, kind=dp\)\n        end if\n)/$1        call fmt2d_dump_ray(isrc, ircv, receiver(ircv)%ray(iray))\n/' "$TMP/inversion2d_dump.f90"
perl -0pi -e 's/\n(end subroutine run_lsqr)/\n  call fmt2d_dump_lsqr(ndata, nparm, nnz, n_vel_parms, smooth_type_in, max_lsqr_iter, lsqr_iter, \&\n       wdv, wdv_iface, w_smooth, w_smooth_iface, eps_min, rnorm_out, bnorm_out, test2_out, \&\n       row_ptr, gindx, gval, residual, gtg_diag, dm_sum, dm_out, mvscale, corr_x_in, corr_z_in)\n$1/' "$TMP/inversion2d_dump.f90"
perl -0pi -e 's/\n(      dmodel_total_sum\(1:nparm\) = dmodel_total_sum\(1:nparm\) \+ dmodel_total\(1:nparm\)\n)/\n      call fmt2d_dump_step(wd_best, wd_best_iface, avg_dmv, avg_dmd, step_len, rms_orig, rms_new, \&\n           ls_prev_idx, nparm, dmodel_total)\n$1/' "$TMP/inversion2d_dump.f90"
perl -0pi -e 's/(        if \(associated\(tf%time_gradient\)\) then\n          tf%time_gradient\(1:2,1:reg%nnode\) = reg%time_gradient\(1:2,1:reg%nnode\)\n        end if\n)(\n        if \(ipath == 3 \.and\. itf_idx == 2\) then)/$1        call fmt2d_dump_field(isrc, itf, reg)\n$2/' "$TMP/inversion2d_dump.f90"
nf=$(grep -c "call fmt2d_dump_field" "$TMP/inversion2d_dump.f90")
if [ "$nf" != 1 ]; then echo "field dump patch applied at $nf sites, expected 1" >&2; exit 1; fi
ns=$(grep -c "call fmt2d_dump_step" "$TMP/inversion2d_dump.f90")
if [ "$ns" != 1 ]; then echo "step dump patch applied at $ns sites, expected 1" >&2; exit 1; fi
nl=$(grep -c "call fmt2d_dump_lsqr" "$TMP/inversion2d_dump.f90")
if [ "$nl" != 1 ]; then echo "lsqr dump patch applied at $nl sites, expected 1" >&2; exit 1; fi
nd=$(grep -c "call fmt2d_dump_ray" "$TMP/inversion2d_dump.f90")
ni=$(grep -c 'include "raydump.inc"' "$TMP/inversion2d_dump.f90")
if [ "$nd" != 1 ] || [ "$ni" != 1 ]; then echo "ray dump patch applied at $nd/$ni sites, expected 1/1" >&2; exit 1; fi
mkdir -p "$TMP/dump" && cd "$TMP/dump"
$FC $FFLAGS -I.. -c "$TMP/inversion2d_dump.f90" -o inversion2d.o
$FC $FFLAGS -I.. -c "$S/invert2d_main.f90" -o invert2d_main.o
$FC $FFLAGS -o "$OUT/invert2d_dump" ../mod_fmt2d.o ../propagate2d.o ../rays2d.o ../read_tomo2d.o ../frechet2d.o inversion2d.o invert2d_main.o
cd "$TMP"
if [ -f "$HERE/ref_fields.f90" ]; then
  $FC $FFLAGS -c "$HERE/ref_fields.f90"
  $FC $FFLAGS -o "$OUT/ref_fields" mod_fmt2d.o propagate2d.o rays2d.o read_tomo2d.o ref_fields.o
fi
echo "$FC $FFLAGS" > "$OUT/COMPILER"
echo "built reference into $OUT with $FC"
