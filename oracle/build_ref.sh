#!/usr/bin/env bash
# Builds oracle/_ref/mom_ref and oracle/_ref/mom_ref_symm: the reference
# Fortran (jmakitalo/for90-mom2, read from /root/reference) linked with our
# own drivers ref_driver.f90 and ref_symm.f90, which dump golden vectors
# (mesh numbering, computeD/H/K, sysmat_pmchwt, srcvec, cs_prtsrf, rcs; the
# latter for a mirror-symmetric problem with every representation) at full
# precision.
#
# TEST INFRASTRUCTURE ONLY.  The reference source never enters the
# repository: this script compiles it from a scratch copy outside the repo
# and deletes the copy; only the linked binary lands in oracle/_ref/ (which
# .gitignore lists).
#
# The reference is written for ifort; the image's only Fortran compiler is
# AMD flang, which rejects two ifort extensions.  The scratch copy gets these
# mechanical dialect fixes, nothing else:
#   1. LOGICAL compared with ==, /= against .TRUE./.FALSE. -> .EQV./.NEQV.
#   2. rank-1 array constructors initialising rank-2 PARAMETERs get RESHAPE
#      (symmetry.f90:16-18 id33r, mesh.f90:976 cmp_triplets indices).
# LAPACK (a third-party dependency of the reference, MKL in its own build)
# is taken from the one LAPACK the image has: the OpenBLAS that scipy ships
# (scipy.libs/libscipy_openblas-*.so), the library the oracle's scipy solves
# also use.  Its entry points carry a "scipy_" prefix, so the reference's
# six LAPACK/BLAS names (zgetrf zgetrs zgetri zgeev zggev zgemm) are bound to
# them by renaming the references in the objects compiled here
# (objcopy --redefine-sym); no code stands in for them.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
FC=${FC:-/opt/rocm/llvm/bin/flang}
if [ ! -d "$REF" ]; then echo "reference not present: $REF" >&2; exit 2; fi
mkdir -p "$OUT"
WORK=$(mktemp -d /tmp/momref.XXXXXX)
trap 'rm -rf "$WORK"' EXIT

MODS="constants linalg aux mesh quad rwgf green greenprd symmetry int srcint nfields dipole bessel source bc nlsurf nlbulk common sysmat ffields cs time solver"
# Fix 1 wraps the primary left of the operator in parentheses, so that the
# operator keeps the precedence of == (a relational operator binds tighter
# than .AND., .EQV. does not): `A .AND. P(x)==.FALSE.` becomes
# `A .AND. (P(x) .EQV. .FALSE.)`.
for m in $MODS; do
  python3 - "$REF/$m.f90" "$WORK/$m.f90" <<'PYEOF'
import re, sys
src = open(sys.argv[1]).read()
pat = re.compile(r'(==|/=)\s*\.(TRUE|FALSE)\.', re.I)
out, pos = [], 0
for mt in pat.finditer(src):
    i = mt.start()
    j = i
    while j > 0 and src[j - 1] == ' ':
        j -= 1
    depth = 0
    while j > 0:
        c = src[j - 1]
        if c == ')':
            depth += 1
        elif c == '(':
            if depth == 0:
                break
            depth -= 1
        elif depth == 0 and not (c.isalnum() or c in '_%'):
            break
        j -= 1
    op = '.EQV.' if mt.group(1) == '==' else '.NEQV.'
    out.append(src[pos:j])
    out.append('(' + src[j:i].rstrip() + ' ' + op + ' .' + mt.group(2).upper() + '.)')
    pos = mt.end()
out.append(src[pos:])
open(sys.argv[2], 'w').write(''.join(out))
PYEOF
done
# rank-2 PARAMETER constructors
sed -i -E '16s/id33r = \(\//id33r = RESHAPE((\//; 18s/1\.0_dp\/\)/1.0_dp\/),(\/3,3\/))/' "$WORK/symmetry.f90"
sed -i -E '976s/indices = \((\/.*\/)\)/indices = RESHAPE((\1),(\/3,6\/))/' "$WORK/mesh.f90"

FFLAGS="-cpp -O2 -fopenmp"
cd "$WORK"
OBJS=""
for m in $MODS; do
  $FC $FFLAGS -c "$m.f90" -o "$m.o"
  OBJS="$OBJS $m.o"
done
$FC $FFLAGS -c "$HERE/ref_driver.f90" -o ref_driver.o
BLAS_SO=$(python3 -c "import glob, os, scipy; print(sorted(glob.glob(os.path.join(os.path.dirname(scipy.__file__) + '.libs', 'libscipy_openblas-*.so')))[0])")
BLAS_DIR=$(dirname "$BLAS_SO")
RENAME=""
for f in zgetrf zgetrs zgetri zgeev zggev zgemm; do
  RENAME="$RENAME --redefine-sym ${f}_=scipy_${f}_"
done
for o in $OBJS; do
  objcopy $RENAME "$o"
done
$FC $FFLAGS -c "$HERE/ref_symm.f90" -o ref_symm.o
for prog in ref_driver:mom_ref ref_symm:mom_ref_symm; do
  $FC -fopenmp -o "$OUT/${prog#*:}" "${prog%%:*}.o" $OBJS -L"$BLAS_DIR" -l:"$(basename "$BLAS_SO")" \
    -Wl,-rpath,"$BLAS_DIR" -Wl,-z,execstack
  echo "built $OUT/${prog#*:}"
done
