! boundary_test.f90 -- calls libmomgpu through the ISO_C_BINDING module
! mom_gpu (for90-mom2_amd/fortran/mom_gpu.f90) the way a Fortran host's
! drop-in bodies do: the domain meshes are passed as flattened
! mesh_container arrays (mom_mesh_create), then sysmat_pmchwt
! (mom_sysmat_pmchwt), srcvec (mom_srcvec_pw), solve_multi_linsys
! (mom_solve_multi_linsys) and cs_prtsrf (mom_cs_prtsrf) run on COMPLEX(dp)
! arrays owned by this program.  TEST INFRASTRUCTURE ONLY
! (tests/test_fortran_boundary_gpu.py writes the input and checks the output
! against the reference's golden vectors).
!
! Input (stream file, argument 1): ndom (int32), N (int64); per domain
! V, F, E (int64), p(3V), fnodes(3F) int32, fedges(3F) int32, fn(3F),
! fcp(3F), fs(9F), fm(9F), farea(F), efaces(2E) int32, ebn(2E) int32,
! elen(E), eparent(E) int32, avelen; omega; ri(2 ndom); nq (int32),
! qw(nq), qnodes(3 nq); theta, phi, psi [rad].  Reals are REAL(8).
! Output (stream file, argument 2): A(2N,2N), q(2N), x(2N) (COMPLEX(8)),
! ipiv(2N) int32, csca, cabs.
PROGRAM boundary_test
  USE ISO_C_BINDING
  USE mom_gpu
  IMPLICIT NONE

  TYPE dom_arrays
     INTEGER(C_INT64_T) :: V, F, E
     REAL(C_DOUBLE), ALLOCATABLE :: p(:), fn(:), fcp(:), fs(:), fm(:), farea(:), elen(:)
     INTEGER(C_INT32_T), ALLOCATABLE :: fnodes(:), fedges(:), efaces(:), ebn(:), eparent(:)
     REAL(C_DOUBLE) :: avelen
  END TYPE dom_arrays

  CHARACTER(LEN=1024) :: fin, fout
  INTEGER :: u, d
  INTEGER(C_INT32_T) :: ndom, nq
  INTEGER(C_INT64_T) :: N, N2
  TYPE(dom_arrays), ALLOCATABLE :: dm(:)
  TYPE(C_PTR) :: ctx
  TYPE(C_PTR), ALLOCATABLE, TARGET :: meshes(:)
  REAL(C_DOUBLE) :: omega, ang(3), csca, cabs
  REAL(C_DOUBLE), ALLOCATABLE :: ri(:), qw(:), qnodes(:)
  COMPLEX(C_DOUBLE_COMPLEX), ALLOCATABLE, TARGET :: A(:,:), LU(:,:), q(:,:), x(:,:)
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: ipiv(:)

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  OPEN(NEWUNIT=u, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(u) ndom, N
  ALLOCATE(dm(ndom), meshes(ndom), ri(2*ndom))
  DO d=1,ndom
     READ(u) dm(d)%V, dm(d)%F, dm(d)%E
     ALLOCATE(dm(d)%p(3*dm(d)%V), dm(d)%fnodes(3*dm(d)%F), dm(d)%fedges(3*dm(d)%F), dm(d)%fn(3*dm(d)%F), &
          dm(d)%fcp(3*dm(d)%F), dm(d)%fs(9*dm(d)%F), dm(d)%fm(9*dm(d)%F), dm(d)%farea(dm(d)%F), &
          dm(d)%efaces(2*dm(d)%E), dm(d)%ebn(2*dm(d)%E), dm(d)%elen(dm(d)%E), dm(d)%eparent(dm(d)%E))
     READ(u) dm(d)%p, dm(d)%fnodes, dm(d)%fedges, dm(d)%fn, dm(d)%fcp, dm(d)%fs, dm(d)%fm, dm(d)%farea, &
          dm(d)%efaces, dm(d)%ebn, dm(d)%elen, dm(d)%eparent, dm(d)%avelen
  END DO
  READ(u) omega, ri, nq
  ALLOCATE(qw(nq), qnodes(3*nq))
  READ(u) qw, qnodes, ang
  CLOSE(u)

  N2 = 2*N
  ALLOCATE(A(N2,N2), LU(N2,N2), q(N2,1), x(N2,1), ipiv(N2))
  CALL mom_check(mom_ctx_create(0_C_INT32_T, ctx), 'mom_ctx_create')
  DO d=1,ndom
     CALL mom_check(mom_mesh_create(ctx, dm(d)%V, dm(d)%p, dm(d)%F, dm(d)%fnodes, dm(d)%fedges, dm(d)%fn, &
          dm(d)%fcp, dm(d)%fs, dm(d)%fm, dm(d)%farea, dm(d)%E, dm(d)%efaces, dm(d)%ebn, dm(d)%elen, &
          dm(d)%eparent, dm(d)%avelen, meshes(d)), 'mom_mesh_create')
  END DO
  ! sysmat_pmchwt (sysmat.f90:18-105)
  CALL mom_check(mom_sysmat_pmchwt(ctx, ndom, meshes, ri, N, omega, nq, qw, qnodes, C_LOC(A), N2), &
       'mom_sysmat_pmchwt')
  ! srcvec (source.f90:136-247), one plane wave in domain 1's medium
  CALL mom_check(mom_srcvec_pw(ctx, meshes(1), N, omega, ri(1), ri(2), 1_C_INT64_T, ang, nq, qw, qnodes, &
       C_LOC(q), N2), 'mom_srcvec_pw')
  ! solve_multi_linsys (linalg.f90:159-180): LU in place, solution in place
  LU(:,:) = A(:,:)
  x(:,:) = q(:,:)
  CALL mom_check(mom_solve_multi_linsys(ctx, N2, C_LOC(LU), N2, 1_C_INT64_T, C_LOC(x), N2, C_LOC(ipiv)), &
       'mom_solve_multi_linsys')
  ! cs_prtsrf (cs.f90:14-104)
  CALL mom_check(mom_cs_prtsrf(ctx, meshes(1), N, omega, ri(1), ri(2), C_LOC(x), ang(1), ang(2), ang(3), &
       nq, qw, qnodes, csca, cabs), 'mom_cs_prtsrf')
  DO d=1,ndom
     CALL mom_check(mom_mesh_destroy(meshes(d)), 'mom_mesh_destroy')
  END DO
  CALL mom_check(mom_ctx_destroy(ctx), 'mom_ctx_destroy')

  OPEN(NEWUNIT=u, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) A, q, x, ipiv, csca, cabs
  CLOSE(u)
END PROGRAM boundary_test
