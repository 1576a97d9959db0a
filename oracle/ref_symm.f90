! ref_symm.f90 -- golden-vector driver for a symmetric problem (mirror planes
! and rotations about z),
! linked against the reference's own modules (built by oracle/build_ref.sh).
! TEST INFRASTRUCTURE ONLY.
!
! It performs the set-up of the `mesh`, `sbnd`, `symm`, `sdom`, `fdom`
! commands (interface.f90:44-123, 120-157, 160-210, 357-366) and the linear
! branch of solve_batch (solver.f90:100-142): sysmat_pmchwt for every
! representation, srcvec of one plane wave, solve_systems, cs_prtsrf and
! rcs.  Outputs (raw little-endian stream files):
!   mesh_<d>.bin  set-up arrays of the master mesh (d=0) and each domain
!   bnd_<d>.bin   the edges' boundary classes (mesh_bnd_*)
!   A.bin         sysmat_pmchwt, 2N x 2N x nga complex (before reduction)
!   q.bin         srcvec, 2N x nga complex
!   x.bin         the expanded solutions of solve_systems, 2N x nga complex
!   post.bin      csca, cabs (cs_prtsrf) and rcs(ntheta,nphi)
!
! Input on stdin, one item per line:
!   mesh file / scale / "ngen name1 name2 .." (mxp, myp, mzp, rN) /
!   nbnd / per boundary: "id name" (xplane, yplane, zplane) / ndom /
!   per domain: "nsurf id1 id2 ..." / wavelength [m] / per domain: "re im"
!   refractive index / "theta phi psi" [deg] / "ntheta nphi" / output dir
PROGRAM ref_symm
  USE sysmat
  USE cs
  USE ffields
  USE solver
  IMPLICIT NONE

  TYPE(batch) :: b
  CHARACTER(LEN=512) :: mshfile, outdir, ln
  ! get_mesh_element_lines (mesh.f90:578-608) declares its file name LEN=256:
  ! the name is passed blank-padded to exactly that length
  CHARACTER(LEN=256) :: mshname
  CHARACTER(LEN=8) :: names(3), bname
  REAL(KIND=dp) :: scale, wl, omega, th, ph, ps, csca, cabs, re, im
  INTEGER :: ndom, d, nsurf, i, ids(16), nth, nph, N, u, ngen, nbnd, bid, bcls, nga, nr
  TYPE(srcdata) :: src
  TYPE(group_action), DIMENSION(:), ALLOCATABLE :: ga
  COMPLEX(KIND=dp), DIMENSION(:,:,:), ALLOCATABLE :: A, x
  COMPLEX(KIND=dp), DIMENSION(:,:), ALLOCATABLE :: q
  REAL(KIND=dp), DIMENSION(:,:), ALLOCATABLE :: rcsdata
  INTEGER, DIMENSION(:), POINTER :: vol_ids

  READ(*,'(A)') mshfile
  READ(*,*) scale
  READ(*,'(A)') ln
  READ(ln,*) ngen
  READ(ln,*) ngen, names(1:ngen)
  READ(*,*) nbnd

  CALL batch_defaults(b)
  mshname = mshfile
  b%mesh = load_mesh(mshname)
  CALL build_mesh(b%mesh, scale)

  DO i=1,nbnd
     READ(*,*) bid, bname
     IF(TRIM(bname)=='xplane') bcls = mesh_bnd_xplane
     IF(TRIM(bname)=='yplane') bcls = mesh_bnd_yplane
     IF(TRIM(bname)=='zplane') bcls = mesh_bnd_zplane
     CALL classify_edges(b%mesh, bid, bcls)
  END DO

  CALL group_id(b%ga)
  DO i=1,ngen
     IF(TRIM(names(i))=='mxp') CALL group_mp(1, ga)
     IF(TRIM(names(i))=='myp') CALL group_mp(2, ga)
     IF(TRIM(names(i))=='mzp') CALL group_mp(3, ga)
     IF(names(i)(1:1)=='r') THEN
        READ(names(i)(2:3),*) nr
        CALL group_rz(nr, ga)
     END IF
     CALL product_group(b%ga, ga)
     DO d=1,SIZE(ga)
        DEALLOCATE(ga(d)%ef)
     END DO
     DEALLOCATE(ga)
  END DO
  nga = SIZE(b%ga)

  READ(*,*) ndom
  ALLOCATE(b%domains(ndom))
  vol_ids => NULL()
  DO d=1,ndom
     READ(*,'(A)') ln
     READ(ln,*) nsurf
     READ(ln,*) nsurf, ids(1:nsurf)
     b%domains(d)%medium_index = d
     b%domains(d)%gf_index = -1
     b%domains(d)%mesh = extract_submesh(b%mesh, ABS(ids(1:nsurf)), vol_ids)
     DO i=1,nsurf
        IF(ids(i)<0) CALL invert_faces(b%domains(d)%mesh, ABS(ids(i)))
     END DO
     CALL build_mesh(b%domains(d)%mesh, 1.0_dp)
  END DO
  CALL determine_edge_couples(b%mesh, 1D-12)
  CALL submesh_edge_connectivity(b%mesh, b%domains(:)%mesh)
  CALL orient_basis(b%mesh, b%domains(:)%mesh)

  READ(*,*) wl
  b%nwl = 1
  ALLOCATE(b%sols(1))
  b%sols(1)%wl = wl
  ALLOCATE(b%media(ndom))
  DO d=1,ndom
     READ(*,*) re, im
     b%media(d)%type = mtype_linear
     ALLOCATE(b%media(d)%prop(1))
     b%media(d)%prop(1)%ri = CMPLX(re, im, KIND=dp)
  END DO
  READ(*,*) th, ph, ps
  src%type = src_pw
  src%pos(:) = 0.0_dp
  src%theta = th*degtorad
  src%phi = ph*degtorad
  src%psi = ps*degtorad
  READ(*,*) nth, nph
  READ(*,'(A)') outdir

  N = b%mesh%nedges
  omega = 2.0_dp*pi*c0/wl

  CALL dump_mesh(b%mesh, TRIM(outdir)//'/mesh_0.bin', TRIM(outdir)//'/bnd_0.bin')
  DO d=1,ndom
     WRITE(ln,'(A,A,I0,A)') TRIM(outdir), '/mesh_', d, '.bin'
     WRITE(mshfile,'(A,A,I0,A)') TRIM(outdir), '/bnd_', d, '.bin'
     CALL dump_mesh(b%domains(d)%mesh, TRIM(ln), TRIM(mshfile))
  END DO

  ALLOCATE(A(2*N,2*N,nga), q(2*N,nga), x(2*N,nga,1))
  CALL sysmat_pmchwt(b, 1, A)
  OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/A.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) A
  CLOSE(u)

  CALL srcvec(b%domains(1)%mesh, N, omega, b%media(1)%prop(1)%ri, b%ga, src, b%qd_tri, q)
  OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/q.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) q
  CLOSE(u)

  x(:,:,1) = q
  CALL solve_systems(b%mesh, b%ga, (1.0_dp,0.0_dp), (1.0_dp,0.0_dp), A, x)
  OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/x.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) x(:,:,1)
  CLOSE(u)

  CALL cs_prtsrf(b%domains(1)%mesh, N, omega, b%media(1)%prop(1)%ri, b%ga, x(:,:,1), src, b%qd_tri, &
       csca, cabs)
  ALLOCATE(rcsdata(nth,nph))
  CALL rcs(b%domains(1)%mesh, N, omega, b%media(1)%prop(1)%ri, b%ga, x(:,:,1), nth, nph, b%qd_tri, rcsdata)
  OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/post.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) csca, cabs, rcsdata
  CLOSE(u)

CONTAINS
  SUBROUTINE dump_mesh(mesh, fname, bname)
    TYPE(mesh_container), INTENT(IN) :: mesh
    CHARACTER(LEN=*), INTENT(IN) :: fname, bname
    INTEGER :: uu, k
    OPEN(NEWUNIT=uu, FILE=fname, ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(uu) mesh%nnodes, mesh%nfaces, mesh%nedges
    WRITE(uu) (mesh%faces(k)%node_indices, k=1,mesh%nfaces)
    WRITE(uu) (mesh%faces(k)%edge_indices, k=1,mesh%nfaces)
    WRITE(uu) (mesh%edges(k)%face_indices, k=1,mesh%nedges)
    WRITE(uu) (mesh%edges(k)%bnode_indices, k=1,mesh%nedges)
    WRITE(uu) (mesh%edges(k)%node_indices, k=1,mesh%nedges)
    WRITE(uu) (mesh%edges(k)%parent_index, k=1,mesh%nedges)
    WRITE(uu) (mesh%nodes(k)%p, k=1,mesh%nnodes)
    WRITE(uu) (mesh%faces(k)%n, k=1,mesh%nfaces)
    WRITE(uu) (mesh%faces(k)%cp, k=1,mesh%nfaces)
    WRITE(uu) (mesh%faces(k)%s, k=1,mesh%nfaces)
    WRITE(uu) (mesh%faces(k)%m, k=1,mesh%nfaces)
    WRITE(uu) (mesh%faces(k)%area, k=1,mesh%nfaces)
    WRITE(uu) (mesh%edges(k)%length, k=1,mesh%nedges)
    WRITE(uu) mesh%avelen
    CLOSE(uu)
    OPEN(NEWUNIT=uu, FILE=bname, ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(uu) (mesh%edges(k)%bnd, k=1,mesh%nedges)
    CLOSE(uu)
  END SUBROUTINE dump_mesh
END PROGRAM ref_symm
