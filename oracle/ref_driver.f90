! ref_driver.f90 -- golden-vector and timing driver linked against the
! reference's own modules (built by oracle/build_ref.sh).  TEST
! INFRASTRUCTURE ONLY.
!
! It performs the set-up of the `mesh`, `symm 1 id`, `sdom`, `fdom` commands
! (interface.f90:44-123, 160-210, 357-366) and then calls the reference
! routines on the hot path, dumping their outputs as raw little-endian
! stream files:
!   mesh_<d>.bin   set-up arrays of the master mesh (d=0) and each domain
!   op_<d>_<X>.bin computeD/H/K (X = D, H, K) of domain d, E x E complex
!   A.bin          sysmat_pmchwt, 2N x 2N complex (column-major)
!   q.bin          srcvec of each plane wave, 2N x nsrc complex
!   xref.bin       (do_solve) the reference's own solve_systems solution of
!                  A X = Q (ZGETRF + ZGETRS, solver.f90:325-357), 2N x nsrc
!   post.bin       per source csca, cabs (cs_prtsrf), then per source
!                  rcs(ntheta,nphi), for the solution read from x.bin or,
!                  without x.bin, for xref
!   times.txt      omp_get_wtime seconds of sysmat_pmchwt, the srcvec loop
!                  and solve_systems (solver.f90:107-140 times the same
!                  three phases)
!   sh.bin         (optional SH block, single source, needs a solution) the
!                  second-harmonic chain of solve_batch's centrosymmetric
!                  branch (solver.f90:143-231): per nonlinear domain
!                  src_coef and src_vec of nlsurf_coef, the accumulated
!                  source, the source after sysmat_pmchwt_nls, the SH
!                  solution of solve_systems and its rcs at 2 omega
!                  (interface.f90:1316-1325)
!
! Input on stdin, one item per line (or "lu" / "n nrhs" / output directory:
! only solve_multi_linsys, linalg.f90:159-180, on a seeded random complex
! n x n system, timed into times.txt):
!   mesh file / scale / ndom / per domain: "nsurf id1 id2 ..." /
!   wavelength [m] / per domain: "re im" refractive index / nsrc /
!   per source: "theta phi psi" [deg] / "dump_ops fill do_solve" with
!   fill 0 (none), 1 (sysmat_pmchwt, written to A.bin) or 2 (sysmat_pmchwt
!   timed, not written) and do_solve 0/1 (needs fill > 0) /
!   "ntheta nphi" / output directory /
!   optional: "1" / per domain "re im" SH refractive index / per domain
!   "0|1" surface chi2 / "nnn_re nnn_im ntt_re ntt_im ttn_re ttn_im"
PROGRAM ref_driver
  USE sysmat
  USE cs
  USE ffields
  USE solver
  USE omp_lib
  IMPLICIT NONE

  TYPE(batch) :: b
  CHARACTER(LEN=512) :: mshfile, outdir, ln
  ! get_mesh_element_lines (mesh.f90:578-608) declares its file name LEN=256:
  ! the name is passed blank-padded to exactly that length
  CHARACTER(LEN=256) :: mshname
  REAL(KIND=dp) :: scale, wl, omega, th, ph, ps, t0, tfill, tsrc, tsolve, re, im
  INTEGER :: ndom, d, nsurf, i, ids(16), dump_ops, dump_sys, do_solve, nth, nph, N, u, ios, nsrc, l
  COMPLEX(KIND=dp), DIMENSION(:,:,:), ALLOCATABLE :: A, xs
  COMPLEX(KIND=dp), DIMENSION(:,:), ALLOCATABLE :: q, M, x
  REAL(KIND=dp), DIMENSION(:), ALLOCATABLE :: csca, cabs
  REAL(KIND=dp), DIMENSION(:,:), ALLOCATABLE :: rcsdata
  REAL(KIND=dp), DIMENSION(:,:,:), ALLOCATABLE :: rcsall
  INTEGER, DIMENSION(:), POINTER :: vol_ids
  TYPE(prdnfo), POINTER :: prd
  LOGICAL :: have_x
  INTEGER :: do_sh, nlsflag, nind
  REAL(KIND=dp) :: chi(6)
  INTEGER, DIMENSION(:), ALLOCATABLE :: ind
  COMPLEX(KIND=dp), DIMENSION(:,:,:,:), ALLOCATABLE :: src_coef
  COMPLEX(KIND=dp), DIMENSION(:,:,:), ALLOCATABLE :: nlx, Ash
  COMPLEX(KIND=dp), DIMENSION(:), ALLOCATABLE :: src_vec, epsp, rhs0
  CHARACTER(LEN=1), DIMENSION(3) :: opn = (/'D','H','K'/)
  INTEGER :: nlu, nrhslu
  INTEGER, DIMENSION(:), ALLOCATABLE :: seed
  REAL(KIND=dp), DIMENSION(:,:), ALLOCATABLE :: rre, rim
  COMPLEX(KIND=dp), DIMENSION(:,:), ALLOCATABLE :: Alu, Blu

  READ(*,'(A)') mshfile
  IF(TRIM(mshfile)=='lu') THEN
     READ(*,*) nlu, nrhslu
     READ(*,'(A)') outdir
     CALL RANDOM_SEED(SIZE=i)
     ALLOCATE(seed(i))
     seed(:) = 20251203
     CALL RANDOM_SEED(PUT=seed)
     ALLOCATE(rre(nlu,nlu), rim(nlu,nlu), Alu(nlu,nlu), Blu(nlu,nrhslu))
     CALL RANDOM_NUMBER(rre)
     CALL RANDOM_NUMBER(rim)
     Alu(:,:) = CMPLX(rre - 0.5_dp, rim - 0.5_dp, KIND=dp)
     DO i=1,nlu
        Alu(i,i) = Alu(i,i) + REAL(nlu, KIND=dp)
     END DO
     Blu(:,:) = Alu(:,1:nrhslu)
     DEALLOCATE(rre, rim)
     t0 = omp_get_wtime()
     CALL solve_multi_linsys(Alu, Blu)
     tsolve = omp_get_wtime() - t0
     OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/times.txt', STATUS='REPLACE')
     WRITE(u,'(A,ES25.17)') 'solve_multi_linsys ', tsolve
     WRITE(u,'(A,I0)') 'threads ', omp_get_max_threads()
     CLOSE(u)
     STOP
  END IF
  READ(*,*) scale
  READ(*,*) ndom

  CALL batch_defaults(b)
  CALL group_id(b%ga)

  mshname = mshfile
  b%mesh = load_mesh(mshname)
  CALL build_mesh(b%mesh, scale)

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
  READ(*,*) nsrc
  ALLOCATE(b%src(nsrc))
  DO l=1,nsrc
     READ(*,*) th, ph, ps
     b%src(l)%type = src_pw
     b%src(l)%pos(:) = 0.0_dp
     b%src(l)%theta = th*degtorad
     b%src(l)%phi = ph*degtorad
     b%src(l)%psi = ps*degtorad
  END DO
  READ(*,*) dump_ops, dump_sys, do_solve
  READ(*,*) nth, nph
  READ(*,'(A)') outdir

  N = b%mesh%nedges
  omega = 2.0_dp*pi*c0/wl
  prd => NULL()

  CALL dump_mesh(b%mesh, TRIM(outdir)//'/mesh_0.bin')
  DO d=1,ndom
     WRITE(ln,'(A,A,I0,A)') TRIM(outdir), '/mesh_', d, '.bin'
     CALL dump_mesh(b%domains(d)%mesh, TRIM(ln))
  END DO

  IF(dump_ops==1) THEN
     DO d=1,ndom
        ALLOCATE(M(b%domains(d)%mesh%nedges, b%domains(d)%mesh%nedges))
        DO i=1,3
           IF(i==1) CALL computeD(omega, b%media(d)%prop(1)%ri, b%domains(d)%mesh, b%ga(1), prd, b%qd_tri, M)
           IF(i==2) CALL computeH(omega, b%media(d)%prop(1)%ri, b%domains(d)%mesh, b%ga(1), prd, b%qd_tri, M)
           IF(i==3) CALL computeK(omega, b%media(d)%prop(1)%ri, b%domains(d)%mesh, b%ga(1), prd, b%qd_tri, M)
           WRITE(ln,'(A,A,I0,A,A,A)') TRIM(outdir), '/op_', d, '_', opn(i), '.bin'
           OPEN(NEWUNIT=u, FILE=TRIM(ln), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
           WRITE(u) M
           CLOSE(u)
        END DO
        DEALLOCATE(M)
     END DO
  END IF

  ! the srcvec loop of solve_batch (solver.f90:123-127), one source per column
  ALLOCATE(q(2*N,nsrc))
  t0 = omp_get_wtime()
  DO l=1,nsrc
     CALL srcvec(b%domains(1)%mesh, N, omega, b%media(1)%prop(1)%ri, b%ga, b%src(l), b%qd_tri, q(:,l:l))
  END DO
  tsrc = omp_get_wtime() - t0
  OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/q.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) q
  CLOSE(u)

  tfill = -1.0_dp
  tsolve = -1.0_dp
  IF(dump_sys>=1) THEN
     ALLOCATE(A(2*N,2*N,1))
     t0 = omp_get_wtime()
     CALL sysmat_pmchwt(b, 1, A)
     tfill = omp_get_wtime() - t0
     IF(dump_sys==1) THEN
        OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/A.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
        WRITE(u) A(:,:,1)
        CLOSE(u)
     END IF
     IF(do_solve==1) THEN
        ! solve_systems (solver.f90:136, 325-357) on all sources at once
        ALLOCATE(xs(2*N,1,nsrc))
        xs(:,1,:) = q(:,:)
        t0 = omp_get_wtime()
        CALL solve_systems(b%mesh, b%ga, (1.0_dp,0.0_dp), (1.0_dp,0.0_dp), A, xs)
        tsolve = omp_get_wtime() - t0
        OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/xref.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
        WRITE(u) xs(:,1,:)
        CLOSE(u)
     END IF
     DEALLOCATE(A)
  END IF

  OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/times.txt', STATUS='REPLACE')
  WRITE(u,'(A,ES25.17)') 'sysmat_pmchwt ', tfill
  WRITE(u,'(A,ES25.17)') 'srcvec ', tsrc
  WRITE(u,'(A,ES25.17)') 'solve_systems ', tsolve
  WRITE(u,'(A,I0)') 'threads ', omp_get_max_threads()
  CLOSE(u)

  INQUIRE(FILE=TRIM(outdir)//'/x.bin', EXIST=have_x)
  IF(have_x) THEN
     ALLOCATE(x(2*N,nsrc))
     OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/x.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', IOSTAT=ios)
     READ(u) x
     CLOSE(u)
  ELSE IF(ALLOCATED(xs)) THEN
     ALLOCATE(x(2*N,nsrc))
     x(:,:) = xs(:,1,:)
     have_x = .TRUE.
  END IF
  IF(have_x .AND. nth>0) THEN
     ALLOCATE(csca(nsrc), cabs(nsrc), rcsall(nth,nph,nsrc))
     DO l=1,nsrc
        CALL cs_prtsrf(b%domains(1)%mesh, N, omega, b%media(1)%prop(1)%ri, b%ga, &
             x(:,l:l), b%src(l), b%qd_tri, csca(l), cabs(l))
        CALL rcs(b%domains(1)%mesh, N, omega, b%media(1)%prop(1)%ri, b%ga, &
             x(:,l:l), nth, nph, b%qd_tri, rcsall(:,:,l))
     END DO
     OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/post.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
     DO l=1,nsrc
        WRITE(u) csca(l), cabs(l)
     END DO
     WRITE(u) rcsall
     CLOSE(u)
  END IF

  do_sh = 0
  READ(*,*,IOSTAT=ios) do_sh
  IF(ios==0 .AND. do_sh==1 .AND. have_x) THEN
     DO d=1,ndom
        READ(*,*) re, im
        b%media(d)%prop(1)%shri = CMPLX(re, im, KIND=dp)
     END DO
     DO d=1,ndom
        READ(*,*) nlsflag
        IF(nlsflag==1) b%media(d)%type = mtype_nls
     END DO
     READ(*,*) chi
     DO d=1,ndom
        b%media(d)%prop(1)%nls%chi2_nnn = CMPLX(chi(1), chi(2), KIND=dp)
        b%media(d)%prop(1)%nls%chi2_ntt = CMPLX(chi(3), chi(4), KIND=dp)
        b%media(d)%prop(1)%nls%chi2_ttn = CMPLX(chi(5), chi(6), KIND=dp)
     END DO
     ! solve_batch, centrosymmetric branch (solver.f90:143-231), identity group
     ALLOCATE(nlx(2*N,1,1), src_coef(2*N,ndom,1,1), src_vec(2*N), epsp(b%mesh%nfaces), ind(N), rhs0(2*N))
     nlx(:,:,:) = 0.0_dp
     src_coef(:,:,:,:) = 0.0_dp
     src_vec(:) = 0.0_dp
     OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/sh.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
     DO d=1,ndom
        IF(b%media(d)%type/=mtype_nls) CYCLE
        CALL determine_epsp(b, d, 1, epsp(1:b%domains(d)%mesh%nfaces))
        nind = b%domains(d)%mesh%nedges
        ind(1:nind) = b%domains(d)%mesh%edges(:)%parent_index
        CALL nlsurf_coef(b%domains(d)%mesh, N, omega, b%media(d)%prop(1)%ri, b%media(d)%prop(1)%shri, &
             epsp(1:b%domains(d)%mesh%nfaces), x(:,1:1), b%ga, 1, b%media(d)%prop(1)%nls, &
             b%qd_tri, (1.0_dp,0.0_dp), (1.0_dp,0.0_dp), src_coef(1:(2*nind),d,1,1), src_vec(1:(2*nind)))
        WRITE(u) epsp(1:b%domains(d)%mesh%nfaces), src_coef(1:(2*nind),d,1,1), src_vec(1:(2*nind))
        nlx(ind(1:nind),1,1) = nlx(ind(1:nind),1,1) + src_vec(1:nind)
        nlx(ind(1:nind)+N,1,1) = nlx(ind(1:nind)+N,1,1) + src_vec((nind+1):(2*nind))
     END DO
     rhs0(:) = nlx(:,1,1)
     ALLOCATE(Ash(2*N,2*N,1))
     CALL sysmat_pmchwt_nls(b, 1, Ash, src_coef, nlx)
     WRITE(u) rhs0, nlx(:,1,1)
     CALL solve_systems(b%mesh, b%ga, (1.0_dp,0.0_dp), (1.0_dp,0.0_dp), Ash, nlx)
     IF(.NOT. ALLOCATED(rcsdata)) ALLOCATE(rcsdata(nth,nph))
     CALL rcs(b%domains(1)%mesh, N, 2.0_dp*omega, b%media(1)%prop(1)%shri, b%ga, nlx(:,:,1), &
          nth, nph, b%qd_tri, rcsdata)
     WRITE(u) nlx(:,1,1), rcsdata
     CLOSE(u)
  END IF

CONTAINS
  SUBROUTINE dump_mesh(mesh, fname)
    TYPE(mesh_container), INTENT(IN) :: mesh
    CHARACTER(LEN=*), INTENT(IN) :: fname
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
  END SUBROUTINE dump_mesh
END PROGRAM ref_driver
