! MODULE: mom_gpu
! ISO_C_BINDING interface of libmomgpu.so (include/mom_gpu.h) for Fortran
! hosts.  The drop-in bodies of sysmat_pmchwt, computeD/H/K, srcvec,
! solve_multi_linsys, cs_prtsrf, far_fields, rcs_ext, nlsurf_coef and
! sysmat_pmchwt_nls (sysmat.f90:18, 272, 359, 446, 112; source.f90:136;
! linalg.f90:159; cs.f90:14; ffields.f90:16, 203; nlsurf.f90:137) call
! these functions, and the mesh/sdom set-up
! (build_mesh, extract_submesh, submesh_edge_connectivity, orient_basis of
! mesh.f90) can be replaced by mom_problem_build.
!
! Conventions: COMPLEX(C_DOUBLE_COMPLEX) arrays are passed by address,
! column-major with an explicit leading dimension; mesh index arrays are
! 0-based (subtract 1 from the reference's 1-based indices); a non-zero
! status is turned into the reference's WRITE + STOP by mom_check.
MODULE mom_gpu
  USE ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(C_INT32_T), PARAMETER :: MOM_PREC_FP64 = 64, MOM_PREC_MIXED = 32

  INTERFACE
     FUNCTION mom_last_error() BIND(C, NAME='mom_last_error') RESULT(p)
       IMPORT :: C_PTR
       TYPE(C_PTR) :: p
     END FUNCTION mom_last_error

     FUNCTION mom_ctx_create(device, ctx) BIND(C, NAME='mom_ctx_create') RESULT(st)
       IMPORT :: C_INT32_T, C_INT, C_PTR
       INTEGER(C_INT32_T), VALUE :: device
       TYPE(C_PTR) :: ctx
       INTEGER(C_INT) :: st
     END FUNCTION mom_ctx_create

     FUNCTION mom_ctx_destroy(ctx) BIND(C, NAME='mom_ctx_destroy') RESULT(st)
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT) :: st
     END FUNCTION mom_ctx_destroy

     FUNCTION mom_mesh_create(ctx, nnodes, p, nfaces, fnodes, fedges, fn, fcp, fs, fm, farea, &
          nedges, efaces, ebn, elen, eparent, avelen, mesh) BIND(C, NAME='mom_mesh_create') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT64_T), VALUE :: nnodes, nfaces, nedges
       REAL(C_DOUBLE), INTENT(IN) :: p(*), fn(*), fcp(*), fs(*), fm(*), farea(*), elen(*)
       INTEGER(C_INT32_T), INTENT(IN) :: fnodes(*), fedges(*), efaces(*), ebn(*), eparent(*)
       REAL(C_DOUBLE), VALUE :: avelen
       TYPE(C_PTR) :: mesh
       INTEGER(C_INT) :: st
     END FUNCTION mom_mesh_create

     FUNCTION mom_mesh_destroy(mesh) BIND(C, NAME='mom_mesh_destroy') RESULT(st)
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: mesh
       INTEGER(C_INT) :: st
     END FUNCTION mom_mesh_destroy

     FUNCTION mom_compute_operator(ctx, mesh, op, omega, ri_re, ri_im, nq, qw, qnodes, M, ldm) &
          BIND(C, NAME='mom_compute_operator') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR, C_CHAR
       TYPE(C_PTR), VALUE :: ctx, mesh, M
       CHARACTER(KIND=C_CHAR), VALUE :: op
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im
       INTEGER(C_INT32_T), VALUE :: nq
       REAL(C_DOUBLE), INTENT(IN) :: qw(*), qnodes(*)
       INTEGER(C_INT64_T), VALUE :: ldm
       INTEGER(C_INT) :: st
     END FUNCTION mom_compute_operator

     FUNCTION mom_sysmat_pmchwt(ctx, ndom, doms, ri, nedges_total, omega, nq, qw, qnodes, A, lda) &
          BIND(C, NAME='mom_sysmat_pmchwt') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: ctx, A
       INTEGER(C_INT32_T), VALUE :: ndom, nq
       TYPE(C_PTR), INTENT(IN) :: doms(*)
       REAL(C_DOUBLE), INTENT(IN) :: ri(*), qw(*), qnodes(*)
       INTEGER(C_INT64_T), VALUE :: nedges_total, lda
       REAL(C_DOUBLE), VALUE :: omega
       INTEGER(C_INT) :: st
     END FUNCTION mom_sysmat_pmchwt

     FUNCTION mom_srcvec_pw(ctx, mesh, nedges_total, omega, ri_re, ri_im, nsrc, angles, nq, qw, qnodes, q, ldq) &
          BIND(C, NAME='mom_srcvec_pw') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: ctx, mesh, q
       INTEGER(C_INT64_T), VALUE :: nedges_total, nsrc, ldq
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im
       REAL(C_DOUBLE), INTENT(IN) :: angles(*), qw(*), qnodes(*)
       INTEGER(C_INT32_T), VALUE :: nq
       INTEGER(C_INT) :: st
     END FUNCTION mom_srcvec_pw

     FUNCTION mom_solve_multi_linsys(ctx, n, A, lda, nrhs, B, ldb, ipiv) &
          BIND(C, NAME='mom_solve_multi_linsys') RESULT(st)
       IMPORT :: C_INT, C_INT64_T, C_PTR
       TYPE(C_PTR), VALUE :: ctx, A, B, ipiv
       INTEGER(C_INT64_T), VALUE :: n, lda, nrhs, ldb
       INTEGER(C_INT) :: st
     END FUNCTION mom_solve_multi_linsys

     FUNCTION mom_system_create(ctx, nedges_total, precision, sys) BIND(C, NAME='mom_system_create') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT64_T), VALUE :: nedges_total
       INTEGER(C_INT32_T), VALUE :: precision
       TYPE(C_PTR) :: sys
       INTEGER(C_INT) :: st
     END FUNCTION mom_system_create

     FUNCTION mom_system_destroy(sys) BIND(C, NAME='mom_system_destroy') RESULT(st)
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: sys
       INTEGER(C_INT) :: st
     END FUNCTION mom_system_destroy

     FUNCTION mom_system_fill(sys, ndom, doms, ri, omega, nq, qw, qnodes) BIND(C, NAME='mom_system_fill') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: sys
       INTEGER(C_INT32_T), VALUE :: ndom, nq
       TYPE(C_PTR), INTENT(IN) :: doms(*)
       REAL(C_DOUBLE), INTENT(IN) :: ri(*), qw(*), qnodes(*)
       REAL(C_DOUBLE), VALUE :: omega
       INTEGER(C_INT) :: st
     END FUNCTION mom_system_fill

     FUNCTION mom_system_factor(sys) BIND(C, NAME='mom_system_factor') RESULT(st)
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: sys
       INTEGER(C_INT) :: st
     END FUNCTION mom_system_factor

     FUNCTION mom_system_solve(sys, nrhs, B, ldb, tol, maxit, iters, berr) BIND(C, NAME='mom_system_solve') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: sys, B
       INTEGER(C_INT64_T), VALUE :: nrhs, ldb
       REAL(C_DOUBLE), VALUE :: tol
       INTEGER(C_INT32_T), VALUE :: maxit
       INTEGER(C_INT32_T) :: iters
       REAL(C_DOUBLE) :: berr
       INTEGER(C_INT) :: st
     END FUNCTION mom_system_solve

     FUNCTION mom_cs_prtsrf(ctx, mesh, nedges_total, omega, ri_re, ri_im, x, theta, phi, psi, nq, qw, qnodes, &
          csca, cabs) BIND(C, NAME='mom_cs_prtsrf') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: ctx, mesh, x
       INTEGER(C_INT64_T), VALUE :: nedges_total
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im, theta, phi, psi
       INTEGER(C_INT32_T), VALUE :: nq
       REAL(C_DOUBLE), INTENT(IN) :: qw(*), qnodes(*)
       REAL(C_DOUBLE) :: csca, cabs
       INTEGER(C_INT) :: st
     END FUNCTION mom_cs_prtsrf

     FUNCTION mom_far_fields(ctx, mesh, nedges_total, omega, ri_re, ri_im, x, r, ndir, theta, phi, nq, qw, &
          qnodes, out) BIND(C, NAME='mom_far_fields') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: ctx, mesh, x
       INTEGER(C_INT64_T), VALUE :: nedges_total, ndir
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im, r
       REAL(C_DOUBLE), INTENT(IN) :: theta(*), phi(*), qw(*), qnodes(*)
       INTEGER(C_INT32_T), VALUE :: nq
       REAL(C_DOUBLE) :: out(*)
       INTEGER(C_INT) :: st
     END FUNCTION mom_far_fields

     FUNCTION mom_rcs_ext(ctx, mesh, nedges_total, omega, ri_re, ri_im, x, ntheta, nphi, filter, fdir, nq, qw, &
          qnodes, rcs) BIND(C, NAME='mom_rcs_ext') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: ctx, mesh, x
       INTEGER(C_INT64_T), VALUE :: nedges_total
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im
       INTEGER(C_INT32_T), VALUE :: ntheta, nphi, filter, nq
       REAL(C_DOUBLE), INTENT(IN) :: fdir(3), qw(*), qnodes(*)
       REAL(C_DOUBLE) :: rcs(*)
       INTEGER(C_INT) :: st
     END FUNCTION mom_rcs_ext

     ! Adjoint sensitivities (no counterpart in the reference)
     FUNCTION mom_system_solve_adjoint(sys, nrhs, B, ldb, tol, maxit, iters, berr) &
          BIND(C, NAME='mom_system_solve_adjoint') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: sys, B
       INTEGER(C_INT64_T), VALUE :: nrhs, ldb
       REAL(C_DOUBLE), VALUE :: tol
       INTEGER(C_INT32_T), VALUE :: maxit
       INTEGER(C_INT32_T) :: iters
       REAL(C_DOUBLE) :: berr
       INTEGER(C_INT) :: st
     END FUNCTION mom_system_solve_adjoint

     FUNCTION mom_cs_abs_gradient(ctx, mesh, nedges_total, omega, ri_re, ri_im, x, nq, qw, qnodes, g, cabs) &
          BIND(C, NAME='mom_cs_abs_gradient') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR, C_DOUBLE_COMPLEX
       TYPE(C_PTR), VALUE :: ctx, mesh, x
       INTEGER(C_INT64_T), VALUE :: nedges_total
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im
       INTEGER(C_INT32_T), VALUE :: nq
       REAL(C_DOUBLE), INTENT(IN) :: qw(*), qnodes(*)
       COMPLEX(C_DOUBLE_COMPLEX) :: g(*)
       REAL(C_DOUBLE) :: cabs
       INTEGER(C_INT) :: st
     END FUNCTION mom_cs_abs_gradient

     FUNCTION mom_pmchwt_deps_apply(ctx, mesh, nedges_total, omega, ri_re, ri_im, nq, qw, qnodes, x, y) &
          BIND(C, NAME='mom_pmchwt_deps_apply') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR, C_DOUBLE_COMPLEX
       TYPE(C_PTR), VALUE :: ctx, mesh, x
       INTEGER(C_INT64_T), VALUE :: nedges_total
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im
       INTEGER(C_INT32_T), VALUE :: nq
       REAL(C_DOUBLE), INTENT(IN) :: qw(*), qnodes(*)
       COMPLEX(C_DOUBLE_COMPLEX) :: y(*)
       INTEGER(C_INT) :: st
     END FUNCTION mom_pmchwt_deps_apply

     ! Second harmonic, surface chi2 (nlsurf.f90:137, sysmat.f90:112)
     FUNCTION mom_nlsurf_coef(ctx, mesh, nedges_total, omega, ri_re, ri_im, epsp, x, chi2, nq, qw, qnodes, &
          src_coef, src_vec) BIND(C, NAME='mom_nlsurf_coef') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR, C_DOUBLE_COMPLEX
       TYPE(C_PTR), VALUE :: ctx, mesh, x, src_coef
       INTEGER(C_INT64_T), VALUE :: nedges_total
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: epsp(*), chi2(3)
       INTEGER(C_INT32_T), VALUE :: nq
       REAL(C_DOUBLE), INTENT(IN) :: qw(*), qnodes(*)
       COMPLEX(C_DOUBLE_COMPLEX) :: src_vec(*)
       INTEGER(C_INT) :: st
     END FUNCTION mom_nlsurf_coef

     FUNCTION mom_sysmat_pmchwt_nls(ctx, ndom, doms, shri, nedges_total, omega, nq, qw, qnodes, nsrc, &
          src_coef, ldsc, src, ldsrc, A, lda) BIND(C, NAME='mom_sysmat_pmchwt_nls') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: ctx, src_coef, src, A
       INTEGER(C_INT32_T), VALUE :: ndom, nq
       TYPE(C_PTR), INTENT(IN) :: doms(*)
       REAL(C_DOUBLE), INTENT(IN) :: shri(*), qw(*), qnodes(*)
       INTEGER(C_INT64_T), VALUE :: nedges_total, nsrc, ldsc, ldsrc, lda
       REAL(C_DOUBLE), VALUE :: omega
       INTEGER(C_INT) :: st
     END FUNCTION mom_sysmat_pmchwt_nls

     FUNCTION mom_system_fill_nls(sys, ndom, doms, shri, omega, nq, qw, qnodes, nsrc, src_coef, ldsc, &
          src, ldsrc) BIND(C, NAME='mom_system_fill_nls') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: sys, src_coef, src
       INTEGER(C_INT32_T), VALUE :: ndom, nq
       TYPE(C_PTR), INTENT(IN) :: doms(*)
       REAL(C_DOUBLE), INTENT(IN) :: shri(*), qw(*), qnodes(*)
       REAL(C_DOUBLE), VALUE :: omega
       INTEGER(C_INT64_T), VALUE :: nsrc, ldsc, ldsrc
       INTEGER(C_INT) :: st
     END FUNCTION mom_system_fill_nls

     ! Problem set-up (mesh + sdom commands): dom_ptr(ndom+1) offsets into
     ! the signed surface ids dom_ids, 0-based, dom_ptr(1) = 0.
     FUNCTION mom_problem_build(nnodes, p, nfaces, fnodes, fids, scale, ndom, dom_ptr, dom_ids, prob) &
          BIND(C, NAME='mom_problem_build') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR
       INTEGER(C_INT64_T), VALUE :: nnodes, nfaces
       REAL(C_DOUBLE), INTENT(IN) :: p(*)
       INTEGER(C_INT32_T), INTENT(IN) :: fnodes(*), fids(*), dom_ptr(*), dom_ids(*)
       REAL(C_DOUBLE), VALUE :: scale
       INTEGER(C_INT32_T), VALUE :: ndom
       TYPE(C_PTR) :: prob
       INTEGER(C_INT) :: st
     END FUNCTION mom_problem_build

     ! path: TRIM(mesh_file)//C_NULL_CHAR
     FUNCTION mom_problem_load_gmsh(path, scale, ndom, dom_ptr, dom_ids, prob) &
          BIND(C, NAME='mom_problem_load_gmsh') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR, C_CHAR
       CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
       REAL(C_DOUBLE), VALUE :: scale
       INTEGER(C_INT32_T), VALUE :: ndom
       INTEGER(C_INT32_T), INTENT(IN) :: dom_ptr(*), dom_ids(*)
       TYPE(C_PTR) :: prob
       INTEGER(C_INT) :: st
     END FUNCTION mom_problem_load_gmsh

     FUNCTION mom_problem_destroy(prob) BIND(C, NAME='mom_problem_destroy') RESULT(st)
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: prob
       INTEGER(C_INT) :: st
     END FUNCTION mom_problem_destroy

     FUNCTION mom_problem_sizes(prob, dom, nedges_total, nnodes, nfaces, nedges) &
          BIND(C, NAME='mom_problem_sizes') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR
       TYPE(C_PTR), VALUE :: prob
       INTEGER(C_INT32_T), VALUE :: dom
       INTEGER(C_INT64_T) :: nedges_total, nnodes, nfaces, nedges
       INTEGER(C_INT) :: st
     END FUNCTION mom_problem_sizes

     FUNCTION mom_problem_domain(prob, dom, p, fnodes, fedges, fn, fcp, fs, fm, farea, efaces, ebn, elen, &
          eparent, avelen) BIND(C, NAME='mom_problem_domain') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE :: prob
       INTEGER(C_INT32_T), VALUE :: dom
       REAL(C_DOUBLE) :: p(*), fn(*), fcp(*), fs(*), fm(*), farea(*), elen(*), avelen
       INTEGER(C_INT32_T) :: fnodes(*), fedges(*), efaces(*), ebn(*), eparent(*)
       INTEGER(C_INT) :: st
     END FUNCTION mom_problem_domain

     FUNCTION mom_problem_topology(prob, dom, enodes, fid, fparent, nparent) &
          BIND(C, NAME='mom_problem_topology') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_PTR
       TYPE(C_PTR), VALUE :: prob
       INTEGER(C_INT32_T), VALUE :: dom
       INTEGER(C_INT32_T) :: enodes(*), fid(*), fparent(*), nparent(*)
       INTEGER(C_INT) :: st
     END FUNCTION mom_problem_topology

     FUNCTION mom_problem_mesh_create(ctx, prob, dom, mesh) BIND(C, NAME='mom_problem_mesh_create') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_PTR
       TYPE(C_PTR), VALUE :: ctx, prob
       INTEGER(C_INT32_T), VALUE :: dom
       TYPE(C_PTR) :: mesh
       INTEGER(C_INT) :: st
     END FUNCTION mom_problem_mesh_create

     ! Vertex-coordinate adjoint: dvert(3*V) += Re d/dP (lambda^T Z_mesh x), or
     ! += d/dP (sigma_abs at fixed x + Re lambda^T q) for the exterior mesh
     FUNCTION mom_pmchwt_vertex_adjoint(ctx, mesh, nedges_total, omega, ri_re, ri_im, nq, qw, qnodes, x, lam, &
          dvert) BIND(C, NAME='mom_pmchwt_vertex_adjoint') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE, C_DOUBLE_COMPLEX
       TYPE(C_PTR), VALUE :: ctx, mesh
       INTEGER(C_INT64_T), VALUE :: nedges_total
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im
       INTEGER(C_INT32_T), VALUE :: nq
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: qw, qnodes
       COMPLEX(C_DOUBLE_COMPLEX), DIMENSION(*), INTENT(IN) :: x, lam
       REAL(C_DOUBLE), DIMENSION(*) :: dvert
       INTEGER(C_INT) :: st
     END FUNCTION mom_pmchwt_vertex_adjoint

     FUNCTION mom_cs_abs_vertex_explicit(ctx, mesh, nedges_total, omega, ri_re, ri_im, x, lam, theta, phi, psi, &
          nq, qw, qnodes, dvert) BIND(C, NAME='mom_cs_abs_vertex_explicit') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE, C_DOUBLE_COMPLEX
       TYPE(C_PTR), VALUE :: ctx, mesh
       INTEGER(C_INT64_T), VALUE :: nedges_total
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im, theta, phi, psi
       INTEGER(C_INT32_T), VALUE :: nq
       COMPLEX(C_DOUBLE_COMPLEX), DIMENSION(*), INTENT(IN) :: x, lam
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: qw, qnodes
       REAL(C_DOUBLE), DIMENSION(*) :: dvert
       INTEGER(C_INT) :: st
     END FUNCTION mom_cs_abs_vertex_explicit

     ! Mirror symmetry (the `sbnd` and `symm` commands; planes 1 = mxp,
     ! 2 = myp, 3 = mzp; boundary classes 2/3/4 = xplane/yplane/zplane)
     FUNCTION mom_problem_set_lines(prob, nlines, nodes, ids) BIND(C, NAME='mom_problem_set_lines') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR
       TYPE(C_PTR), VALUE :: prob
       INTEGER(C_INT64_T), VALUE :: nlines
       INTEGER(C_INT32_T), DIMENSION(*), INTENT(IN) :: nodes, ids
       INTEGER(C_INT) :: st
     END FUNCTION mom_problem_set_lines

     FUNCTION mom_problem_classify_edges(prob, id, bnd, nclassified) &
          BIND(C, NAME='mom_problem_classify_edges') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR
       TYPE(C_PTR), VALUE :: prob
       INTEGER(C_INT32_T), VALUE :: id, bnd
       INTEGER(C_INT64_T) :: nclassified
       INTEGER(C_INT) :: st
     END FUNCTION mom_problem_classify_edges

     FUNCTION mom_problem_edge_bnd(prob, dom, bnd) BIND(C, NAME='mom_problem_edge_bnd') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_PTR
       TYPE(C_PTR), VALUE :: prob
       INTEGER(C_INT32_T), VALUE :: dom
       INTEGER(C_INT32_T), DIMENSION(*) :: bnd
       INTEGER(C_INT) :: st
     END FUNCTION mom_problem_edge_bnd

     FUNCTION mom_symm_group(ngen, planes, nga, ef, detj, jdiag, id) BIND(C, NAME='mom_symm_group') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_DOUBLE
       INTEGER(C_INT32_T), VALUE :: ngen
       INTEGER(C_INT32_T), DIMENSION(*), INTENT(IN) :: planes
       INTEGER(C_INT32_T) :: nga
       REAL(C_DOUBLE), DIMENSION(*) :: ef, detj, jdiag
       INTEGER(C_INT32_T), DIMENSION(*) :: id
       INTEGER(C_INT) :: st
     END FUNCTION mom_symm_group

     FUNCTION mom_symm_group_c(ngen, gens, nga, ef, detj, jmat, id) BIND(C, NAME='mom_symm_group_c') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_DOUBLE
       INTEGER(C_INT32_T), VALUE :: ngen
       INTEGER(C_INT32_T), DIMENSION(*), INTENT(IN) :: gens
       INTEGER(C_INT32_T) :: nga
       REAL(C_DOUBLE), DIMENSION(*) :: ef, detj, jmat
       INTEGER(C_INT32_T), DIMENSION(*) :: id
       INTEGER(C_INT) :: st
     END FUNCTION mom_symm_group_c

     FUNCTION mom_symm_edge_bc(nedges_total, bnd, ngen, planes, rep, id, dim) &
          BIND(C, NAME='mom_symm_edge_bc') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T
       INTEGER(C_INT64_T), VALUE :: nedges_total
       INTEGER(C_INT32_T), DIMENSION(*), INTENT(IN) :: bnd, planes
       INTEGER(C_INT32_T), VALUE :: ngen, rep
       INTEGER(C_INT32_T), DIMENSION(*) :: id
       INTEGER(C_INT64_T) :: dim
       INTEGER(C_INT) :: st
     END FUNCTION mom_symm_edge_bc

     FUNCTION mom_system_create_order(ctx, order, precision, sys) BIND(C, NAME='mom_system_create_order') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR
       TYPE(C_PTR), VALUE :: ctx
       INTEGER(C_INT64_T), VALUE :: order
       INTEGER(C_INT32_T), VALUE :: precision
       TYPE(C_PTR) :: sys
       INTEGER(C_INT) :: st
     END FUNCTION mom_system_create_order

     FUNCTION mom_system_fill_symm(nrep, systems, ndom, doms, ri, nedges_total, bnd, ngen, planes, omega, &
          nq, qw, qnodes) BIND(C, NAME='mom_system_fill_symm') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE
       INTEGER(C_INT32_T), VALUE :: nrep, ndom, ngen, nq
       TYPE(C_PTR), DIMENSION(*), INTENT(IN) :: systems, doms
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: ri, qw, qnodes
       INTEGER(C_INT64_T), VALUE :: nedges_total
       INTEGER(C_INT32_T), DIMENSION(*), INTENT(IN) :: bnd, planes
       REAL(C_DOUBLE), VALUE :: omega
       INTEGER(C_INT) :: st
     END FUNCTION mom_system_fill_symm

     FUNCTION mom_srcvec_pw_symm(ctx, mesh, nedges_total, omega, ri_re, ri_im, ngen, planes, nsrc, angles, &
          nq, qw, qnodes, q, ldq) BIND(C, NAME='mom_srcvec_pw_symm') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE, C_DOUBLE_COMPLEX
       TYPE(C_PTR), VALUE :: ctx, mesh
       INTEGER(C_INT64_T), VALUE :: nedges_total, nsrc, ldq
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im
       INTEGER(C_INT32_T), VALUE :: ngen, nq
       INTEGER(C_INT32_T), DIMENSION(*), INTENT(IN) :: planes
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: angles, qw, qnodes
       COMPLEX(C_DOUBLE_COMPLEX), DIMENSION(*) :: q
       INTEGER(C_INT) :: st
     END FUNCTION mom_srcvec_pw_symm

     FUNCTION mom_cs_prtsrf_symm(ctx, mesh, nedges_total, omega, ri_re, ri_im, ngen, planes, x, theta, phi, psi, &
          nq, qw, qnodes, csca, cabs) BIND(C, NAME='mom_cs_prtsrf_symm') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE, C_DOUBLE_COMPLEX
       TYPE(C_PTR), VALUE :: ctx, mesh
       INTEGER(C_INT64_T), VALUE :: nedges_total
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im, theta, phi, psi
       INTEGER(C_INT32_T), VALUE :: ngen, nq
       INTEGER(C_INT32_T), DIMENSION(*), INTENT(IN) :: planes
       COMPLEX(C_DOUBLE_COMPLEX), DIMENSION(*), INTENT(IN) :: x
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: qw, qnodes
       REAL(C_DOUBLE) :: csca, cabs
       INTEGER(C_INT) :: st
     END FUNCTION mom_cs_prtsrf_symm

     FUNCTION mom_rcs_ext_symm(ctx, mesh, nedges_total, omega, ri_re, ri_im, ngen, planes, x, ntheta, nphi, &
          filter, fdir, nq, qw, qnodes, rcs) BIND(C, NAME='mom_rcs_ext_symm') RESULT(st)
       IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE, C_DOUBLE_COMPLEX
       TYPE(C_PTR), VALUE :: ctx, mesh
       INTEGER(C_INT64_T), VALUE :: nedges_total
       REAL(C_DOUBLE), VALUE :: omega, ri_re, ri_im
       INTEGER(C_INT32_T), VALUE :: ngen, ntheta, nphi, filter, nq
       INTEGER(C_INT32_T), DIMENSION(*), INTENT(IN) :: planes
       COMPLEX(C_DOUBLE_COMPLEX), DIMENSION(*), INTENT(IN) :: x
       REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: fdir, qw, qnodes
       REAL(C_DOUBLE), DIMENSION(*) :: rcs
       INTEGER(C_INT) :: st
     END FUNCTION mom_rcs_ext_symm
  END INTERFACE

CONTAINS

  ! The reference reports failures with WRITE + STOP (linalg.f90:170-179);
  ! the drop-in keeps that behaviour on any non-zero status.
  SUBROUTINE mom_check(st, what)
    INTEGER(C_INT), INTENT(IN) :: st
    CHARACTER(LEN=*), INTENT(IN) :: what
    CHARACTER(KIND=C_CHAR), DIMENSION(:), POINTER :: msg
    INTEGER :: n
    IF(st==0) RETURN
    CALL C_F_POINTER(mom_last_error(), msg, (/512/))
    n = 0
    DO WHILE(n < 512)
       IF(msg(n+1)==C_NULL_CHAR) EXIT
       n = n + 1
    END DO
    WRITE(*,*) what, ' failed (status ', st, '): ', msg(1:n)
    STOP
  END SUBROUTINE mom_check

END MODULE mom_gpu
