/*
 * mom_gpu.h -- C-ABI of the MI355X PMCHWT hot path (libmomgpu.so).
 *
 * Drop-in boundary for the for90-mom2 routines that fill the dense PMCHWT
 * system, build the plane-wave right-hand sides and factor/solve the
 * system.  Every entry point is a plain C function over pointers and
 * 64-bit sizes; a Fortran host binds them with ISO_C_BINDING.
 *
 * Conventions (those of the reference):
 *  - complex numbers are interleaved (re, im) doubles, i.e. Fortran
 *    COMPLEX(KIND=dp) / C double _Complex; matrices are column-major with
 *    a leading dimension;
 *  - index arrays in mom_mesh_create are 0-based (the Fortran shim
 *    subtracts 1), -1 marks "none";
 *  - all calls are synchronous at return; work runs on the context's HIP
 *    stream on one GPU.
 * Status: 0 = success, > 0 = LAPACK-style INFO (a zero pivot in column
 * INFO, 1-based), < 0 = one of the MOM_ERR_* codes below.  The message of
 * the last failure on the calling thread is returned by mom_last_error().
 */
#ifndef MOM_GPU_H
#define MOM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOM_OK 0
#define MOM_ERR_ARG (-1)         /* invalid argument */
#define MOM_ERR_UNSUPPORTED (-2) /* request outside the implemented path */
#define MOM_ERR_HIP (-3)         /* HIP runtime failure */
#define MOM_ERR_NOMEM (-4)       /* device allocation failed */
#define MOM_ERR_NOCONV (-5)      /* iterative refinement did not converge */
#define MOM_ERR_STATE (-6)       /* call out of order (e.g. solve before factor) */

typedef struct mom_ctx mom_ctx;
typedef struct mom_mesh mom_mesh;
typedef struct mom_system mom_system;

const char *mom_last_error(void);

/* One context per host thread and device; owns the HIP stream. */
int mom_ctx_create(int32_t device, mom_ctx **ctx);
int mom_ctx_destroy(mom_ctx *ctx);
/* The fill's tile choice in this context: the pair kernel's source faces per
 * wave (2, 4 or 8), picked on the first fill with the 13-point rule by timing
 * each candidate on the first slab of source faces (0 before that fill), and
 * the measured times [ms] of the candidates 2, 4, 8 (ms may be NULL).  The
 * choice changes no result: each face pair's arithmetic is the same. */
int mom_ctx_fill_tuning(const mom_ctx *ctx, int32_t *spw, double *ms);
/* Event timing of the fill kernels of this context (no counterpart in the
 * reference; for measurement).  out[6] (may be NULL) receives the totals
 * accumulated since timing was last switched: pair-kernel ms, pair-kernel
 * launches, point pairs evaluated by them (test x source quadrature points),
 * assembly-kernel ms, assembly launches, near-part ms.  enable = 1 switches
 * timing on, 0 off (both reset the totals after reading them), -1 only
 * reads.  Timing adds one event pair per launch. */
int mom_ctx_fill_profile(mom_ctx *ctx, int32_t enable, double *out);

/*
 * Upload one domain mesh (a flattened mesh_container, mesh.f90:75-83, as
 * built by build_mesh, extract_submesh and orient_basis):
 *   p[3*nnodes]          node coordinates (x,y,z per node) [m]
 *   fnodes[3*nfaces]     face node indices
 *   fedges[3*nfaces]     face edge indices (local edge numbering)
 *   fn, fcp[3*nfaces]    unit normal, centroid
 *   fs, fm[9*nfaces]     s(:,e), m(:,e) of each face side e, e = 0..2
 *   farea[nfaces]
 *   efaces, ebn[2*nedges] "+"/"-" face and free vertex of each edge
 *   elen[nedges]          edge length
 *   eparent[nedges]       global RWG index (parent_index) of each edge
 *   avelen                average edge length (near-pair threshold 3*avelen)
 * Geometry is taken as given, so the kernels see the host's exact bits.
 */
int mom_mesh_create(mom_ctx *ctx, int64_t nnodes, const double *p, int64_t nfaces,
                    const int32_t *fnodes, const int32_t *fedges, const double *fn,
                    const double *fcp, const double *fs, const double *fm, const double *farea,
                    int64_t nedges, const int32_t *efaces, const int32_t *ebn, const double *elen,
                    const int32_t *eparent, double avelen, mom_mesh **mesh);
int mom_mesh_destroy(mom_mesh *mesh);
/* Number of near (singular-subtraction) face pairs of the mesh. */
int64_t mom_mesh_near_pairs(const mom_mesh *mesh);
/* Classify the near face pairs of `mesh` by the centroids and average edge
 * length of `from` (same context, same face count) instead of its own: its
 * near list and the fill's near/far test (singular subtraction or plain
 * quadrature per pair) then follow `from`.  No counterpart in the reference,
 * which classifies every mesh by its own geometry (sysmat.f90:233-260); for
 * central differences of a perturbed mesh against the unperturbed one with
 * the classification frozen.  `from` may be `mesh` itself (a no-op). */
int mom_mesh_classify_like(mom_mesh *mesh, const mom_mesh *from);

/*
 * computeD / computeH / computeK (sysmat.f90:272-347, 359-434, 446-519),
 * identity group, free-space Green's function:
 *   op = 'D', 'H' or 'K'; k = ri*omega/c0; triangle rule of nq points with
 *   weights qw[nq] and barycentric nodes qnodes[nq*3] (column-major, as
 *   quad_data%nodes); M[E x E] (ldm >= E) is overwritten.
 */
int mom_compute_operator(mom_ctx *ctx, const mom_mesh *mesh, char op, double omega, double ri_re,
                         double ri_im, int32_t nq, const double *qw, const double *qnodes,
                         double *M, int64_t ldm);

/*
 * sysmat_pmchwt (sysmat.f90:18-105), identity group, free space:
 * A[2N x 2N] (lda >= 2N, N = nedges_total) is overwritten with the PMCHWT
 * matrix of the ndom domains; ri[2*ndom] holds each domain's refractive
 * index (re, im).  Domain meshes map their edges to global indices through
 * eparent.
 */
int mom_sysmat_pmchwt(mom_ctx *ctx, int32_t ndom, const mom_mesh *const *doms, const double *ri,
                      int64_t nedges_total, double omega, int32_t nq, const double *qw,
                      const double *qnodes, double *A, int64_t lda);

/*
 * srcvec (source.f90:136-247) for nsrc plane waves (type 'pw'), identity
 * group: angles[3*nsrc] = (theta, phi, psi) in radians per source;
 * q[2N x nsrc] (ldq >= 2N) is overwritten: rows 0..N-1 <f, E_inc>,
 * rows N..2N-1 <f, H_inc>.
 */
int mom_srcvec_pw(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega,
                  double ri_re, double ri_im, int64_t nsrc, const double *angles, int32_t nq,
                  const double *qw, const double *qnodes, double *q, int64_t ldq);

/*
 * solve_multi_linsys (linalg.f90:159-180): LU with partial pivoting of
 * A[n x n] (overwritten by L\U in LAPACK ZGETRF layout, ipiv[n] 1-based row
 * interchanges, may be NULL) and solve of B[n x nrhs] (overwritten by X).
 * Runs in complex128.  Returns INFO > 0 for an exactly zero pivot.
 */
int mom_solve_multi_linsys(mom_ctx *ctx, int64_t n, double *A, int64_t lda, int64_t nrhs,
                           double *B, int64_t ldb, int32_t *ipiv);

/*
 * Device-resident PMCHWT system: the matrix stays in HBM between fill,
 * factorisation and solves (solver.f90:109-136 without host round trips).
 *   precision MOM_PREC_FP64:  complex128 matrix, complex128 LU.
 *   precision MOM_PREC_MIXED: complex128 matrix kept for the residual, an
 *     equilibrated complex64 copy factorised on the FP32 matrix cores, and
 *     GMRES-based iterative refinement to a complex128 backward error.
 */
#define MOM_PREC_FP64 64
#define MOM_PREC_MIXED 32

int mom_system_create(mom_ctx *ctx, int64_t nedges_total, int32_t precision, mom_system **sys);
int mom_system_destroy(mom_system *sys);
/* Fill (same arguments as mom_sysmat_pmchwt, result kept on the device). */
int mom_system_fill(mom_system *sys, int32_t ndom, const mom_mesh *const *doms, const double *ri,
                    double omega, int32_t nq, const double *qw, const double *qnodes);
/* Copy the complex128 system matrix to the host (before factorisation). */
int mom_system_get_matrix(mom_system *sys, double *A, int64_t lda);
/* Upload a complex128 matrix instead of filling it. */
int mom_system_set_matrix(mom_system *sys, const double *A, int64_t lda);
/* Factorise (returns INFO > 0 on a zero pivot). */
int mom_system_factor(mom_system *sys);
/*
 * Solve for B[2N x nrhs] (host, overwritten by X).  In MOM_PREC_MIXED, the
 * refinement stops at backward error ||b - Zx|| / (||Z|| ||x|| + ||b||)
 * (infinity norms) <= tol or after maxit refinement steps; iters and berr
 * (may be NULL) return the steps taken and the final backward error of the
 * worst column.  MOM_ERR_NOCONV if tol is not met.  From 4 right-hand sides
 * on, the triangular solves and the refinement run on the whole block
 * (matrix products per 256-row block, one residual product per step);
 * columns whose refinement stalls are finished one by one by GMRES, or,
 * when that is estimated to cost more than an FP64 factorisation, the
 * resident complex128 matrix is factored and every column solved directly:
 * the system then keeps FP64 factors until the next fill (berr = -1).
 */
int mom_system_solve(mom_system *sys, int64_t nrhs, double *B, int64_t ldb, double tol,
                     int32_t maxit, int32_t *iters, double *berr);
/*
 * Adjoint solve on the same factorisation: Z^H X = B (conjugate transpose;
 * B[2N x nrhs] host, overwritten by X), the Z^H lambda = dsigma/dx solve of
 * an adjoint sensitivity.  MOM_PREC_FP64: direct solve with the factors.
 * MOM_PREC_MIXED: the FP32 factors as preconditioner (R M^-H C) and GMRES-IR
 * against the complex128 Z^H to backward error ||b - Z^H x|| / (||Z||_1 ||x||
 * + ||b||) (infinity norms) <= tol, one right-hand side at a time.  There is
 * no counterpart in the reference; linalg.f90:137-180 only solves with Z.
 */
int mom_system_solve_adjoint(mom_system *sys, int64_t nrhs, double *B, int64_t ldb, double tol,
                             int32_t maxit, int32_t *iters, double *berr);
/*
 * Adjoint sensitivities (no counterpart in the reference).  For the
 * solution x of Z x = b and an observable sigma(x) with dsigma = Re(g^T dx):
 *   dsigma/dp = -Re(lambda^T (dZ/dp) x),  lambda = conj(mu),
 *   Z^H mu = conj(g)  (mom_system_solve_adjoint).
 * mom_cs_abs_gradient: sigma_abs of cs_prtsrf (cs.f90:14-104, medium index
 * ri, domain mesh normally domain 1) for the solution x[4N] and its gradient
 * g[4N] (complex).  mom_pmchwt_deps_apply: y[4N] = (dZ/deps) x for the
 * relative permittivity eps = ri^2 of the medium of one domain mesh (the
 * derivative of that domain's PMCHWT blocks, sysmat.f90:18-105, from a fill
 * of the k-derivatives of its operators, dk/deps = k/(2 eps), and the eps
 * dependence of the H prefactor and of Zsi).  With p = Re eps, dZ/dp = dZ/deps;
 * with p = Im eps, dZ/dp = i dZ/deps.
 */
int mom_cs_abs_gradient(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                        double ri_im, const double *x, int32_t nq, const double *qw, const double *qnodes,
                        double *g, double *cabs);
int mom_pmchwt_deps_apply(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                          double ri_im, int32_t nq, const double *qw, const double *qnodes, const double *x,
                          double *y);
/*
 * Vertex-coordinate adjoint (the vertex part of mom_fill_adjoint; no
 * counterpart in the reference): with x the solution and lambda = Z^-T g
 * (g = dsigma/dx, i.e. conj of mom_system_solve_adjoint's solution for
 * conj(g)),
 *   dsigma/dP = dsigma/dP|_x + Re(lambda^T dq/dP) - Re(lambda^T dZ/dP x).
 * mom_pmchwt_vertex_adjoint adds Re d/dP (lambda^T Z_mesh x) of one domain's
 * blocks of the PMCHWT matrix (sysmat.f90:56-98) to dvert[3 V] (this mesh's
 * nodes, host); mom_cs_abs_vertex_explicit adds d/dP (sigma_abs at fixed x +
 * Re lambda^T q) for the exterior mesh and one plane wave (cs.f90:14-104,
 * source.f90:136-247).  Each face pair's (or face's) contribution is rebuilt
 * from its nodes with the fill's formulas and differentiated: far pairs and
 * the near pairs' smoothed point sums in closed form, the near pairs'
 * int.f90 analytic parts and the face terms by the fourth-order central
 * difference (h = 1e-3 of the mean edge length) in every coordinate of
 * their nodes; the near/far classification stays that of the mesh (or the
 * one frozen by mom_mesh_classify_like).  Every sum has a fixed order: the
 * result is the same bit for bit from run to run.
 * x, lambda: [2N] complex (host).  The caller sums the domains' dvert through
 * their nodes' master indices.
 */
int mom_pmchwt_vertex_adjoint(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                              double ri_im, int32_t nq, const double *qw, const double *qnodes, const double *x,
                              const double *lam, double *dvert);
int mom_cs_abs_vertex_explicit(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                               double ri_im, const double *x, const double *lam, double theta, double phi, double psi,
                               int32_t nq, const double *qw, const double *qnodes, double *dvert);
/* Build the plane-wave right-hand sides on the device and solve in place of
 * mom_srcvec_pw + mom_system_solve; X[2N x nsrc] receives the solutions. */
int mom_system_solve_pw(mom_system *sys, const mom_mesh *mesh, double omega, double ri_re,
                        double ri_im, int64_t nsrc, const double *angles, int32_t nq,
                        const double *qw, const double *qnodes, double tol, int32_t maxit,
                        double *X, int64_t ldx, int32_t *iters, double *berr);
/* Pivoting of the complex64 factorisation of a MOM_PREC_MIXED system (no
 * counterpart in the reference): 0 = partial pivoting as LAPACK's ZGETRF
 * (the default), 1 = tournament pivoting (communication-avoiding LU: the
 * pivot rows of each 32-column block chosen by partial pivoting inside
 * groups of 256 rows and a reduction tree over their choices, one launch
 * per tree level instead of one exchange per column).  The pivots differ
 * from LAPACK's; the complex64 factors only precondition the complex128
 * refinement, so the solution is the same to the refinement's tolerance.
 * The complex128 factorisation always pivots partially. */
int mom_system_set_pivoting(mom_system *sys, int32_t mode);
/* Precision of the factorisations of a system created MOM_PREC_MIXED (no
 * counterpart in the reference): MOM_PREC_FP64 factors the complex128
 * matrix itself, MOM_PREC_MIXED returns to the complex64 factors with
 * refinement.  Takes effect at the next fill, or at once on a system filled
 * but not yet factored.  A batch of wavelengths uses it to factor directly
 * in FP64 where the complex64 factors stop converging (interior permittivity
 * near zero).  A system created MOM_PREC_FP64 has no complex64 copy and
 * accepts only MOM_PREC_FP64. */
int mom_system_set_precision(mom_system *sys, int32_t precision);
/* Launch kind of the partial-pivoting panels (no counterpart in the
 * reference): 1 (default) one cooperative launch per 32-column block where
 * the device runs the whole grid at once, 0 one ordinary launch per column
 * (the same pivots, about 15% slower at 30720 unknowns).  Systems driven
 * from worker threads, several at once, use 0: cooperative launches made
 * from such threads were followed by a segmentation fault at process exit. */
int mom_system_set_cooperative(mom_system *sys, int32_t allow);
/* Event timing of the kernel groups of mom_system_factor (reported by
 * mom_system_stats): 0 off, 1 the depth-256 trailing updates only (cheap
 * enough for timed runs), 2 every group (adds host work per launch). */
int mom_system_profile(mom_system *sys, int32_t enable);
/*
 * Wall times of the last fill / factor / solve [s] and, with profiling
 * enabled, the factorisation's kernel time by group (23 doubles):
 * out[0] fill, out[1] factor, out[2] solve;
 * out[3] kernel time of the trailing updates (depth NB = 256) [s], out[4]
 * their number of launches, out[5] their algorithmic flops (8 M N K each);
 * out[6] panel factorisation (pivot search, row swaps in the panel, column
 * updates) [s], out[7] triangular solves of the U blocks [s], out[8] row
 * interchanges outside the panels [s];
 * out[9] applications of the FP32 factors and out[10] FP64 matrix-vector
 * products in the last mixed-precision refinement, out[11..18] its backward
 * error after outer steps 0..7 (-1 where no step was taken);
 * out[19] time of the shallower updates (inner panel and triangular-solve
 * blocks, depth 32) with their panel transposes [s], out[20] their number
 * of launches, out[21] their algorithmic flops; out[22] 1 if the last block
 * solve moved a mixed system to FP64 factors (see mom_system_solve).
 */
int mom_system_stats(const mom_system *sys, double *out);

/*
 * cs_prtsrf (cs.f90:14-104) for a plane-wave source: scattering and
 * absorption cross-sections of the solution x[2N] over the domain mesh
 * (normally domain 1), medium index ri.
 */
int mom_cs_prtsrf(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega,
                  double ri_re, double ri_im, const double *x, double theta, double phi, double psi,
                  int32_t nq, const double *qw, const double *qnodes, double *csca, double *cabs);

/*
 * far_fields (ffields.f90:16-122), identity group, for ndir directions at
 * once: the solution x[2N] (host, complex128) radiates through the domain
 * mesh (normally domain 1) in the medium of index ri; for direction d,
 * angles theta[d], phi[d] [rad], distance r, out[8d .. 8d+7] = et, ep, ht,
 * hp (complex, interleaved).
 */
int mom_far_fields(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                   double ri_im, const double *x, double r, int64_t ndir, const double *theta, const double *phi,
                   int32_t nq, const double *qw, const double *qnodes, double *out);
/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mom_far_fields with every array on the device and no
 * allocation -- x[4N] (complex128 solution), theta[ndir], phi[ndir] and
 * out[8 ndir] are device pointers; the work areas live in a fixed 64 MiB
 * device array per GPU (set up by mom_ctx_create).  It runs the same
 * kernels as mom_far_fields and returns the same values.  A problem needing
 * more scratch than 64 MiB (about 144 F Q + 64 ndir (F Q / 4096 + 3) bytes
 * for F faces of Q points; F = 20480, Q = 13 and the 2664 directions of a
 * 37 x 72 grid fit) is rejected with MOM_ERR_ARG.  Calls that overlap on one GPU (two contexts on the same
 * device at once) race on the shared array; the call is synchronous.
 */
int mom_far_fields_noalloc(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                           double ri_im, const double *x, double r, int64_t ndir, const double *theta,
                           const double *phi, int32_t nq, const double *qw, const double *qnodes, double *out);
/*
 * rcs_ext (ffields.f90:203-249): bistatic radar cross-section on the grid
 * theta = pi i/(ntheta-1), phi = 2 pi j/(nphi-1) (both ends included, pi as
 * the reference's), r = 1: rcs[i + ntheta j] = |et|^2 + |ep|^2, or with
 * filter != 0, |fdir . (et v_theta + ep v_phi)|^2 (no conjugation, as dotc).
 * rcs(...) of the reference is filter = 0.
 */
int mom_rcs_ext(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                double ri_im, const double *x, int32_t ntheta, int32_t nphi, int32_t filter, const double *fdir,
                int32_t nq, const double *qw, const double *qnodes, double *rcs);

/*
 * Second harmonic from a surface chi2 (centrosymmetric media), identity
 * group, free space: the centrosymmetric branch of solve_batch
 * (solver.f90:143-231) is
 *   per nonlinear domain d: epsp (determine_epsp, host), then
 *     mom_nlsurf_coef -> src_coef_d, src_vec_d; src_vec_d is added to the SH
 *     source at the domain's global unknowns (rows parent, parent + N);
 *   mom_sysmat_pmchwt_nls (or mom_system_fill_nls) -> SH matrix, source
 *     updated; then a solve, and mom_rcs_ext at 2 omega in domain 1.
 *
 * nlsurf_coef (nlsurf.f90:137-338) of one domain mesh: omega and ri are the
 * fundamental angular frequency and the domain's fundamental index,
 * epsp[2*nfaces] the permittivity outside each face (complex), x[4N] the
 * fundamental solution (complex, J then M), chi2[6] = chi2_nnn, chi2_ntt,
 * chi2_ttn (complex).  Outputs (complex, local edge numbering):
 * src_coef[4*nedges] = (M, J) expansion coefficients of the surface
 * polarisation source, src_vec[4*nedges] = its (E, H) source terms.  The RWG
 * Gram systems are solved by conjugate gradients on the sparse matrix
 * (MOM_ERR_NOCONV if the relative residual stays above 1e-12).
 */
int mom_nlsurf_coef(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                    double ri_im, const double *epsp, const double *x, const double *chi2, int32_t nq,
                    const double *qw, const double *qnodes, double *src_coef, double *src_vec);
/*
 * sysmat_pmchwt_nls (sysmat.f90:112-229): A[2N x 2N] (lda >= 2N) becomes the
 * PMCHWT matrix at the SH angular frequency omega (twice the fundamental)
 * with the SH indices shri[2*ndom], and for each source s the products of
 * every domain's block with its src_coef are subtracted from src[:, s]
 * (complex, ldsrc >= 2N).  src_coef is complex [ldsc x ndom x nsrc] as the
 * reference's src_coef(2N, ndom, 1, nsrc): column (d, s) starts at element
 * ldsc (d + ndom s) and holds domain d's (M, J) coefficients in its first
 * 2 nedges_d entries (zero for linear domains).
 */
int mom_sysmat_pmchwt_nls(mom_ctx *ctx, int32_t ndom, const mom_mesh *const *doms, const double *shri,
                          int64_t nedges_total, double omega, int32_t nq, const double *qw, const double *qnodes,
                          int64_t nsrc, const double *src_coef, int64_t ldsc, double *src, int64_t ldsrc,
                          double *A, int64_t lda);
/* The same into the device-resident system (then mom_system_factor and
 * mom_system_solve on src). */
int mom_system_fill_nls(mom_system *sys, int32_t ndom, const mom_mesh *const *doms, const double *shri,
                        double omega, int32_t nq, const double *qw, const double *qnodes, int64_t nsrc,
                        const double *src_coef, int64_t ldsc, double *src, int64_t ldsrc);

/*
 * Transposes for adjoint sensitivities of the second-harmonic chain (no
 * counterparts in the reference; convention as mom_pmchwt_vertex_adjoint:
 * an observable s with ds = Re(g^T dx) has g the non-conjugated gradient).
 *
 * mom_far_fields_adjoint: for adjoint values v[8 ndir] (complex,
 * interleaved) of the outputs (et, ep, ht, hp) of mom_far_fields with the
 * same arguments, g[4N] = L^T v, where out = L x: Re(g^T dx) =
 * Re(sum_d v_d . dout_d).  For s = sum_d W_d (|et|^2 + |ep|^2),
 * v = (2 W conj(et), 2 W conj(ep), 0, 0).
 *
 * mom_nlsurf_coef_adjoint: the outputs of mom_nlsurf_coef are holomorphic
 * and quadratic in x; for adjoint values src_coef_bar[4 nedges] and
 * src_vec_bar[4 nedges] (complex, the outputs' layout) it returns
 * xbar[4N] = J^T (src_coef_bar, src_vec_bar) with J the Jacobian at x
 * (complex, non-conjugated).
 *
 * mom_domain_block_apply: y[4N] = Z_d x (trans = 0) or Z_d^T x (trans = 1)
 * for the PMCHWT blocks of one domain mesh alone (sysmat.f90:56-98 for that
 * domain, medium index ri), at the domain's unknowns (zero elsewhere).
 */
int mom_far_fields_adjoint(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                           double ri_im, double r, int64_t ndir, const double *theta, const double *phi,
                           const double *v, int32_t nq, const double *qw, const double *qnodes, double *g);
int mom_nlsurf_coef_adjoint(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                            double ri_im, const double *epsp, const double *x, const double *chi2, int32_t nq,
                            const double *qw, const double *qnodes, const double *src_coef_bar,
                            const double *src_vec_bar, double *xbar);
int mom_domain_block_apply(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                           double ri_im, int32_t trans, int32_t nq, const double *qw, const double *qnodes,
                           const double *x, double *y);

/*
 * Problem set-up on the host (no GPU work): the master mesh and the domain
 * meshes that the `mesh` and `sdom` commands build (interface.f90:357-366,
 * 44-105), identity group, no periodic boundaries.  Replaces build_mesh
 * (mesh.f90:1044-1137), compute_basis_data (1265-1322), extract_submesh
 * (290-376), submesh_edge_connectivity (378-452) and orient_basis (454-558):
 * the same edge numbering, orientation and geometry bits, found with hash
 * maps instead of nested scans.
 *   p[3*nnodes]        node coordinates in file units (scaled by `scale`)
 *   fnodes[3*nfaces]   face nodes, 0-based; fids[nfaces] physical ids
 *   domain d takes the surface ids dom_ids[dom_ptr[d] .. dom_ptr[d+1]-1]
 *   (signed: a negative id reverses that surface's faces), dom_ptr[0] = 0.
 */
typedef struct mom_problem mom_problem;
int mom_problem_build(int64_t nnodes, const double *p, int64_t nfaces, const int32_t *fnodes,
                      const int32_t *fids, double scale, int32_t ndom, const int32_t *dom_ptr,
                      const int32_t *dom_ids, mom_problem **prob);
/* The same from a Gmsh 2.1/2.2 ASCII file, read as load_mesh_gmsh does
 * (mesh.f90:610-736): nodes at their node number, triangles with their first
 * tag as the surface id, other elements skipped. */
int mom_problem_load_gmsh(const char *path, double scale, int32_t ndom, const int32_t *dom_ptr,
                          const int32_t *dom_ids, mom_problem **prob);
int mom_problem_destroy(mom_problem *prob);
/* Sizes of domain dom (dom = -1: the master mesh); nedges_total is the
 * master mesh's edge count, the number of RWG functions per current.  Any
 * output pointer may be NULL. */
int mom_problem_sizes(const mom_problem *prob, int32_t dom, int64_t *nedges_total, int64_t *nnodes,
                      int64_t *nfaces, int64_t *nedges);
/* Copy out domain dom (dom = -1: the master mesh) in mom_mesh_create's
 * layout; eparent is -1 throughout for the master.  Any output pointer may
 * be NULL. */
int mom_problem_domain(const mom_problem *prob, int32_t dom, double *p, int32_t *fnodes, int32_t *fedges,
                       double *fn, double *fcp, double *fs, double *fm, double *farea, int32_t *efaces,
                       int32_t *ebn, double *elen, int32_t *eparent, double *avelen);
/* The rest of the mesh_container (mesh.f90:75-83) of domain dom (-1:
 * master): enodes[2*nedges] edge nodes, fid[nfaces] physical ids,
 * fparent[nfaces] / nparent[nnodes] master face / node of each face / node
 * (the identity on the master).  Any output pointer may be NULL. */
int mom_problem_topology(const mom_problem *prob, int32_t dom, int32_t *enodes, int32_t *fid,
                         int32_t *fparent, int32_t *nparent);
/* mom_mesh_create of domain dom (0-based) straight from the set-up. */
int mom_problem_mesh_create(mom_ctx *ctx, const mom_problem *prob, int32_t dom, mom_mesh **mesh);

/* ---------------------------------------------------------------------------
 * Symmetry (the `symm` and `sbnd` commands): the mirrors of the planes
 * x = 0, y = 0, z = 0, the rotations about the z axis (rn) and their
 * products, groups of up to 8 elements.  The mesh holds one symmetry cell;
 * its edges on a mirror plane are boundary edges (one face), classified by
 * the line elements that the mesh file puts on them.  planes[ngen] (the
 * generators, up to 4) lists them in the order of the `symm` line: 1 = mxp,
 * 2 = myp, 3 = mzp, 100 + n = rn (2 <= n <= 8, group_rz, symmetry.f90:
 * 105-156); rn with n > 2 combines only with mzp (with mxp or myp the group
 * is not abelian, MOM_ERR_ARG).  A rotation group is for bodies that no rotation cut crosses (n
 * copies of a particle around the z axis): the boundary classes of rotation
 * cuts (rz1, rz2) and of periodic cells are not supported (MOM_ERR_ARG), as
 * the reference sets no edge coupling for them (mesh.f90:195-258).
 */
/* Line elements (0-based master node pairs, ids) for classification, for a
 * problem built from arrays; mom_problem_load_gmsh keeps the file's lines. */
int mom_problem_set_lines(mom_problem *prob, int64_t nlines, const int32_t *nodes, const int32_t *ids);
/* classify_edges (mesh.f90:155-174), the `sbnd id name` command
 * (interface.f90:120-157): master edges whose node pair is a line of that id
 * get boundary class bnd (mesh_bnd_*: 2 xplane, 3 yplane, 4 zplane, mesh.f90:
 * 16-25); domain edges inherit their parent's class (mesh.f90:427-445).
 * nclassified (may be NULL) counts the matches as the reference prints them. */
int mom_problem_classify_edges(mom_problem *prob, int32_t id, int32_t bnd, int64_t *nclassified);
/* Boundary class of every edge of domain dom (-1: master), 1 = none. */
int mom_problem_edge_bnd(const mom_problem *prob, int32_t dom, int32_t *bnd);
/* read_symm (interface.f90:160-210): the group of the generators, built by
 * group_mp, group_rz and product_group (symmetry.f90:61-156, 173-241), for
 * a group whose Jacobians are all diagonal (the mirror planes): nga
 * elements; ef[g*nga + r] field action of element g in representation r;
 * detj[g]; jdiag[3g..3g+2] the diagonal Jacobian; id[g] (1 identity, 2/3/4
 * the pure mirrors, 0 products).  Outputs may be NULL. */
int mom_symm_group(int32_t ngen, const int32_t *planes, int32_t *nga, double *ef, double *detj, double *jdiag,
                   int32_t *id);
/* The same for any group of the generators above (MOM_ERR_ARG from
 * mom_symm_group for a group with rotations): ef[2 (g*nga + r) + {0, 1}]
 * the complex field action (conj(exp(i 2 pi k l / n)) for group_rz),
 * jmat[9g..9g+8] the Jacobian row-major (J[3i + j] = J_ij), id[g] (5 the
 * last rotation of group_rz, as the reference marks it).  Outputs may be
 * NULL. */
int mom_symm_group_c(int32_t ngen, const int32_t *gens, int32_t *nga, double *ef, double *detj, double *jmat,
                     int32_t *id);
/* edge_bc (bc.f90:28-118) of representation rep (0-based) for the master
 * edges' boundary classes bnd[N]: id[2N] = -1 for the unknowns the
 * representation removes, 0 otherwise; dim = the kept count, the order of
 * reduce_system's system (bc.f90:170-193).  The representation's system is
 * made with mom_system_create_order(dim + (dim & 1)): an odd order gets one
 * decoupled unknown with a unit diagonal, whose right-hand side is 0. */
int mom_symm_edge_bc(int64_t nedges_total, const int32_t *bnd, int32_t ngen, const int32_t *planes, int32_t rep,
                     int32_t *id, int64_t *dim);
/* A device system of even order n for one representation (filled by
 * mom_system_fill_symm; factor, solve and the stats work as for
 * mom_system_create's systems). */
int mom_system_create_order(mom_ctx *ctx, int64_t order, int32_t precision, mom_system **sys);
/* sysmat_pmchwt (sysmat.f90:18-105) for the group: systems[r], r < nga,
 * receive representation r's matrix in reduced numbering (the unknowns of
 * mom_symm_edge_bc with id != -1, in order), the image of every source face
 * under every admissible element (admissible_ga, bc.f90:237-257) being
 * integrated once and added to all representations.  bnd[N]: master edge
 * classes.  Other arguments as mom_system_fill. */
int mom_system_fill_symm(int32_t nrep, mom_system *const *systems, int32_t ndom, const mom_mesh *const *doms,
                         const double *ri, int64_t nedges_total, const int32_t *bnd, int32_t ngen,
                         const int32_t *planes, double omega, int32_t nq, const double *qw, const double *qnodes);
/* srcvec (source.f90:136-247) with src_field_frags' projection on the
 * representations (source.f90:249-315): q[:, r + nga*s] (2N rows, ldq) is
 * plane wave s (angles as mom_srcvec_pw) in representation r, unreduced. */
int mom_srcvec_pw_symm(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                       double ri_im, int32_t ngen, const int32_t *planes, int64_t nsrc, const double *angles,
                       int32_t nq, const double *qw, const double *qnodes, double *q, int64_t ldq);
/* cs_prtsrf (cs.f90:14-104) and rcs_ext (ffields.f90:203-249) over every
 * image of the mesh, x[2N * nga] the expanded solutions of the nga
 * representations (expand_solution, bc.f90:199-232: removed unknowns 0). */
int mom_cs_prtsrf_symm(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                       double ri_im, int32_t ngen, const int32_t *planes, const double *x, double theta, double phi,
                       double psi, int32_t nq, const double *qw, const double *qnodes, double *csca, double *cabs);
int mom_rcs_ext_symm(mom_ctx *ctx, const mom_mesh *mesh, int64_t nedges_total, double omega, double ri_re,
                     double ri_im, int32_t ngen, const int32_t *planes, const double *x, int32_t ntheta, int32_t nphi,
                     int32_t filter, const double *fdir, int32_t nq, const double *qw, const double *qnodes,
                     double *rcs);

#ifdef __cplusplus
}
#endif
#endif /* MOM_GPU_H */
