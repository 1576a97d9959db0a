/*
 * mom_gpu_testing.h -- test hooks of libmomgpu.so.
 *
 * These entry points have no counterpart in the reference; they exist so
 * that the test suite can check single kernels of the hot path (matrix-core
 * lane layout, the trailing-update GEMM) in isolation through the C-ABI.
 */
#ifndef MOM_GPU_TESTING_H
#define MOM_GPU_TESTING_H

#include <stdint.h>

#include "mom_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Lane layout probe of the 16x16x4 matrix-core instruction used by the LU
 * (precision 64: v_mfma_f64_16x16x4f64, 32: v_mfma_f32_16x16x4f32): for
 * every pair (L0, L1) of lanes one wave multiplies A = e(L0), B = e(L1)
 * (lane L0 / L1 holds 1, all other lanes 0) and writes its 64 x 4
 * accumulators to out[((L0*64 + L1)*64 + lane)*4 + v] (1048576 doubles).
 */
int mom_test_mfma_layout(mom_ctx *ctx, int32_t precision, double *out);

/*
 * C -= A B with the trailing-update kernel of the LU, complex column-major:
 * A[M x K] (lda), B[K x N] (ldb), C[M x N] (ldc), interleaved (re, im),
 * in complex128 (precision 64) or complex64 (precision 32, values rounded
 * on upload and widened on return).
 */
int mom_test_gemm_sub(mom_ctx *ctx, int32_t precision, int64_t M, int64_t N, int64_t K, const double *A,
                      int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc);

/*
 * Panel factorisation path of mom_system_factor and mom_solve_multi_linsys:
 * 0 (default) factors each inner block of 32 columns in one cooperative
 * launch when its workgroups fit on the device at once, 1 always uses one
 * launch per column.  Process-wide; lets the tests check both paths.
 */
int mom_test_panel_mode(int32_t mode);

/*
 * Look-ahead of the blocked LU: 1 (default) factors panel k+1 on a second
 * stream while the trailing update of step k runs, 0 runs every step on
 * one stream.  Process-wide; lets the tests check both schedules.
 */
int mom_test_lookahead(int32_t enable);

/*
 * U blocks of the complex64 LU: 1 (default) blocked substitution, 0 one
 * depth-256 product with E = I - inv(L11) where the shapes allow it.
 * Process-wide; lets the tests check both.
 */
int mom_test_u_solve(int32_t mode);

/*
 * E = I - inv(L11) for a unit lower triangular 256 x 256 block L11 (the
 * strict lower part of L[256 x 256] (ldl), complex128, rounded to complex64
 * on upload), as the complex64 LU forms it after each panel; E (256 x 256,
 * leading dimension 256) is returned widened to complex128.
 */
int mom_test_unit_lower_inverse(mom_ctx *ctx, const double *L, int64_t ldl, double *E);

/* Node-coordinate adjoint: 1 = every face pair by fourth-order differences
 * (the closed forms of the far pairs and of the near pairs' smoothed point
 * loop, and the forward-mode derivatives of the near pairs' analytic parts,
 * are checked against it), 0 = the default. */
int mom_test_vertex_fd(int32_t enable);

/* Block refinement of many right-hand sides in the mixed path: 1 = take the
 * branch that refactors the resident complex128 matrix in FP64 whenever a
 * column is above the tolerance after the first residual, 0 = the default
 * (taken when finishing the stalled columns by GMRES would cost more).
 * Process-wide; lets the tests run that branch on any matrix. */
int mom_test_batch_fallback(int32_t force);

/* Panels of the complex64 factorisation (mixed path) of every system:
 * 1 = tournament pivoting whatever the system chose (mom_system_set_pivoting),
 * 0 = each system's own choice.  Process-wide. */
int mom_test_tournament(int32_t enable);

/* Complex64 factorisations with tournament pivoting replay a HIP graph
 * captured on the system's first factorisation (captured again when a
 * buffer it baked in moves): enable 1 (default) or 0 switches that
 * process-wide, -1 leaves it; counts (if not NULL) receives the process's
 * captures [0] and replays [1] so far. */
int mom_test_factor_graph(int32_t enable, int64_t *counts);

/* The fill's pair kernel in a lossless medium (Im k = 0) skips the
 * exponential of e^{ikR}, which is exactly 1 there: 1 (default) takes that
 * variant, 0 the general one (the same bits, checked by a test). */
int mom_test_fill_lossless(int32_t enable);

/* The fill's pair-kernel variant of a context: 2, 4 or 8 source faces per
 * wave, or 0 to let the next fill's autotuner choose again.  Every variant
 * gives the same bits (checked by a test). */
int mom_test_fill_spw(mom_ctx *ctx, int32_t spw);

#ifdef __cplusplus
}
#endif
#endif
