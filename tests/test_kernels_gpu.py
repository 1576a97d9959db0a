"""Single-kernel GPU checks through the test hooks: the matrix-core lane
layout the LU kernels assume, and the trailing-update GEMM against numpy."""
import numpy as np
import pytest

from mom2amd import capi

pytestmark = pytest.mark.gpu


def expected_layout(prec):
    """A[i][k] in lane i + 16k, B[k][j] in lane j + 16k.  D[i][j]: f32 in
    lane j + 16*(i//4), accumulator i%4; f64 in lane j + 16*(i%4),
    accumulator i//4."""
    exp = np.zeros((64, 64, 64, 4))
    for L0 in range(64):
        i, k = L0 % 16, L0 // 16
        for L1 in range(64):
            j, k2 = L1 % 16, L1 // 16
            if k == k2:
                if prec == 32:
                    exp[L0, L1, j + 16 * (i // 4), i % 4] = 1.0
                else:
                    exp[L0, L1, j + 16 * (i % 4), i // 4] = 1.0
    return exp


@pytest.mark.parametrize("prec", [64, 32])
def test_mfma_lane_layout(gpu_ctx, prec):
    out = np.zeros(64 * 64 * 64 * 4)
    capi.check(capi.lib().mom_test_mfma_layout(gpu_ctx.h, prec, out))
    out = out.reshape(64, 64, 64, 4)
    exp = expected_layout(prec)
    if not np.array_equal(out, exp):
        for L0, L1 in [(0, 0), (1, 0), (0, 1), (5, 3), (17, 18), (33, 40), (63, 63)]:
            nz = np.argwhere(out[L0, L1] != 0)
            print("prec %d A-lane %d B-lane %d -> nonzero at (lane, acc) %s" % (prec, L0, L1, nz.tolist()),
                  flush=True)
    assert np.array_equal(out, exp)


@pytest.mark.parametrize("prec", [64, 32])
# (M, N, K, pad): leading dimensions M + pad and M + pad + 2.  An even pad
# with K = 32 or 256 takes the complex64 direct-transfer kernel, an odd pad
# (or another K) the register-staged one.
@pytest.mark.parametrize("shape", [(32, 32, 32, 4), (16, 64, 4, 3), (100, 70, 45, 3), (300, 257, 256, 3),
                                   (300, 257, 256, 4), (513, 130, 32, 3), (1000, 900, 256, 4), (64, 700, 32, 4),
                                   (2, 5, 256, 4), (258, 1, 32, 4), (4096, 4096, 256, 4)])
def test_gemm_sub(gpu_ctx, prec, shape):
    M, N, K, pad = shape
    rng = np.random.default_rng(M * 1000 + N + K)
    cplx = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    lda, ldb, ldc = M + pad, K + 1, M + pad + 2
    A = cplx(lda, K)
    B = cplx(ldb, N)
    C = cplx(ldc, N)
    if prec == 32:
        A, B, C = (X.astype(np.complex64).astype(np.complex128) for X in (A, B, C))
    ref = C.copy()
    ref[:M, :] -= A[:M, :] @ B[:K, :]
    Cv = np.asfortranarray(C).reshape(-1, order="F").view(np.float64).copy()
    Av = np.asfortranarray(A).reshape(-1, order="F").view(np.float64).copy()
    Bv = np.asfortranarray(B).reshape(-1, order="F").view(np.float64).copy()
    capi.check(capi.lib().mom_test_gemm_sub(gpu_ctx.h, prec, M, N, K, Av, lda, Bv, ldb, Cv, ldc))
    got = Cv.view(np.complex128).reshape(ldc, N, order="F")
    err = np.abs(got - ref).max() / np.abs(ref).max()
    tol = 1e-13 if prec == 64 else 1e-5
    print("gemm prec %d %s err %.3e" % (prec, shape, err), flush=True)
    assert err < tol
    # rows outside M untouched
    assert np.array_equal(got[M:, :], C[M:, :].astype(np.complex64).astype(np.complex128) if prec == 32 else C[M:, :])


@pytest.mark.parametrize("seed", [0, 1])
def test_unit_lower_inverse(gpu_ctx, seed):
    """E = I - inv(L11) formed in complex64 for the U-block product of the
    complex64 LU, against numpy in complex128 on the same rounded L."""
    import scipy.linalg as sl

    rng = np.random.default_rng(seed)
    G = rng.standard_normal((700, 256)) + 1j * rng.standard_normal((700, 256))
    _, Lfull, _ = sl.lu(G)
    ldl = 260
    L = np.zeros((ldl, 256), np.complex128, order="F")
    L[:256, :] = np.tril(Lfull[:256, :], -1) + 9.0 * np.triu(np.ones((256, 256)))  # only the strict lower part is read
    E = np.zeros((256, 256), np.complex128, order="F")
    capi.check(capi.lib().mom_test_unit_lower_inverse(gpu_ctx.h, L.ctypes.data, ldl, E.ctypes.data))
    L32 = np.tril(L[:256, :].astype(np.complex64).astype(np.complex128), -1) + np.eye(256)
    E_ref = np.eye(256) - np.linalg.inv(L32)
    scale = np.abs(E_ref).max()
    err = np.abs(E - E_ref).max() / scale
    print("seed %d: max|E| %.3g, rel err %.3e" % (seed, scale, err), flush=True)
    assert np.all(np.triu(E) == 0)
    assert err < 3e-7
