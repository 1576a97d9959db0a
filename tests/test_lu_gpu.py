"""GPU parity of the dense complex LU / solves against LAPACK (scipy) and of
the mixed-precision GMRES-IR path against the complex128 solution.

LAPACK is the reference's own solver (ZGETRF/ZGETRS via solve_multi_linsys,
linalg.f90:159-180).  Pins: identical pivot sequence on matrices without
near-ties, factors within 1e-12 relative, backward error ~ machine eps."""
import numpy as np
import pytest
import scipy.linalg as sl

from mom2amd import MU0, System, capi, solve_multi_linsys
from mom2amd.capi import MOM_PREC_FP64, MOM_PREC_MIXED, MomError

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[(0, 1, 1), (0, 0, 1), (1, 1, 1), (0, 1, 0)],
                ids=["coop_panel", "coop_panel_no_lookahead", "column_panel", "coop_panel_inverse_u"])
def panel_mode(request):
    """Both panel factorisation paths (one cooperative launch per inner
    block of 32 columns, one launch per column), with the look-ahead
    schedule (panel k+1 on a second stream beside the trailing update of
    step k) and without it, and the complex64 U blocks by substitution and
    by the product with I - inv(L11)."""
    mode, lookahead, u_subst = request.param
    capi.check(capi.lib().mom_test_panel_mode(mode))
    capi.check(capi.lib().mom_test_lookahead(lookahead))
    capi.check(capi.lib().mom_test_u_solve(u_subst))
    yield request.param
    capi.check(capi.lib().mom_test_panel_mode(0))
    capi.check(capi.lib().mom_test_lookahead(1))
    capi.check(capi.lib().mom_test_u_solve(1))


def rand_complex(n, m=None, seed=0, diag=0.0):
    rng = np.random.default_rng(seed)
    m = n if m is None else m
    A = rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))
    if diag:
        A += diag * np.eye(n, m)
    return np.asfortranarray(A)


def berr(A, x, b):
    r = b - A @ x
    return np.abs(r).max() / (np.abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())


@pytest.mark.parametrize("n", [7, 64, 300, 517, 1100, 3000])
def test_lu_matches_lapack(gpu_ctx, panel_mode, n):
    A = rand_complex(n, seed=n)
    B = rand_complex(n, 3, seed=n + 1)
    LU, X, ipiv = solve_multi_linsys(gpu_ctx, A, B)
    lu_ref, piv_ref = sl.lu_factor(A)
    print("n=%d pivots equal: %s" % (n, np.array_equal(ipiv - 1, piv_ref)), flush=True)
    assert np.array_equal(ipiv - 1, piv_ref)
    err = np.abs(LU - lu_ref).max() / np.abs(lu_ref).max()
    print("n=%d LU rel err %.3e" % (n, err), flush=True)
    assert err < 1e-12 * max(1, n / 100)
    X_ref = sl.lu_solve((lu_ref, piv_ref), B)
    assert np.abs(X - X_ref).max() / np.abs(X_ref).max() < 1e-10
    assert berr(A, X[:, 0], B[:, 0]) < 1e-14


def test_lu_zero_pivot_reports_info(gpu_ctx, panel_mode):
    A = rand_complex(40, seed=3)
    A[:, 5] = 0.0
    with pytest.raises(MomError) as e:
        solve_multi_linsys(gpu_ctx, A, rand_complex(40, 1, seed=4))
    assert e.value.status == 6


@pytest.mark.parametrize("n2", [512, 1200])
def test_fp64_system_solve(gpu_ctx, n2):
    A = rand_complex(n2, seed=10 + n2)
    b = rand_complex(n2, 2, seed=11)
    S = System(gpu_ctx, n2 // 2, MOM_PREC_FP64)
    S.set_matrix(A)
    S.factor()
    x = S.solve(b)
    x_ref = np.linalg.solve(A, b)
    assert np.abs(x - x_ref).max() / np.abs(x_ref).max() < 1e-10


@pytest.mark.parametrize("prec", [MOM_PREC_FP64, MOM_PREC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("nrhs", [5, 24])
def test_system_solve_many_rhs(gpu_ctx, prec, nrhs):
    """A block of right-hand sides goes through the batched solves (matrix
    products per 256-row block, block refinement); every column must match
    numpy's solution and, in the mixed path, reach the FP64 backward error."""
    n2 = 1500
    A = rand_complex(n2, seed=31)
    A[n2 // 2:, :] *= 1e-5
    A[:, n2 // 2:] *= 3e3
    B = rand_complex(n2, nrhs, seed=32)
    S = System(gpu_ctx, n2 // 2, prec)
    S.set_matrix(A)
    S.factor()
    X = S.solve(B, tol=1e-13, maxit=20)
    st = S.stats()
    print("nrhs=%d prec=%d: history %s, %d block applications" % (
        nrhs, prec, ["%.1e" % v for v in st["berr_history"] if v >= 0], st["n_minv"]), flush=True)
    X_ref = np.linalg.solve(A, B)
    for j in range(nrhs):
        assert berr(A, X[:, j], B[:, j]) < (1e-13 if prec == MOM_PREC_MIXED else 1e-14), j
    assert np.abs(X - X_ref).max() / np.abs(X_ref).max() < 1e-7
    # the first column alone (vector path) gives the same solution
    x1 = S.solve(B[:, :1], tol=1e-13, maxit=20)
    assert np.abs(x1[:, 0] - X[:, 0]).max() / np.abs(X[:, 0]).max() < 1e-9


@pytest.mark.parametrize("n2", [512, 1500])
def test_mixed_system_reaches_fp64_backward_error(gpu_ctx, panel_mode, n2):
    A = rand_complex(n2, seed=20 + n2)
    # rows and columns of very different scales, as the PMCHWT M-block has
    A[n2 // 2:, :] *= 1e-5
    A[:, n2 // 2:] *= 3e3
    b = rand_complex(n2, 1, seed=21)
    S = System(gpu_ctx, n2 // 2, MOM_PREC_MIXED)
    S.set_matrix(A)
    S.factor()
    try:
        x = S.solve(b, tol=1e-13, maxit=20)
    finally:
        st = S.stats()
        print("n=%d mixed: backward errors %s, %d factor solves" % (
            n2, ["%.1e" % v for v in st["berr_history"]], st["n_minv"]), flush=True)
    print("n=%d mixed: iters %d berr %.3e" % (n2, S.iters, S.berr), flush=True)
    assert S.berr <= 1e-13
    assert berr(A, x[:, 0], b[:, 0]) < 1e-13
    x_ref = np.linalg.solve(A, b)
    # forward error is bounded by cond(A) * backward error
    assert np.abs(x - x_ref).max() / np.abs(x_ref).max() < 1e-7


def test_mixed_factor_accuracy_at_scale(gpu_ctx):
    """The FP32 factors of an n = 8192 random system must give a first
    solution whose backward error is that of a complex64 LU (a few 1e-6):
    this catches stage races in the trailing-update kernels and inverse
    rounding in the triangular solves, which leave the refinement converging
    but from a much worse start."""
    n2 = 8192
    A = rand_complex(n2, seed=8192)
    b = rand_complex(n2, 1, seed=8193)
    S = System(gpu_ctx, n2 // 2, MOM_PREC_MIXED)
    S.set_matrix(A)
    S.factor()
    S.solve(b, tol=1e-13, maxit=30)
    hist = S.stats()["berr_history"]
    print("n=%d mixed backward errors %s" % (n2, ["%.1e" % v for v in hist]), flush=True)
    assert hist[0] < 1e-5
    assert hist[1] < 1e-7
    S.close()


@pytest.mark.parametrize("where", ["entry", "column"])
def test_lu_nan_reports_without_fault(gpu_ctx, panel_mode, where):
    """NaN in the matrix (one entry, or a whole column so that no pivot
    candidate compares): the factorisation ends normally, every pivot row
    lies inside the matrix, and the NaN shows either as INFO > 0 (zero
    pivot, the reference's STOP path) or in the solution."""
    n = 300
    A = rand_complex(n, seed=5)
    if where == "entry":
        A[17, 3] = np.nan
    else:
        A[:, 3] = np.nan
    B = rand_complex(n, 1, seed=6)
    try:
        _, X, ipiv = solve_multi_linsys(gpu_ctx, A, B)
    except MomError as e:
        assert e.status > 0, e
        return
    assert ((ipiv >= np.arange(1, n + 1)) & (ipiv <= n)).all()
    assert not np.isfinite(X).all()


def test_block_refinement_falls_back_to_fp64(gpu_ctx):
    """A matrix with condition number 1e9 defeats the FP32 factors (kappa
    eps32 > 1): every column of a block solve stalls, the system factors its
    complex128 matrix instead and solves directly; the next matrix set
    brings the mixed path back."""
    n = 1024
    rng = np.random.default_rng(21)
    U, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = np.asfortranarray((U * np.logspace(0, -9, n)) @ V.conj().T)
    B = rand_complex(n, 8, seed=22)
    S = System(gpu_ctx, n // 2, MOM_PREC_MIXED)
    try:
        S.set_matrix(A)
        S.factor()
        X = S.solve(B)
        st = S.stats()
        print("ill-conditioned block: fell back %s, refinement history %s" % (st["fell_back_to_fp64"],
                                                                             st["berr_history"]), flush=True)
        assert st["fell_back_to_fp64"]
        for j in range(B.shape[1]):
            assert berr(A, X[:, j], B[:, j]) < 1e-14, j
        x1 = S.solve(B[:, :1])  # later solves use the FP64 factors
        assert np.abs(x1[:, 0] - X[:, 0]).max() <= 1e-12 * np.abs(X[:, 0]).max()
        A2 = rand_complex(n, seed=23, diag=30.0)
        S.set_matrix(A2)
        S.factor()
        X2 = S.solve(B)
        st2 = S.stats()
        assert not st2["fell_back_to_fp64"] and S.berr > 0
        for j in range(B.shape[1]):
            assert berr(A2, X2[:, j], B[:, j]) < 1e-13, j
    finally:
        S.close()


@pytest.mark.parametrize("n2", [66, 1100, 3000])
def test_tournament_panels(gpu_ctx, n2):
    """The complex64 factorisation with tournament pivoting
    (mom_system_set_pivoting 1) against the same system factored with
    partial pivoting (the default):
    both first solutions have the backward error of a complex64 LU, and the
    refined solutions agree with LAPACK's to the same backward error.  Sizes
    cover a single partial inner block, several panels with a partial last
    one, and several groups of 1024 rows (two tree levels)."""
    A = rand_complex(n2, seed=40 + n2)
    A[n2 // 2:, :] *= 1e-5
    A[:, n2 // 2:] *= 3e3
    b = rand_complex(n2, 1, seed=41)
    x_ref = np.linalg.solve(A, b)
    out = {}
    for mode in (1, 0):
        S = System(gpu_ctx, n2 // 2, MOM_PREC_MIXED)
        try:
            S.set_pivoting(mode)
            S.set_matrix(A)
            S.factor()
            x = S.solve(b, tol=1e-13, maxit=20)
            st = S.stats()
        finally:
            S.close()
        hist = st["berr_history"]
        print("n=%d tournament=%d: backward errors %s" % (n2, mode, ["%.1e" % v for v in hist]), flush=True)
        assert hist[0] < 1e-4, hist
        assert berr(A, x[:, 0], b[:, 0]) < 1e-13
        # forward errors: cond(A) (about 1e6 with these row and column scales) times the backward error
        assert np.abs(x - x_ref).max() / np.abs(x_ref).max() < 1e-5
        out[mode] = x
    assert np.abs(out[1] - out[0]).max() / np.abs(out[0]).max() < 1e-5


def _tournament_solve(S, A, b):
    S.set_matrix(A)
    S.factor()
    x = S.solve(b, tol=1e-13, maxit=20)
    return x, S.stats()["berr_history"]


def test_factor_graph_replays_match_eager(gpu_ctx):
    """The tournament factorisation replayed from its captured HIP graph
    gives bit for bit the solutions and refinement history of the same
    launches made eagerly: one system captures on its first factorisation and
    replays on the next ones, with eager work (uploads, solves, another
    system's factorisation that may move the shared GEMM scratch) between the
    replays; every replay is checked against the eager result of its
    matrix."""
    L = capi.lib()
    cnt = np.zeros(2, dtype=np.int64)
    mats = [rand_complex(1100, seed=60), rand_complex(1100, seed=61)]
    for A in mats:
        A[550:, :] *= 1e-4
    b = rand_complex(1100, 2, seed=62)
    eager = []
    capi.check(L.mom_test_factor_graph(0, None))
    try:
        S = System(gpu_ctx, 550, MOM_PREC_MIXED)
        S.set_pivoting(1)
        for A in mats:
            eager.append(_tournament_solve(S, A, b))
        S.close()
    finally:
        capi.check(L.mom_test_factor_graph(1, None))
    S = System(gpu_ctx, 550, MOM_PREC_MIXED)
    S2 = System(gpu_ctx, 1500, MOM_PREC_MIXED)
    try:
        S.set_pivoting(1)
        S2.set_pivoting(1)
        capi.check(L.mom_test_factor_graph(-1, cnt.ctypes.data_as(capi.C.POINTER(capi._i64))))
        c0 = cnt.copy()
        for rep, i in enumerate([0, 1, 0, 1]):
            if rep == 2:
                # a larger system of the same context in between
                _tournament_solve(S2, rand_complex(3000, seed=63), rand_complex(3000, 1, seed=64))
            x, hist = _tournament_solve(S, mats[i], b)
            assert np.array_equal(x, eager[i][0]), rep
            assert hist == eager[i][1], rep
            assert berr(mats[i], x[:, 0], b[:, 0]) < 1e-13
        capi.check(L.mom_test_factor_graph(-1, cnt.ctypes.data_as(capi.C.POINTER(capi._i64))))
        assert cnt[0] - c0[0] >= 2  # S, then S2 (S may capture again after S2 moved the scratch)
        assert cnt[1] - c0[1] >= 2  # S replayed at least on its second factorisation and after
    finally:
        S.close()
        S2.close()


def test_set_precision(gpu_ctx):
    """mom_system_set_precision: a mixed system factored in FP64 for one
    matrix (the refinement then has nothing to do: no preconditioner
    applications) and back to complex64 factors for the next; a system
    created FP64 refuses the mixed path."""
    A = rand_complex(600, seed=70)
    b = rand_complex(600, 1, seed=71)
    S = System(gpu_ctx, 300, MOM_PREC_MIXED)
    try:
        for prec, minv in ((MOM_PREC_FP64, False), (MOM_PREC_MIXED, True)):
            S.set_precision(prec)
            S.set_matrix(A)
            S.factor()
            x = S.solve(b, tol=1e-13)
            assert berr(A, x[:, 0], b[:, 0]) < 1e-13
            assert (S.stats()["n_minv"] > 0) == minv
    finally:
        S.close()
    S = System(gpu_ctx, 300, MOM_PREC_FP64)
    try:
        with pytest.raises(MomError):
            S.set_precision(MOM_PREC_MIXED)
    finally:
        S.close()
