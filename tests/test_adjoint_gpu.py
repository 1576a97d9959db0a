"""GPU checks of the adjoint solve Z^H X = B on the device-resident
factorisation (mom_system_solve_adjoint): against LAPACK on the same matrix,
and through the duality b2^H Z^-1 b1 = (Z^-H b2)^H b1 with the forward solve
at a size with many 256-row blocks.  Tolerances: FP64 1e-10 relative to
LAPACK; mixed backward error <= 1e-12 and 1e-7 relative to LAPACK, the
bar of the forward mixed solves."""
import numpy as np
import pytest
import scipy.linalg as sl

from cases import sphere
from mom2amd import DeviceMesh, System
from mom2amd.capi import MOM_PREC_FP64, MOM_PREC_MIXED

pytestmark = pytest.mark.gpu


def _system(ctx, case, prec):
    dms = [DeviceMesh(ctx, d) for d in case["problem"].domains]
    S = System(ctx, case["problem"].nedges, prec)
    S.fill(dms, case["ri"], case["omega"])
    return S, dms


@pytest.mark.parametrize("prec", [MOM_PREC_FP64, MOM_PREC_MIXED], ids=["fp64", "mixed"])
def test_adjoint_solve_matches_lapack(gpu_ctx, msh_dir, prec):
    case = sphere(msh_dir, 2)
    S, dms = _system(gpu_ctx, case, prec)
    try:
        A = S.get_matrix()
        S.factor()
        rng = np.random.default_rng(7)
        n = A.shape[0]
        for nrhs in (1, 3, 5):
            B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
            X = S.solve_adjoint(B)
            R = sl.solve(A.conj().T, B)
            err = np.abs(X - R).max() / np.abs(R).max()
            be = (np.abs(B - A.conj().T @ X).max(axis=0) /
                  (np.abs(A).sum(axis=0).max() * np.abs(X).max(axis=0) + np.abs(B).max(axis=0))).max()
            print("prec %d nrhs %d: rel err %.2e, backward error %.2e, iters %d" % (prec, nrhs, err, be, S.iters),
                  flush=True)
            assert be <= 1e-12
            # mixed: the same forward-error bar as the forward mixed solves (the
            # backward error above is the refinement's target)
            assert err < (1e-10 if prec == MOM_PREC_FP64 else 1e-7)
    finally:
        S.close()
        for d in dms:
            d.close()


@pytest.mark.parametrize("prec", [MOM_PREC_FP64, MOM_PREC_MIXED], ids=["fp64", "mixed"])
def test_adjoint_duality_many_blocks(gpu_ctx, msh_dir, prec):
    """s=3: 2N = 3840 unknowns, 15 blocks of 256 rows."""
    case = sphere(msh_dir, 3)
    S, dms = _system(gpu_ctx, case, prec)
    try:
        S.factor()
        n = 2 * case["problem"].nedges
        rng = np.random.default_rng(11)
        b1 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        b2 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        x = S.solve(b1)[:, 0]
        y = S.solve_adjoint(b2)[:, 0]
        lhs, rhs = np.vdot(b2, x), np.vdot(y, b1)
        err = abs(lhs - rhs) / abs(lhs)
        print("prec %d: b2^H Z^-1 b1 = %s, (Z^-H b2)^H b1 = %s, rel diff %.2e" % (prec, lhs, rhs, err), flush=True)
        # mixed: both solves carry forward errors up to 1e-7 relative (the
        # normwise backward error does not see the 1e-5 scaling of the M block)
        assert err < (1e-10 if prec == MOM_PREC_FP64 else 1e-6)
    finally:
        S.close()
        for d in dms:
            d.close()
