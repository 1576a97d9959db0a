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


def test_cs_abs_gradient(gpu_ctx, msh_dir):
    """sigma_abs of mom_cs_abs_gradient equals cs_prtsrf's, and since
    sigma_abs is quadratic in x, central differences along any direction v
    give Re(g^T v) up to rounding."""
    from conftest import golden
    from mom2amd import cs_abs_gradient, cs_prtsrf

    g0 = golden("sphere_s1")
    case = sphere(msh_dir, 1)
    N = case["problem"].nedges
    dm = DeviceMesh(gpu_ctx, case["problem"].domains[0])
    try:
        x = g0["x"]
        ang = np.asarray(g0["angles_deg"]) * np.pi / 180
        sab, g = cs_abs_gradient(gpu_ctx, dm, N, case["omega"], 1.0, x)
        ref = cs_prtsrf(gpu_ctx, dm, N, case["omega"], 1.0, x, *ang)[1]
        assert abs(sab / ref - 1) < 1e-12, (sab, ref)
        rng = np.random.default_rng(3)
        for _ in range(4):
            v = (rng.standard_normal(2 * N) + 1j * rng.standard_normal(2 * N)) * np.abs(x).max()
            h = 1e-3
            sp = cs_prtsrf(gpu_ctx, dm, N, case["omega"], 1.0, x + h * v, *ang)[1]
            sm = cs_prtsrf(gpu_ctx, dm, N, case["omega"], 1.0, x - h * v, *ang)[1]
            fd = (sp - sm) / (2 * h)
            an = np.dot(g, v).real
            print("directional derivative: adjoint %.12e fd %.12e" % (an, fd), flush=True)
            assert abs(an - fd) <= 1e-9 * (abs(fd) + abs(sp))
    finally:
        dm.close()


def test_deps_apply_matches_finite_differences(gpu_ctx, msh_dir):
    """(dZ/deps) x of the interior domain against central differences of the
    filled PMCHWT matrix in eps (h = 1e-4 |eps|, truncation ~1e-8)."""
    from mom2amd import pmchwt_deps_apply, sysmat_pmchwt

    case = sphere(msh_dir, 1)
    N = case["problem"].nedges
    dms = [DeviceMesh(gpu_ctx, d) for d in case["problem"].domains]
    try:
        eps = complex(case["ri"][1]) ** 2
        rng = np.random.default_rng(9)
        x = rng.standard_normal(2 * N) + 1j * rng.standard_normal(2 * N)
        x[N:] *= 300.0
        y = pmchwt_deps_apply(gpu_ctx, dms[1], N, case["omega"], case["ri"][1], x)
        for hdir in (1.0, 1j):
            h = 1e-4 * abs(eps) * hdir
            Zp = sysmat_pmchwt(gpu_ctx, dms, [1.0, np.sqrt(eps + h)], N, case["omega"])
            Zm = sysmat_pmchwt(gpu_ctx, dms, [1.0, np.sqrt(eps - h)], N, case["omega"])
            yfd = (Zp - Zm) @ x / (2 * h)
            for blk, sl in (("J", slice(0, N)), ("M", slice(N, 2 * N))):
                err = np.abs(y[sl] - yfd[sl]).max() / np.abs(yfd[sl]).max()
                print("dZ/deps x (h along %s), %s rows: rel err %.2e" % (hdir, blk, err), flush=True)
                assert err < 1e-6
    finally:
        for d in dms:
            d.close()


@pytest.mark.parametrize("prec", [MOM_PREC_FP64, MOM_PREC_MIXED], ids=["fp64", "mixed"])
def test_sigma_abs_eps_gradient_matches_finite_differences(gpu_ctx, msh_dir, prec):
    """d sigma_abs / d eps_in by the adjoint (one Z^H solve and one
    derivative fill) against central differences of the whole FP64 solve."""
    from mom2amd import cs_prtsrf, sigma_abs_eps_gradient

    case = sphere(msh_dir, 2)
    N = case["problem"].nedges
    ang = np.array([0.0, 90.0, 0.0]) * np.pi / 180
    dms = [DeviceMesh(gpu_ctx, d) for d in case["problem"].domains]
    eps = complex(case["ri"][1]) ** 2

    def solve(e, p):
        S = System(gpu_ctx, N, p)
        ri = [1.0 + 0j, np.sqrt(e)]
        S.fill(dms, ri, case["omega"])
        S.factor()
        x = S.solve_pw(dms[0], case["omega"], 1.0, ang)[:, 0]
        return S, x, ri

    try:
        S, x, ri = solve(eps, prec)
        try:
            sab, dre, dim = sigma_abs_eps_gradient(gpu_ctx, S, x, dms, 1, case["omega"], ri)
        finally:
            S.close()
        fd = []
        for hdir in (1.0, 1j):
            h = 1e-4 * abs(eps) * hdir
            vals = []
            for e in (eps + h, eps - h):
                S2, x2, _ = solve(e, MOM_PREC_FP64)
                S2.close()
                vals.append(cs_prtsrf(gpu_ctx, dms[0], N, case["omega"], 1.0, x2, *ang)[1])
            fd.append((vals[0] - vals[1]) / (2 * abs(h)))
        print("prec %d: sigma_abs %.6e; d/dRe eps adjoint %.8e fd %.8e; d/dIm eps adjoint %.8e fd %.8e"
              % (prec, sab, dre, fd[0], dim, fd[1]), flush=True)
        scale = max(abs(fd[0]), abs(fd[1]))
        assert abs(dre - fd[0]) < 1e-6 * scale and abs(dim - fd[1]) < 1e-6 * scale
    finally:
        for d in dms:
            d.close()
