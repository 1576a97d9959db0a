"""GPU parity of the second-harmonic (surface chi2) path: nlsurf_coef,
sysmat_pmchwt_nls, the SH solve and the SH far field, against the
reference's own run on the dimer (tests/golden/dimer_s1.npz) and against the
oracle on a larger mesh.  Tolerances: nlsurf_coef 1e-12, SH source 1e-10
(relative to the largest entry), SH solution and RCS 1e-8 (FP64 and mixed)."""
import numpy as np
import pytest

import oracle
from cases import dimer, sphere
from conftest import golden
from mom2amd import (DeviceMesh, System, determine_epsp, nlsurf_coef, rcs, solve_sh, sysmat_pmchwt,
                     sysmat_pmchwt_nls)
from mom2amd.capi import MOM_PREC_FP64, MOM_PREC_MIXED

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


@pytest.fixture(scope="module")
def dimer_case(gpu_ctx, msh_dir):
    g = golden("dimer_s1")
    case = dimer(msh_dir, 1)
    dms = [DeviceMesh(gpu_ctx, d) for d in case["problem"].domains]
    yield g, case, dms
    for d in dms:
        d.close()


def test_nlsurf_coef_golden(gpu_ctx, dimer_case):
    g, case, dms = dimer_case
    P = case["problem"]
    for d in (1, 2):
        ep = determine_epsp(P, d, list(g["ri"]))
        sc, sv = nlsurf_coef(gpu_ctx, dms[d], P.nedges, case["omega"], g["ri"][d], ep, g["x"], g["sh_chi2"])
        e1, e2 = rel(sc, g["sh_src_coef_%d" % d]), rel(sv, g["sh_src_vec_%d" % d])
        print("domain %d: src_coef %.3e src_vec %.3e" % (d, e1, e2), flush=True)
        assert e1 < 1e-12 and e2 < 1e-12


def test_sysmat_nls_golden(gpu_ctx, dimer_case):
    """The SH matrix is the PMCHWT matrix at 2 omega with the SH indices; the
    source after sysmat_pmchwt_nls matches the reference's."""
    g, case, dms = dimer_case
    P = case["problem"]
    N = P.nedges
    sc = np.zeros((2 * N, 3), np.complex128)
    for d in (1, 2):
        E = dms[d].nedges
        sc[:2 * E, d] = g["sh_src_coef_%d" % d]
    om2 = 2 * case["omega"]
    A, rhs = sysmat_pmchwt_nls(gpu_ctx, dms, list(g["sh_shri"]), N, om2, sc, g["sh_rhs0"])
    e = rel(rhs, g["sh_rhs"])
    print("SH source after sysmat_pmchwt_nls: %.3e" % e, flush=True)
    assert e < 1e-10
    A0 = sysmat_pmchwt(gpu_ctx, dms, list(g["sh_shri"]), N, om2)
    assert rel(A, A0) < 1e-13
    # two sources: the second with a scaled src_coef
    sc2 = np.stack([sc, 0.5j * sc], axis=2)
    b2 = np.stack([g["sh_rhs0"], 0.5j * g["sh_rhs0"]], axis=1)
    _, rhs2 = sysmat_pmchwt_nls(gpu_ctx, dms, list(g["sh_shri"]), N, om2, sc2, b2)
    assert rel(rhs2[:, 0], g["sh_rhs"]) < 1e-10
    assert rel(rhs2[:, 1], 0.5j * g["sh_rhs"]) < 1e-10


@pytest.mark.parametrize("prec", [MOM_PREC_FP64, MOM_PREC_MIXED], ids=["fp64", "mixed"])
def test_sh_chain_golden(gpu_ctx, dimer_case, prec):
    g, case, dms = dimer_case
    P = case["problem"]
    out = solve_sh(gpu_ctx, P, dms, case["omega"], list(g["ri"]), list(g["sh_shri"]), list(g["sh_nls"]),
                   g["sh_chi2"], g["x"], precision=prec)
    ex = rel(out["x"], g["sh_x"])
    r = rcs(gpu_ctx, dms[0], P.nedges, 2 * case["omega"], complex(g["sh_shri"][0]), out["x"], 37, 72)
    er = rel(r, g["sh_rcs"])
    print("prec %d: SH x %.3e, SH rcs %.3e, iters %d berr %.1e" % (prec, ex, er, out["iters"], out["berr"]),
          flush=True)
    assert ex < 1e-8 and er < 1e-8


def test_nlsurf_coef_matches_oracle_sphere(gpu_ctx, msh_dir):
    """Sphere s=3 (1280 faces, 1920 edges) with surface chi2 on the inner
    domain, all three chi2 components non-zero."""
    case = sphere(msh_dir, 3)
    P = case["problem"]
    op = oracle.OracleProblem(case["path"], 1e-9, case["domains"])
    N = P.nedges
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(2 * N) + 1j * rng.standard_normal(2 * N))
    x[N:] *= 300.0
    chi2 = (1.0 + 0.2j, -0.3 + 0.1j, 0.25 - 0.05j)
    ri = case["ri"]
    ep = determine_epsp(P, 1, ri)
    dm = DeviceMesh(gpu_ctx, P.domains[1])
    try:
        sc, sv = nlsurf_coef(gpu_ctx, dm, N, case["omega"], ri[1], ep, x, chi2)
    finally:
        dm.close()
    rsc, rsv = op.nlsurf_coef(1, case["omega"], ri[1], op.epsp(1, ri), x, np.array(chi2))
    e1, e2 = rel(sc, rsc), rel(sv, rsv)
    print("sphere s=3: src_coef %.3e src_vec %.3e" % (e1, e2), flush=True)
    assert e1 < 1e-12 and e2 < 1e-12


def test_system_fill_nls_without_sources(gpu_ctx, dimer_case):
    """With every src_coef zero the SH fill is the plain PMCHWT fill and the
    source is unchanged."""
    g, case, dms = dimer_case
    N = case["problem"].nedges
    S = System(gpu_ctx, N, MOM_PREC_FP64)
    try:
        b = g["sh_rhs0"].copy()
        out = S.fill_nls(dms, list(g["sh_shri"]), 2 * case["omega"], np.zeros((2 * N, 3)), b)
        assert np.array_equal(out, b)
        A = S.get_matrix()
    finally:
        S.close()
    A0 = sysmat_pmchwt(gpu_ctx, dms, list(g["sh_shri"]), N, 2 * case["omega"])
    assert rel(A, A0) < 1e-13
