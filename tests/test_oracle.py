"""Pins of the CPU oracle (oracle/mom_oracle.c), no GPU needed.

1. Golden vectors written by the reference Fortran itself (tests/golden/,
   made by tests/golden/make_golden.py from oracle/_ref/mom_ref): mesh
   numbering and geometry bit-exact, the PMCHWT matrix, srcvec, cs_prtsrf and
   the far field to rounding level.
2. Closed forms: the bit values of the float32-rounded constants and
   quadrature tables (constants.f90:10-16, quad.f90:115-171), polynomial
   exactness of the 13-point rule.
3. Physics: the PMCHWT solution of a sphere converges to the Mie series as
   the mesh is refined (second order), and the far-field operator pairs are
   symmetric (D = D^T, H = H^T, K = K^T on pairs without singular
   subtraction).
"""
import math
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from cases import sphere, dimer
from conftest import golden
from mie import mie_cross_sections
from mom2amd import DEGTORAD, constants as pc


def test_constants_bits():
    c = oracle.constants()
    assert c["pi"] == 3.14159274101257324
    assert c["mu0"] == 4 * 3.14159274101257324 * 1e-7
    assert c["c0"] == 2.997924580003e8 and c["eta0"] == 376.7303134622 and c["eps0"] == 8.8541878176e-12
    # product-side constants are the same numbers
    assert (pc.PI, pc.MU0, pc.C0, pc.ETA0, pc.EPS0) == (c["pi"], c["mu0"], c["c0"], c["eta0"], c["eps0"])


def test_gl13_table_bits_and_exactness():
    w, n = oracle.quad_gl13()
    wp, np_ = pc.tri_quad_data("tri_gl13")
    assert np.array_equal(w, wp) and np.array_equal(n, np_)
    assert w[0] == -0.149570047855377197
    assert abs(w.sum() - 0.999999992549419403) < 1e-18
    assert np.array_equal(n[1], [0.479308068752288818, 0.260345965623855591, 0.260345965623855591])
    # degree-7 exactness on the reference triangle (up to the float32 rounding of the table)
    for a in range(5):
        for b in range(5 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2) * 2
            got = np.sum(w * n[:, 0] ** a * n[:, 1] ** b)
            assert abs(got - exact) < 3e-8


@pytest.mark.parametrize("name", ["sphere_s1", "sphere_s2", "dimer_s1"])
def test_oracle_mesh_matches_reference(msh_dir, name):
    g = golden(name)
    case = sphere(msh_dir, 1 if name == "sphere_s1" else 2) if name.startswith("sphere") else dimer(msh_dir, 1)
    op = oracle.OracleProblem(case["path"], 1e-9, case["domains"])
    for d in range(len(case["domains"]) + 1):
        m = op.mesh(d - 1)
        for k in ("fnodes", "fedges", "efaces", "ebn", "enodes"):
            assert np.array_equal(m[k], g["mesh%d_%s" % (d, k)]), (name, d, k)
        if d > 0:
            assert np.array_equal(m["eparent"], g["mesh%d_eparent" % d])
        assert m["avelen"] == float(g["mesh%d_avelen" % d])
        if "mesh%d_p" % d in g:
            for k in ("p", "fn", "fcp", "fs", "fm", "farea", "elen"):
                assert np.array_equal(m[k], g["mesh%d_%s" % (d, k)]), (name, d, k)


@pytest.mark.parametrize("name", ["sphere_s1", "sphere_s2", "dimer_s1"])
def test_host_mesh_builder_matches_reference(msh_dir, name):
    """The product's host set-up (mom2amd.mesh) reproduces the reference's
    numbering, orientation and geometry bit for bit."""
    g = golden(name)
    case = sphere(msh_dir, 1 if name == "sphere_s1" else 2) if name.startswith("sphere") else dimer(msh_dir, 1)
    P = case["problem"]
    for d, m in enumerate([P.master] + P.domains):
        for k in ("fnodes", "fedges", "efaces", "ebn", "enodes"):
            assert np.array_equal(getattr(m, k), g["mesh%d_%s" % (d, k)]), (name, d, k)
        if d > 0:
            assert np.array_equal(m.eparent, g["mesh%d_eparent" % d])
        assert m.avelen == float(g["mesh%d_avelen" % d])
        if "mesh%d_p" % d in g:
            for k in ("p", "fn", "fcp", "fs", "fm", "farea", "elen"):
                assert np.array_equal(getattr(m, k), g["mesh%d_%s" % (d, k)]), (name, d, k)


def _blockwise(A, B):
    n = A.shape[0] // 2
    return max(np.abs(A[i * n:(i + 1) * n, j * n:(j + 1) * n] - B[i * n:(i + 1) * n, j * n:(j + 1) * n]).max() /
               np.abs(B[i * n:(i + 1) * n, j * n:(j + 1) * n]).max() for i in range(2) for j in range(2))


def test_oracle_sysmat_matches_reference_s1(msh_dir):
    g = golden("sphere_s1")
    case = sphere(msh_dir, 1)
    op = oracle.OracleProblem(case["path"], 1e-9, case["domains"])
    A = op.sysmat(case["omega"], case["ri"])
    assert _blockwise(A, g["A"]) < 1e-14


@pytest.mark.parametrize("name", ["sphere_s2", "dimer_s1"])
def test_oracle_sysmat_matches_reference_sampled(msh_dir, name):
    g = golden(name)
    case = sphere(msh_dir, 2) if name == "sphere_s2" else dimer(msh_dir, 1)
    op = oracle.OracleProblem(case["path"], 1e-9, case["domains"])
    A = op.sysmat(case["omega"], case["ri"])
    got, ref = A[g["A_i"], g["A_j"]], g["A_v"]
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-14
    N = A.shape[0] // 2
    fro = np.array([[np.linalg.norm(A[i * N:(i + 1) * N, j * N:(j + 1) * N]) for j in range(2)] for i in range(2)])
    assert np.abs(fro / g["block_fro"] - 1).max() < 1e-14


@pytest.mark.parametrize("name", ["sphere_s1", "sphere_s2", "dimer_s1"])
def test_oracle_rhs_and_postprocessing_match_reference(msh_dir, name):
    g = golden(name)
    case = {"sphere_s1": lambda: sphere(msh_dir, 1), "sphere_s2": lambda: sphere(msh_dir, 2),
            "dimer_s1": lambda: dimer(msh_dir, 1)}[name]()
    op = oracle.OracleProblem(case["path"], 1e-9, case["domains"])
    ang = np.array(g["angles_deg"]) * DEGTORAD
    q = op.srcvec(case["omega"], case["ri"][0], *ang)
    assert np.array_equal(q, g["q"])
    cs, ca = op.cross_sections(case["omega"], case["ri"][0], g["x"], *ang)
    assert abs(cs / g["csca"] - 1) < 1e-14 and abs(ca / g["cabs"] - 1) < 1e-14
    r = op.rcs(case["omega"], case["ri"][0], g["x"], 37, 72)
    assert np.abs(r - g["rcs"]).max() / np.abs(g["rcs"]).max() < 1e-13


def test_mie_convergence(msh_dir):
    """sigma_sca, sigma_abs of the oracle's PMCHWT solution approach the Mie
    series with second-order convergence in the mesh size."""
    ss, sa, _ = mie_cross_sections(50e-9, pc.omega_of(600e-9) / pc.C0, 1.0, np.sqrt(-9.4 + 1.5j))
    errs = []
    for s in (1, 2):
        case = sphere(msh_dir, s)
        op = oracle.OracleProblem(case["path"], 1e-9, case["domains"])
        A = op.sysmat(case["omega"], case["ri"])
        ang = np.array([0.0, 90.0, 0.0]) * DEGTORAD
        x = sl.solve(A, op.srcvec(case["omega"], case["ri"][0], *ang))
        cs, ca = op.cross_sections(case["omega"], case["ri"][0], x, *ang)
        errs.append((abs(cs / ss - 1), abs(ca / sa - 1)))
    assert errs[1][0] < 0.08 and errs[1][1] < 0.06
    # halving the edge length divides the error by ~4 (>= 3 allowed)
    assert errs[0][0] / errs[1][0] > 3 and errs[0][1] / errs[1][1] > 3


def test_far_pairs_symmetric(msh_dir):
    """Galerkin D, H, K are symmetric; entries coupling only far face pairs
    (no singular subtraction) must be symmetric to rounding."""
    case = sphere(msh_dir, 2)
    op = oracle.OracleProblem(case["path"], 1e-9, case["domains"])
    m = op.mesh(0)
    cp, thr = m["fcp"], 3 * m["avelen"]
    E = m["enodes"].shape[0]
    efaces = m["efaces"]
    # edge pairs whose four face pairs are all far
    d = cp[:, None, :] - cp[None, :, :]
    near = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]) < thr * thr
    fa = efaces[:, 0][:, None], efaces[:, 1][:, None]
    fb = efaces[:, 0][None, :], efaces[:, 1][None, :]
    far = ~(near[fa[0], fb[0]] | near[fa[0], fb[1]] | near[fa[1], fb[0]] | near[fa[1], fb[1]])
    for opn in "DHK":
        M = op.operator(0, opn, case["omega"], case["ri"][0])
        asym = np.abs(M - M.T)[far].max() / np.abs(M[far]).max()
        assert asym < 1e-12, (opn, asym)


def _dimer_sh(msh_dir):
    g = golden("dimer_s1")
    case = dimer(msh_dir, 1)
    op = oracle.OracleProblem(case["path"], 1e-9, case["domains"])
    return g, case, op


def test_oracle_sh_chain_matches_reference(msh_dir):
    """The second-harmonic chain of solve_batch (solver.f90:143-231) --
    determine_epsp, nlsurf_coef with its two Gram solves, sysmat_pmchwt_nls
    and the SH solve -- against the reference's own run on the dimer."""
    g, case, op = _dimer_sh(msh_dir)
    out = op.sh_chain(case["omega"], list(g["ri"]), list(g["sh_shri"]), list(g["sh_nls"]), g["sh_chi2"], g["x"])
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()
    for d in (1, 2):
        assert np.array_equal(out["epsp_%d" % d], g["sh_epsp_%d" % d])
        assert rel(out["src_coef_%d" % d], g["sh_src_coef_%d" % d]) < 1e-13
        assert rel(out["src_vec_%d" % d], g["sh_src_vec_%d" % d]) < 1e-13
    assert rel(out["rhs0"], g["sh_rhs0"]) < 1e-13
    assert rel(out["rhs"], g["sh_rhs"]) < 1e-13
    assert rel(out["x"], g["sh_x"]) < 1e-12
    r = op.rcs(2 * case["omega"], complex(g["sh_shri"][0]), out["x"], 37, 72)
    assert rel(r, g["sh_rcs"]) < 1e-12


def test_oracle_gram_closed_form(msh_dir):
    """rwg_moments against exact integration: f_m . f_n is quadratic on a
    face, so the edge-midpoint rule (weights A/3) integrates it exactly; the
    13-point table differs only by its float32 rounding."""
    case = sphere(msh_dir, 1)
    op = oracle.OracleProblem(case["path"], 1e-9, case["domains"])
    F = op.rwg_moments(0)
    m = op.mesh(0)
    E = F.shape[0]
    ref = np.zeros((E, E))
    for f in range(m["fnodes"].shape[0]):
        P = m["p"][m["fnodes"][f]]
        mids = [(P[0] + P[1]) / 2, (P[1] + P[2]) / 2, (P[2] + P[0]) / 2]
        A = m["farea"][f]
        fun = []
        for q in range(3):
            e = m["fedges"][f, q]
            plus = m["efaces"][e, 0] == f
            v = m["p"][m["ebn"][e, 0 if plus else 1]]
            c = m["elen"][e] / (2 * A) * (1 if plus else -1)
            fun.append((e, [c * (x - v) for x in mids]))
        for ea, fa in fun:
            for eb, fb in fun:
                ref[ea, eb] += A / 3 * sum(np.dot(x, y) for x, y in zip(fa, fb))
    assert np.allclose(F, F.T)
    assert np.abs(F - ref).max() / np.abs(ref).max() < 1e-7
    assert np.linalg.eigvalsh(F).min() > 0


def test_oracle_operators_match_reference_one_by_one(msh_dir):
    """computeD, computeH and computeK (sysmat.f90:272-519) of both domains of
    the s = 1 sphere, each against the reference's own dump of that operator
    (not only through their sums in sysmat_pmchwt)."""
    g = golden("ops_s1")
    case = sphere(msh_dir, 1)
    op = oracle.OracleProblem(case["path"], 1e-9, case["domains"])
    for d in range(2):
        for name in "DHK":
            R = g["op_%d_%s" % (d, name)]
            M = op.operator(d, name, case["omega"], case["ri"][d])
            err = np.abs(M - R).max() / np.abs(R).max()
            assert err < 1e-13, (d, name, err)
