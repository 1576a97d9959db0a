"""GPU parity of the batched spectrum path (many wavelengths, many plane
waves on one reused System; solver.f90:75-140) against the reference's own
run of the same problem (tests/golden/spectrum_s2.npz: the 80/40/40 nm
ellipsoid from icosphere s=2, three wavelengths of `wlrg 64 400d-9 900d-9`
with the Drude index, 8 plane waves solved together by its solve_systems).

Tolerances: matrix entries 1e-10 of their block's largest entry, sources
1e-12, solutions x and sigma_sca, sigma_abs, rcs 1e-8 (relative), for the
mixed path (partial-pivoting panels, and tournament-pivoting panels), the
FP64 path and the mixed path's FP64-factor branch, one wavelength at a time
and two at a time (lanes)."""
import os

import numpy as np
import pytest

from conftest import golden
from mom2amd import DEGTORAD, DeviceMesh, capi, cs_prtsrf, load_problem, meshgen, omega_of, rcs, srcvec, sysmat_pmchwt
from mom2amd.capi import MOM_PREC_FP64, MOM_PREC_MIXED
from mom2amd.spectrum import drude_index, solve_spectrum

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spectrum_case(msh_dir):
    g = golden("spectrum_s2")
    nodes, faces = meshgen.icosphere(2, 1.0)
    path = os.path.join(str(msh_dir), "ellipsoid2.msh")
    meshgen.write_gmsh(path, nodes * np.array([80.0, 40.0, 40.0]), faces, 15)
    prob = load_problem(path, 1e-9, [[15], [-15]])
    return g, prob


def _relerr(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_spectrum_setup_and_fill_match_reference(gpu_ctx, spectrum_case):
    g, prob = spectrum_case
    for d, dm in enumerate([prob.master] + list(prob.domains)):
        assert np.array_equal(dm.fedges, g["mesh%d_fedges" % d]), d
        if d > 0:  # the reference sets parent_index on the domain meshes only
            assert np.array_equal(dm.eparent, g["mesh%d_eparent" % d]), d
    dms = [DeviceMesh(gpu_ctx, d) for d in prob.domains]
    N = prob.nedges
    for i, wl in enumerate(g["wl"]):
        ri = list(g["ri_%d" % i])
        assert ri[1] == drude_index(wl)
        om = omega_of(wl)
        A = sysmat_pmchwt(gpu_ctx, dms, ri, N, om)
        fro = np.array([[np.linalg.norm(A[a * N:(a + 1) * N, b * N:(b + 1) * N]) for b in range(2)] for a in range(2)])
        assert np.abs(fro / g["block_fro_%d" % i] - 1).max() < 1e-10
        scale = np.array([[np.abs(A[a * N:(a + 1) * N, b * N:(b + 1) * N]).max() for b in range(2)] for a in range(2)])
        ii, jj = g["A_i_%d" % i], g["A_j_%d" % i]
        err = (np.abs(A[ii, jj] - g["A_v_%d" % i]) / scale[ii // N, jj // N]).max()
        q = srcvec(gpu_ctx, dms[0], N, om, ri[0], g["angles_deg"] * DEGTORAD)
        qerr = _relerr(q, g["q_%d" % i])
        print("wl=%.1f nm: sampled A err %.2e, q err %.2e" % (wl * 1e9, err, qerr), flush=True)
        assert err < 1e-10
        assert qerr < 1e-12


@pytest.mark.parametrize("mode", ["mixed", "mixed_tournament", "fp64", "mixed_fp64_fallback",
                                  "mixed_tournament_2lanes", "fp64_2lanes", "mixed_fp64_fallback_2lanes"])
def test_spectrum_solutions_match_reference(gpu_ctx, spectrum_case, mode):
    """Every path of solve_spectrum against the reference's run.  The
    *_fp64_fallback modes force the block refinement's FP64-factor branch:
    the wavelength after a fallback is then factored directly in FP64 (the
    driver's policy), the one after that tries complex64 again.  The 2lanes
    modes solve the three wavelengths two at a time on two contexts."""
    g, prob = spectrum_case
    dms = [DeviceMesh(gpu_ctx, d) for d in prob.domains]
    N = prob.nedges
    ang = g["angles_deg"]
    prec = MOM_PREC_FP64 if mode.startswith("fp64") else MOM_PREC_MIXED
    fallback = "fp64_fallback" in mode
    capi.check(capi.lib().mom_test_batch_fallback(1 if fallback else 0))
    try:
        X, info = solve_spectrum(gpu_ctx, dms, N, g["wl"], lambda wl: [1.0 + 0j, drude_index(wl)], ang,
                                 precision=prec, pivoting=1 if "tournament" in mode else 0,
                                 lanes=2 if mode.endswith("2lanes") else 1)
    finally:
        capi.check(capi.lib().mom_test_batch_fallback(0))
    if mode == "mixed_fp64_fallback":
        assert [r["fell_back"] for r in info] == [True, False, True]
        assert [r["precision"] for r in info] == ["mixed", "fp64", "mixed"]
    if mode.endswith("2lanes"):
        assert sorted(set(r["lane"] for r in info)) == [0, 1]
    for i, wl in enumerate(g["wl"]):
        om = omega_of(wl)
        x_ref = g["x_%d" % i]
        xerr = _relerr(X[i], x_ref)
        cs = np.array([cs_prtsrf(gpu_ctx, dms[0], N, om, 1.0, X[i][:, k], *(ang[k] * DEGTORAD))
                       for k in range(ang.shape[0])])
        serr = max(np.abs(cs[:, 0] / g["csca_%d" % i] - 1).max(), np.abs(cs[:, 1] / g["cabs_%d" % i] - 1).max())
        rerr = _relerr(rcs(gpu_ctx, dms[0], N, om, 1.0, X[i][:, 0], 37, 72), g["rcs0_%d" % i])
        print("%s wl=%.1f nm: x err %.2e, sigma err %.2e, rcs err %.2e, steps %d, berr %.1e, fp64 fallback %s"
              % (mode, wl * 1e9, xerr, serr, rerr, info[i]["iters"], info[i]["berr"], info[i]["fell_back"]),
              flush=True)
        assert xerr < 1e-8
        assert serr < 1e-8
        assert rerr < 1e-8
        if fallback:
            assert info[i]["fell_back"] or info[i]["precision"] == "fp64"
        elif prec == MOM_PREC_MIXED:
            assert info[i]["fell_back"] or info[i]["berr"] <= 1e-12
    for d in dms:
        d.close()
