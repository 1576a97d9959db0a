"""GPU parity of the fill, RHS and cross-section kernels against the oracle
and the reference's golden vectors.  Tolerance: matrix entries within
1e-10 relative to the largest entry of their PMCHWT block (FP64 path)."""
import numpy as np
import pytest

import oracle
from cases import blockwise_relerr, dimer, sphere
from conftest import golden
from mom2amd import DEGTORAD, DeviceMesh, computeD, computeH, computeK, cs_prtsrf, srcvec, sysmat_pmchwt

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _dev(ctx, case):
    return [DeviceMesh(ctx, d) for d in case["problem"].domains]


@pytest.mark.parametrize("s", [1, 2])
def test_operators_match_oracle(gpu_ctx, msh_dir, s):
    case = sphere(msh_dir, s)
    op = oracle.OracleProblem(case["path"], 1e-9, case["domains"])
    dms = _dev(gpu_ctx, case)
    for d in range(2):
        for name, fn in (("D", computeD), ("H", computeH), ("K", computeK)):
            M = fn(gpu_ctx, dms[d], case["omega"], case["ri"][d])
            R = op.operator(d, name, case["omega"], case["ri"][d])
            err = np.abs(M - R).max() / np.abs(R).max()
            print("s=%d dom=%d %s err=%.3e" % (s, d, name, err), flush=True)
            assert err < TOL, (s, d, name, err)


def test_sysmat_golden_sphere_s1(gpu_ctx, msh_dir):
    g = golden("sphere_s1")
    case = sphere(msh_dir, 1)
    dms = _dev(gpu_ctx, case)
    A = sysmat_pmchwt(gpu_ctx, dms, case["ri"], case["problem"].nedges, case["omega"])
    err = blockwise_relerr(A, g["A"])
    print("sphere s=1 sysmat vs reference: %.3e" % err, flush=True)
    assert err < TOL


@pytest.mark.parametrize("name", ["sphere_s2", "dimer_s1"])
def test_sysmat_golden_sampled(gpu_ctx, msh_dir, name):
    g = golden(name)
    case = sphere(msh_dir, 2) if name == "sphere_s2" else dimer(msh_dir, 1)
    dms = _dev(gpu_ctx, case)
    A = sysmat_pmchwt(gpu_ctx, dms, case["ri"], case["problem"].nedges, case["omega"])
    N = A.shape[0] // 2
    # block norms pin the whole matrix, sampled entries pin positions
    fro = np.array([[np.linalg.norm(A[i * N:(i + 1) * N, j * N:(j + 1) * N]) for j in range(2)] for i in range(2)])
    assert np.abs(fro / g["block_fro"] - 1).max() < TOL
    ref = g["A_v"]
    got = A[g["A_i"], g["A_j"]]
    scale = np.array([[np.abs(A[i * N:(i + 1) * N, j * N:(j + 1) * N]).max() for j in range(2)] for i in range(2)])
    s = scale[g["A_i"] // N, g["A_j"] // N]
    err = (np.abs(got - ref) / s).max()
    print("%s sampled sysmat vs reference: %.3e" % (name, err), flush=True)
    assert err < TOL


@pytest.mark.parametrize("name", ["sphere_s1", "sphere_s2", "dimer_s1"])
def test_srcvec_and_cs_golden(gpu_ctx, msh_dir, name):
    g = golden(name)
    case = {"sphere_s1": lambda: sphere(msh_dir, 1), "sphere_s2": lambda: sphere(msh_dir, 2),
            "dimer_s1": lambda: dimer(msh_dir, 1)}[name]()
    dm = DeviceMesh(gpu_ctx, case["problem"].domains[0])
    ang = np.array(g["angles_deg"]) * DEGTORAD
    q = srcvec(gpu_ctx, dm, case["problem"].nedges, case["omega"], case["ri"][0], ang)[:, 0]
    err_q = np.abs(q - g["q"]).max() / np.abs(g["q"]).max()
    cs, ca = cs_prtsrf(gpu_ctx, dm, case["problem"].nedges, case["omega"], case["ri"][0], g["x"], *ang)
    print("%s q err %.3e csca %.3e cabs %.3e" % (name, err_q, cs / g["csca"] - 1, ca / g["cabs"] - 1), flush=True)
    assert err_q < 1e-12
    assert abs(cs / g["csca"] - 1) < 1e-12 and abs(ca / g["cabs"] - 1) < 1e-12


def test_near_pair_count_matches_oracle_rule(gpu_ctx, msh_dir):
    case = sphere(msh_dir, 2)
    d = case["problem"].domains[0]
    cp = d.fcp
    diff = cp[:, None, :] - cp[None, :, :]
    ds = diff[..., 0] * diff[..., 0] + diff[..., 1] * diff[..., 1] + diff[..., 2] * diff[..., 2]
    thr = d.avelen * 3
    assert DeviceMesh(gpu_ctx, d).near_pairs == int((ds < thr * thr).sum())


def test_phase_range_is_checked_before_launch(gpu_ctx, msh_dir):
    """The point-pair loop reduces k R modulo pi/2 exactly only while
    Re(k) times the mesh extent stays below 2^19 rad; beyond that the fill
    refuses with an argument error (checked on the host, before any kernel
    runs) instead of returning inaccurate entries."""
    from mom2amd.capi import MomError

    case = sphere(msh_dir, 1)
    dm = DeviceMesh(gpu_ctx, case["problem"].domains[0])
    with pytest.raises(MomError) as e:
        computeD(gpu_ctx, dm, case["omega"] * 1e7, case["ri"][0])
    assert e.value.status < 0 and "phase range" in str(e.value)
    # the context stays usable
    M = computeD(gpu_ctx, dm, case["omega"], case["ri"][0])
    assert np.isfinite(M).all()


@pytest.mark.parametrize("name", ["sphere_s1", "sphere_s2", "dimer_s1"])
def test_rcs_golden(gpu_ctx, msh_dir, name):
    """rcs on the 37 x 72 grid against the reference's own values for the
    same solution x (ffields.f90:203-265)."""
    from mom2amd import rcs

    g = golden(name)
    case = {"sphere_s1": lambda: sphere(msh_dir, 1), "sphere_s2": lambda: sphere(msh_dir, 2),
            "dimer_s1": lambda: dimer(msh_dir, 1)}[name]()
    dm = DeviceMesh(gpu_ctx, case["problem"].domains[0])
    r = rcs(gpu_ctx, dm, case["problem"].nedges, case["omega"], case["ri"][0], g["x"], 37, 72)
    err = np.abs(r - g["rcs"]).max() / np.abs(g["rcs"]).max()
    print("%s rcs err %.3e" % (name, err), flush=True)
    assert err < 1e-12


def test_far_fields_relations(gpu_ctx, msh_dir):
    """Closed-form relations of the far field that a slip in far_fields'
    assembly would break: H = (ri/eta0) r x E (ht = -(ri/eta0) ep, hp =
    (ri/eta0) et), the exp(ikr)/r dependence on the distance, the filter
    (the filtered values for x, y and z sum to the unfiltered ones, since
    v_theta, v_phi are orthonormal) and the grid of rcs_ext."""
    from mom2amd import ETA0, far_fields, rcs_ext

    g = golden("sphere_s1")
    case = sphere(msh_dir, 1)
    dm = DeviceMesh(gpu_ctx, case["problem"].domains[0])
    N, om, ri = case["problem"].nedges, case["omega"], case["ri"][0]
    rng = np.random.default_rng(3)
    th, ph = rng.uniform(0, np.pi, 50), rng.uniform(0, 2 * np.pi, 50)
    et, ep, ht, hp = far_fields(gpu_ctx, dm, N, om, ri, g["x"], 1.0, th, ph)
    s = np.abs(et).max()
    assert np.abs(ht + ri / ETA0 * ep).max() < 1e-12 * s * abs(ri / ETA0)
    assert np.abs(hp - ri / ETA0 * et).max() < 1e-12 * s * abs(ri / ETA0)
    # distance: the phase k r is about 1e7 rad at r = 1 m, so the prefactor
    # carries an absolute phase rounding of |k r| u (a few 1e-9 relative)
    k = ri * om / 2.997924580003e8
    et2, ep2, _, _ = far_fields(gpu_ctx, dm, N, om, ri, g["x"], 3.0, th, ph)
    fac = np.exp(1j * k * 2.0) / 3.0
    tol = 1e-15 * abs(k) * 3.0
    assert np.abs(et2 - fac * et).max() < tol * s
    assert np.abs(ep2 - fac * ep).max() < tol * s
    full = rcs_ext(gpu_ctx, dm, N, om, ri, g["x"], 9, 12)
    parts = [rcs_ext(gpu_ctx, dm, N, om, ri, g["x"], 9, 12, filter=True, fdir=e) for e in np.eye(3)]
    assert np.abs(sum(parts) - full).max() < 1e-12 * full.max()
    # the grid: theta = pi i/8, phi = 2 pi j/11 (pi as the reference's float32-rounded value)
    from mom2amd import PI
    T, P = np.meshgrid(PI * np.arange(9) / 8, 2 * PI * np.arange(12) / 11, indexing="ij")
    et3, ep3, _, _ = far_fields(gpu_ctx, dm, N, om, ri, g["x"], 1.0, T, P)
    assert np.abs(np.abs(et3) ** 2 + np.abs(ep3) ** 2 - full).max() < 1e-12 * full.max()


def test_mesh_from_native_setup(gpu_ctx, msh_dir):
    """mom_problem_mesh_create uploads the set-up's domains directly; the
    meshes it makes give the same near lists and matrix as meshes uploaded
    from the copied-out arrays."""
    import ctypes as C

    from mom2amd import capi, meshgen

    case = dimer(msh_dir, 1)
    L = capi.lib()
    nodes, faces, ids = meshgen.read_gmsh(case["path"])
    doms = case["domains"]
    ptr = np.cumsum([0] + [len(s) for s in doms]).astype(np.int32)
    di = np.array([s for d in doms for s in d], np.int32)
    h = C.c_void_p()
    capi.check(L.mom_problem_build(len(nodes), np.ascontiguousarray(nodes, np.float64).ravel(), len(faces),
                                   np.ascontiguousarray(faces, np.int32).ravel(),
                                   np.ascontiguousarray(ids, np.int32), 1e-9, len(doms), ptr, di, C.byref(h)))

    class Handle:
        def __init__(self, d):
            self.h = C.c_void_p()
            capi.check(L.mom_problem_mesh_create(gpu_ctx.h, h, d, C.byref(self.h)))

    try:
        native = [Handle(d) for d in range(len(doms))]
        with pytest.raises(capi.MomError):
            Handle(len(doms))
    finally:
        L.mom_problem_destroy(h)
    try:
        copied = _dev(gpu_ctx, case)
        for a, b in zip(native, copied):
            assert L.mom_mesh_near_pairs(a.h) == b.near_pairs
        N = case["problem"].nedges
        A = sysmat_pmchwt(gpu_ctx, native, case["ri"], N, case["omega"])
        B = sysmat_pmchwt(gpu_ctx, copied, case["ri"], N, case["omega"])
        err = np.abs(A - B).max() / np.abs(B).max()
        print("native set-up vs copied arrays: %.3e" % err, flush=True)
        assert err < 1e-14
    finally:
        for a in native:
            L.mom_mesh_destroy(a.h)


def test_operators_match_reference_one_by_one(gpu_ctx, msh_dir):
    """computeD/H/K of each domain against the reference's own dump of each
    operator (tests/golden/ops_s1.npz), so that a slip that cancels in D+H or
    in the sum of K over the domains cannot hide."""
    g = golden("ops_s1")
    case = sphere(msh_dir, 1)
    dms = _dev(gpu_ctx, case)
    for d in range(2):
        for name, fn in (("D", computeD), ("H", computeH), ("K", computeK)):
            R = g["op_%d_%s" % (d, name)]
            M = fn(gpu_ctx, dms[d], case["omega"], case["ri"][d])
            err = np.abs(M - R).max() / np.abs(R).max()
            print("dom=%d %s vs reference: %.3e" % (d, name, err), flush=True)
            assert err < TOL, (d, name, err)
