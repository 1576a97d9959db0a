"""Vertex-coordinate adjoint of sigma_abs (mom_pmchwt_vertex_adjoint,
mom_cs_abs_vertex_explicit, mom2amd.sigma_abs_vertex_gradient).

Pinned by
- rigid translation: the gradient summed over the nodes is the derivative
  of sigma_abs when every node moves by the same vector.  For an exact
  rule that sum would vanish (a plane wave only changes phase); the
  reference's gl13 table has barycentric coordinates that sum to 1 +- 3e-8
  (float32 values, quad.f90:277-315), so its quadrature points move by
  (1 + eps) t and the discrete sigma_abs changes by about 2e-8 per nm.  The
  test compares the node sum with central differences of the whole FP64
  solve under translation (h = 0.1 nm), which checks that the three terms
  of the adjoint balance to that level;
- central differences of the whole FP64 solve (set-up, fill, LU, solve,
  cs_prtsrf) in single node coordinates, h = 1e-3 nm on a 50 nm sphere
  (the near/far classification does not change at that size), and on all
  three coordinates of 32 seeded nodes of a radially perturbed sphere
  (icosphere s = 2, R = 60 nm, l <= 6 harmonics of 3 nm) with the
  classification of every perturbed mesh frozen to the unperturbed one
  (mom_mesh_classify_like), as the gradient assumes.
Tolerances: the node sum against the translation derivative <= 1e-9 of the
summed magnitudes (1e-7 for the mixed solve, whose x and lambda carry
forward errors near 1e-8 that the cancelling sum exposes); the derivatives
<= 2e-6 of the largest gradient component."""
import numpy as np
import pytest

from cases import N_IN_SPHERE
from mom2amd import (DEGTORAD, DeviceMesh, System, build_problem, cs_prtsrf, meshgen, omega_of,
                     sigma_abs_vertex_gradient)
from mom2amd.capi import MOM_PREC_FP64, MOM_PREC_MIXED

pytestmark = pytest.mark.gpu

ANG = np.array([40.0, 25.0, 30.0]) * DEGTORAD
OMEGA = omega_of(600e-9)
RI = [1.0 + 0j, complex(N_IN_SPHERE)]


def _solve(ctx, nodes, faces, prec):
    P = build_problem(nodes, faces, np.full(faces.shape[0], 15), 1e-9, [[15], [-15]])
    dms = [DeviceMesh(ctx, d) for d in P.domains]
    S = System(ctx, P.nedges, prec)
    S.fill(dms, RI, OMEGA)
    S.factor()
    x = S.solve_pw(dms[0], OMEGA, RI[0], ANG)[:, 0]
    return P, dms, S, x


def _sigma_abs(ctx, nodes, faces):
    P, dms, S, x = _solve(ctx, nodes, faces, MOM_PREC_FP64)
    try:
        return cs_prtsrf(ctx, dms[0], P.nedges, OMEGA, RI[0], x, *ANG)[1]
    finally:
        S.close()
        for d in dms:
            d.close()


def _gradient(ctx, nodes, faces, prec):
    P, dms, S, x = _solve(ctx, nodes, faces, prec)
    try:
        return sigma_abs_vertex_gradient(ctx, S, x, dms, P, OMEGA, RI, ANG)
    finally:
        S.close()
        for d in dms:
            d.close()


@pytest.mark.parametrize("prec", [MOM_PREC_FP64, MOM_PREC_MIXED], ids=["fp64", "mixed"])
def test_vertex_gradient_sums_to_the_translation_derivative(gpu_ctx, prec):
    nodes, faces = meshgen.icosphere(2, 50.0)
    sab, grad = _gradient(gpu_ctx, nodes, faces, prec)
    tot, mag = grad.sum(axis=0), np.abs(grad).sum(axis=0)
    h = 0.1
    dt = []
    for c in range(3):
        vals = []
        for sgn in (1.0, -1.0):
            p = nodes.copy()
            p[:, c] += sgn * h
            vals.append(_sigma_abs(gpu_ctx, p, faces))
        dt.append((vals[0] - vals[1]) / (2 * h * 1e-9))
    dt = np.array(dt)
    print("prec %d: sigma_abs %.6e, node sum %s, translation derivative %s, magnitude %s"
          % (prec, sab, tot, dt, mag), flush=True)
    assert np.all(np.abs(tot - dt) <= (1e-9 if prec == MOM_PREC_FP64 else 1e-7) * mag)
    # the translation derivative is the quadrature table's 3e-8 defect, not a missing term
    assert np.all(np.abs(dt) <= 1e-5 * mag)


def test_vertex_gradient_matches_finite_differences(gpu_ctx):
    nodes, faces = meshgen.icosphere(2, 50.0)
    sab, grad = _gradient(gpu_ctx, nodes, faces, MOM_PREC_FP64)
    scale = np.abs(grad).max()
    h = 1e-3  # nm; the gradient is per metre
    rng = np.random.default_rng(5)
    picks = [(int(v), int(c)) for v, c in zip(rng.integers(0, nodes.shape[0], 5), rng.integers(0, 3, 5))]
    picks.append((int(np.argmax(np.abs(grad).max(axis=1))), int(np.argmax(np.abs(grad).max(axis=0)))))
    for v, c in picks:
        vals = []
        for sgn in (1.0, -1.0):
            p = nodes.copy()
            p[v, c] += sgn * h
            vals.append(_sigma_abs(gpu_ctx, p, faces))
        fd = (vals[0] - vals[1]) / (2 * h * 1e-9)
        print("node %d coord %d: adjoint %.10e fd %.10e (rel to max %.1e)"
              % (v, c, grad[v, c], fd, abs(grad[v, c] - fd) / scale), flush=True)
        assert abs(grad[v, c] - fd) <= 2e-6 * scale


def test_far_pair_closed_form_matches_differences(gpu_ctx):
    """The closed-form derivatives (far pairs, the near pairs' smoothed point
    loop and, by forward mode, their int.f90 analytic parts) against
    fourth-order differences of the same pair sums
    (mom_test_vertex_fd), per domain and node coordinate: <= 1e-8 of the
    largest component."""
    import ctypes as C  # noqa: F401
    from mom2amd import capi
    from mom2amd.gpu import _quad

    nodes, faces = meshgen.icosphere(2, 50.0)
    P, dms, S, x = _solve(gpu_ctx, nodes, faces, MOM_PREC_FP64)
    L = capi.lib()
    try:
        rng = np.random.default_rng(2)
        lam = rng.standard_normal(x.size) + 1j * rng.standard_normal(x.size)
        Q, w, n = _quad(None)
        xv = np.ascontiguousarray(x).view(np.float64)
        lv = np.ascontiguousarray(lam).view(np.float64)
        for d, (dm, mesh) in enumerate(zip(dms, P.domains)):
            out = []
            for fd in (0, 1):
                L.mom_test_vertex_fd(fd)
                dv = np.zeros(3 * mesh.nnodes)
                capi.check(L.mom_pmchwt_vertex_adjoint(gpu_ctx.h, dm.h, P.nedges, OMEGA, RI[d].real, RI[d].imag, Q,
                                                       w, n, xv, lv, dv))
                out.append(dv)
            err = np.abs(out[0] - out[1]).max() / np.abs(out[1]).max()
            print("domain %d: closed form vs differences %.2e" % (d, err), flush=True)
            assert err < 1e-8
    finally:
        L.mom_test_vertex_fd(0)
        S.close()
        for dm in dms:
            dm.close()


def test_perturbed_sphere_gradient_matches_frozen_differences(gpu_ctx):
    nodes, faces = meshgen.perturbed_sphere(2, 60.0, 3.0)
    base = build_problem(nodes, faces, np.full(faces.shape[0], 15), 1e-9, [[15], [-15]])
    bdms = [DeviceMesh(gpu_ctx, d) for d in base.domains]
    nnear = [d.near_pairs for d in bdms]
    sab, grad = _gradient(gpu_ctx, nodes, faces, MOM_PREC_FP64)
    scale = np.abs(grad).max()
    h = 1e-3  # nm
    rng = np.random.default_rng(20251204)
    verts = rng.choice(nodes.shape[0], 32, replace=False)
    worst = 0.0
    try:
        for v in verts:
            for c in range(3):
                vals = []
                for sgn in (1.0, -1.0):
                    p = nodes.copy()
                    p[v, c] += sgn * h
                    P = build_problem(p, faces, np.full(faces.shape[0], 15), 1e-9, [[15], [-15]])
                    dms = [DeviceMesh(gpu_ctx, d) for d in P.domains]
                    for dm, bm in zip(dms, bdms):
                        dm.classify_like(bm)
                    assert [d.near_pairs for d in dms] == nnear
                    S = System(gpu_ctx, P.nedges, MOM_PREC_FP64)
                    try:
                        S.fill(dms, RI, OMEGA)
                        S.factor()
                        x = S.solve_pw(dms[0], OMEGA, RI[0], ANG)[:, 0]
                        vals.append(cs_prtsrf(gpu_ctx, dms[0], P.nedges, OMEGA, RI[0], x, *ANG)[1])
                    finally:
                        S.close()
                        for d in dms:
                            d.close()
                fd = (vals[0] - vals[1]) / (2 * h * 1e-9)
                err = abs(grad[v, c] - fd) / scale
                worst = max(worst, err)
                assert err <= 2e-6, (int(v), c, grad[v, c], fd)
    finally:
        for d in bdms:
            d.close()
    print("perturbed s=2 sphere, 32 nodes x 3 coordinates, frozen classification: sigma_abs %.6e, "
          "max |adjoint - fd| / max |grad| = %.2e" % (sab, worst), flush=True)


def test_classify_like_keeps_or_freezes_the_near_list(gpu_ctx):
    """mom_mesh_classify_like: a mesh classified like itself fills the same
    bits; a mesh scaled by 1.5 classified like the original keeps the
    original's near-pair count; meshes of different face counts are refused."""
    from mom2amd import capi

    nodes, faces = meshgen.icosphere(1, 50.0)
    P = build_problem(nodes, faces, np.full(faces.shape[0], 15), 1e-9, [[15], [-15]])
    dms = [DeviceMesh(gpu_ctx, d) for d in P.domains]
    S = System(gpu_ctx, P.nedges, MOM_PREC_FP64)
    S.fill(dms, RI, OMEGA)
    A0 = S.get_matrix()
    for d in dms:
        d.classify_like(d)
    S.fill(dms, RI, OMEGA)
    assert np.array_equal(S.get_matrix(), A0)
    S.close()
    P2 = build_problem(nodes * 1.5, faces, np.full(faces.shape[0], 15), 1e-9, [[15], [-15]])
    d2 = DeviceMesh(gpu_ctx, P2.domains[0])
    d2.classify_like(dms[0])
    assert d2.near_pairs == dms[0].near_pairs
    n3, f3 = meshgen.icosphere(2, 50.0)
    P3 = build_problem(n3, f3, np.full(f3.shape[0], 15), 1e-9, [[15], [-15]])
    d3 = DeviceMesh(gpu_ctx, P3.domains[0])
    with pytest.raises(capi.MomError):
        d3.classify_like(dms[0])
    for d in dms + [d2, d3]:
        d.close()


def test_vertex_gradient_is_bitwise_reproducible(gpu_ctx):
    """Every sum of the node-coordinate adjoint has a fixed order (pair
    tiles, near-list entries and faces, then the faces into the nodes on the
    host): the gradient of the same solve computed twice is bitwise equal."""
    nodes, faces = meshgen.icosphere(2, 50.0)
    P, dms, S, x = _solve(gpu_ctx, nodes, faces, MOM_PREC_FP64)
    try:
        _, g1 = sigma_abs_vertex_gradient(gpu_ctx, S, x, dms, P, OMEGA, RI, ANG)
        _, g2 = sigma_abs_vertex_gradient(gpu_ctx, S, x, dms, P, OMEGA, RI, ANG)
    finally:
        S.close()
        for d in dms:
            d.close()
    g1, g2 = np.asarray(g1), np.asarray(g2)
    assert np.abs(g1).max() > 0
    assert np.array_equal(g1, g2)
