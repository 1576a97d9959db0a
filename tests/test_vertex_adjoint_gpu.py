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
  (the near/far classification does not change at that size).
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
    """The closed-form derivatives (far pairs, and the near pairs' smoothed
    point loop) against fourth-order differences of the same pair sums
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
