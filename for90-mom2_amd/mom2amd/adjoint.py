"""Adjoint sensitivities of the PMCHWT solution over libmomgpu (no
counterpart in the reference).

For x solving Z x = b and sigma(x) with dsigma = Re(g^T dx):
    dsigma/dp = -Re(lambda^T (dZ/dp) x),   lambda = conj(Z^-H conj(g)).
sigma_abs_eps_gradient applies this to the absorption cross-section of
cs_prtsrf (cs.f90:14-104) and the permittivity of one domain's medium.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .gpu import _quad

__all__ = ["cs_abs_gradient", "pmchwt_deps_apply", "sigma_abs_eps_gradient", "sigma_abs_vertex_gradient"]


def cs_abs_gradient(ctx, dmesh, nedges_total, omega, ri, x, quad=None):
    """(sigma_abs, g) with dsigma_abs = Re(g^T dx) (mom_cs_abs_gradient)."""
    Q, w, n = _quad(quad)
    g = np.zeros(4 * nedges_total)
    ca = C.c_double()
    ri = complex(ri)
    capi.check(capi.lib().mom_cs_abs_gradient(ctx.h, dmesh.h, nedges_total, omega, ri.real, ri.imag,
                                              np.ascontiguousarray(x, np.complex128).view(np.float64), Q, w, n, g,
                                              C.byref(ca)))
    return ca.value, g.view(np.complex128)


def pmchwt_deps_apply(ctx, dmesh, nedges_total, omega, ri, x, quad=None):
    """(dZ/deps) x for eps = ri^2 of the medium of `dmesh`'s domain."""
    Q, w, n = _quad(quad)
    y = np.zeros(4 * nedges_total)
    ri = complex(ri)
    capi.check(capi.lib().mom_pmchwt_deps_apply(ctx.h, dmesh.h, nedges_total, omega, ri.real, ri.imag, Q, w, n,
                                                np.ascontiguousarray(x, np.complex128).view(np.float64), y))
    return y.view(np.complex128)


def sigma_abs_eps_gradient(ctx, system, x, dmeshes, dom, omega, ri_list, quad=None, tol=1e-12):
    """sigma_abs of the solution x of the factored `system` and its
    derivatives with respect to Re and Im of eps = ri^2 of domain `dom`:
    returns (sigma_abs, dsigma/dRe(eps), dsigma/dIm(eps))."""
    N = system.N
    sab, g = cs_abs_gradient(ctx, dmeshes[0], N, omega, ri_list[0], x, quad)
    mu = system.solve_adjoint(np.conj(g), tol=tol)[:, 0]
    lam = np.conj(mu)
    y = pmchwt_deps_apply(ctx, dmeshes[dom], N, omega, ri_list[dom], x, quad)
    t = np.dot(lam, y)
    return sab, -t.real, t.imag


def sigma_abs_vertex_gradient(ctx, system, x, dmeshes, problem, omega, ri_list, angles, quad=None, tol=1e-12):
    """sigma_abs of the solution x of the factored `system` (plane wave of
    `angles` = (theta, phi, psi) [rad]) and its gradient with respect to the
    master mesh's node coordinates: returns (sigma_abs, dsigma/dP[V, 3]).

    dsigma/dP = dsigma/dP|_x + Re(lambda^T dq/dP) - Re(lambda^T dZ/dP x)
    with lambda = Z^-T g: one adjoint solve on the same factors, the
    per-domain contraction of the fill's node derivative
    (mom_pmchwt_vertex_adjoint) and the exterior domain's source and
    cross-section terms (mom_cs_abs_vertex_explicit), summed into the master
    nodes.  The near/far classification is that of the unperturbed mesh."""
    Q, w, n = _quad(quad)
    N = system.N
    sab, g = cs_abs_gradient(ctx, dmeshes[0], N, omega, ri_list[0], x, quad)
    lam = np.conj(system.solve_adjoint(np.conj(g), tol=tol)[:, 0])
    xv = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).view(np.float64)
    lv = np.ascontiguousarray(lam).view(np.float64)
    L = capi.lib()
    grad = np.zeros((problem.master.nnodes, 3))
    for d, (dm, mesh) in enumerate(zip(dmeshes, problem.domains)):
        r = complex(ri_list[d])
        dv = np.zeros(3 * mesh.nnodes)
        capi.check(L.mom_pmchwt_vertex_adjoint(ctx.h, dm.h, N, omega, r.real, r.imag, Q, w, n, xv, lv, dv))
        part = -dv
        if d == 0:
            ex = np.zeros(3 * mesh.nnodes)
            capi.check(L.mom_cs_abs_vertex_explicit(ctx.h, dm.h, N, omega, r.real, r.imag, xv, lv,
                                                    float(angles[0]), float(angles[1]), float(angles[2]),
                                                    Q, w, n, ex))
            part = part + ex
        np.add.at(grad, mesh.node_parent, part.reshape(-1, 3))
    return sab, grad
