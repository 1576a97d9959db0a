"""Second harmonic from a surface chi2 (centrosymmetric media), identity
group, free space: the centrosymmetric branch of solve_batch
(solver.f90:143-231) over libmomgpu.

    for each nonlinear domain d:
        epsp = determine_epsp(problem, d, ri)          (solver.f90:393-433, host)
        src_coef_d, src_vec_d = nlsurf_coef(...)       (nlsurf.f90:137-338)
        nlx[parent] += src_vec_d[:E]; nlx[parent + N] += src_vec_d[E:]
    A, nlx = sysmat_pmchwt_nls(...)                    (sysmat.f90:112-229)
    nlx = A^-1 nlx                                     (solve_systems)
and the SH far field is rcs(..., 2 omega, shri of domain 1, nlx)
(interface.f90:1316-1325).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .constants import C0, EPS0
from .gpu import System, _domain_handles, _quad, _ri_array

__all__ = ["determine_epsp", "nlsurf_coef", "sysmat_pmchwt_nls", "solve_sh", "far_fields_adjoint",
           "nlsurf_coef_adjoint", "domain_block_apply", "power_grid", "edge_mask", "sh_power", "sh_power_translation_gradient"]


def determine_epsp(problem, dom, ri_list):
    """Permittivity just outside each face of domain `dom` (solver.f90:393-433):
    eps0 ri^2 of the first other domain holding the same master face, with
    the fundamental indices ri_list."""
    dm = problem.domains[dom]
    owner = {}
    for d, other in enumerate(problem.domains):
        if d == dom:
            continue
        for f in other.fparent.tolist():
            owner.setdefault(f, d)
    out = np.empty(dm.nfaces, dtype=np.complex128)
    for i, f in enumerate(dm.fparent.tolist()):
        d = owner.get(f)
        if d is None:
            raise RuntimeError("Could not find epsp!")
        r = complex(ri_list[d])
        out[i] = (r * r) * EPS0
    return out


def nlsurf_coef(ctx, dmesh, nedges_total, omega, ri, epsp, x, chi2, quad=None):
    """nlsurf_coef (nlsurf.f90:137-338) of one domain: (src_coef[2E],
    src_vec[2E]).  omega, ri: fundamental angular frequency and the domain's
    fundamental index; x: fundamental solution [2N]; chi2 = (nnn, ntt, ttn)."""
    Q, w, n = _quad(quad)
    E = dmesh.nedges
    sc, sv = np.zeros(4 * E), np.zeros(4 * E)
    ri = complex(ri)
    capi.check(capi.lib().mom_nlsurf_coef(
        ctx.h, dmesh.h, nedges_total, omega, ri.real, ri.imag,
        np.ascontiguousarray(epsp, np.complex128).view(np.float64),
        np.ascontiguousarray(x, np.complex128).view(np.float64),
        np.ascontiguousarray(np.asarray(chi2, np.complex128)).view(np.float64), Q, w, n, sc, sv))
    return sc.view(np.complex128), sv.view(np.complex128)


def _coef_array(src_coef, N2, ndom):
    sc = np.asarray(src_coef, dtype=np.complex128)
    if sc.ndim == 2:
        sc = sc[:, :, None]
    if sc.shape[0] != N2 or sc.shape[1] != ndom:
        raise ValueError("src_coef must be (2N, ndom[, nsrc])")
    return np.ascontiguousarray(sc.reshape(-1, order="F")).view(np.float64), sc.shape[2]


def sysmat_pmchwt_nls(ctx, dmeshes, shri_list, nedges_total, omega, src_coef, src, quad=None):
    """sysmat_pmchwt_nls (sysmat.f90:112-229): returns (A, src) with A the SH
    PMCHWT matrix at the SH angular frequency omega and src (2N[, nsrc])
    minus the blocks' products with src_coef (2N, ndom[, nsrc])."""
    Q, w, n = _quad(quad)
    N2 = 2 * nedges_total
    scv, nsrc = _coef_array(src_coef, N2, len(dmeshes))
    b = np.asfortranarray(np.asarray(src, np.complex128).reshape(N2, -1)).copy(order="F")
    bv = np.ascontiguousarray(b.reshape(-1, order="F").view(np.float64))
    A = np.zeros(2 * N2 * N2)
    ptr, keep = _domain_handles(dmeshes)
    capi.check(capi.lib().mom_sysmat_pmchwt_nls(ctx.h, len(dmeshes), ptr, _ri_array(shri_list), nedges_total, omega,
                                                Q, w, n, nsrc, scv, N2, bv, N2, A, N2))
    out = bv.view(np.complex128).reshape(N2, nsrc, order="F")
    return A.view(np.complex128).reshape(N2, N2, order="F"), (out[:, 0] if np.ndim(src) == 1 else out)


def _fill_nls(system, dmeshes, shri_list, omega, src_coef, src, quad=None):
    Q, w, n = _quad(quad)
    N2 = 2 * system.N
    scv, nsrc = _coef_array(src_coef, N2, len(dmeshes))
    b = np.asfortranarray(np.asarray(src, np.complex128).reshape(N2, -1)).copy(order="F")
    bv = np.ascontiguousarray(b.reshape(-1, order="F").view(np.float64))
    ptr, keep = _domain_handles(dmeshes)
    capi.check(capi.lib().mom_system_fill_nls(system.h, len(dmeshes), ptr, _ri_array(shri_list), omega, Q, w, n,
                                              nsrc, scv, N2, bv, N2))
    out = bv.view(np.complex128).reshape(N2, nsrc, order="F")
    return out[:, 0] if np.ndim(src) == 1 else out


System.fill_nls = _fill_nls


def solve_sh(ctx, problem, dmeshes, omega, ri_list, shri_list, nls, chi2, x, precision=capi.MOM_PREC_MIXED,
             tol=1e-12, quad=None, system=None):
    """The centrosymmetric SH branch of solve_batch for one source x (the
    fundamental solution): returns dict(src_coef (2N, ndom), rhs (SH source
    after sysmat_pmchwt_nls), x (SH solution), iters, berr)."""
    N = problem.nedges
    ndom = len(dmeshes)
    nlx = np.zeros(2 * N, np.complex128)
    src_coef = np.zeros((2 * N, ndom), np.complex128)
    for d in range(ndom):
        if not nls[d]:
            continue
        ep = determine_epsp(problem, d, ri_list)
        sc, sv = nlsurf_coef(ctx, dmeshes[d], N, omega, ri_list[d], ep, x, chi2, quad)
        E = dmeshes[d].nedges
        ind = problem.domains[d].eparent
        src_coef[:2 * E, d] = sc
        nlx[ind] += sv[:E]
        nlx[ind + N] += sv[E:]
    own = system is None
    S = System(ctx, N, precision) if own else system
    try:
        rhs = S.fill_nls(dmeshes, shri_list, 2.0 * omega, src_coef, nlx, quad)
        S.factor()
        xs = S.solve(rhs, tol=tol)[:, 0]
        return dict(src_coef=src_coef, rhs=rhs, x=xs, iters=S.iters, berr=S.berr)
    finally:
        if own:
            S.close()


# ---------------------------------------------------------------------------
# Adjoint of the SH chain (no counterpart in the reference).  Convention of
# mom_pmchwt_vertex_adjoint: an observable s with ds = Re(g^T dx).

def far_fields_adjoint(ctx, dmesh, nedges_total, omega, ri, r, theta, phi, v, quad=None):
    """g[2N] = L^T v for the far-field map out = L x of far_fields (same
    arguments): v (ndir, 4) complex adjoint values of (et, ep, ht, hp)."""
    Q, w, n = _quad(quad)
    th = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
    ph = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
    vv = np.ascontiguousarray(np.asarray(v, np.complex128).reshape(th.size, 4)).view(np.float64).reshape(-1)
    g = np.zeros(4 * nedges_total)
    ri = complex(ri)
    capi.check(capi.lib().mom_far_fields_adjoint(ctx.h, dmesh.h, nedges_total, omega, ri.real, ri.imag, r, th.size,
                                                 th, ph, vv, Q, w, n, g))
    return g.view(np.complex128)


def nlsurf_coef_adjoint(ctx, dmesh, nedges_total, omega, ri, epsp, x, chi2, src_coef_bar, src_vec_bar, quad=None):
    """xbar[2N] = J^T (src_coef_bar, src_vec_bar) for the Jacobian J of
    nlsurf_coef at the fundamental solution x (holomorphic, non-conjugated)."""
    Q, w, n = _quad(quad)
    xb = np.zeros(4 * nedges_total)
    ri = complex(ri)
    c128 = lambda a: np.ascontiguousarray(a, np.complex128).view(np.float64)
    capi.check(capi.lib().mom_nlsurf_coef_adjoint(
        ctx.h, dmesh.h, nedges_total, omega, ri.real, ri.imag, c128(epsp), c128(x),
        c128(np.asarray(chi2, np.complex128)), Q, w, n, c128(src_coef_bar), c128(src_vec_bar), xb))
    return xb.view(np.complex128)


def domain_block_apply(ctx, dmesh, nedges_total, omega, ri, x, trans=False, quad=None):
    """Z_d x (or Z_d^T x) for one domain's PMCHWT blocks in global numbering."""
    Q, w, n = _quad(quad)
    y = np.zeros(4 * nedges_total)
    ri = complex(ri)
    capi.check(capi.lib().mom_domain_block_apply(ctx.h, dmesh.h, nedges_total, omega, ri.real, ri.imag,
                                                 1 if trans else 0, Q, w, n,
                                                 np.ascontiguousarray(x, np.complex128).view(np.float64), y))
    return y.view(np.complex128)


def power_grid(ntheta=37, nphi=72):
    """(theta, phi, W): the rcs_ext grid (theta = pi i/(ntheta-1), phi =
    2 pi j/(nphi-1), both ends included, flattened with i fastest) and
    solid-angle weights of the trapezoid rule (the phi = 2 pi column repeats
    phi = 0 and has weight 0), so that sum W |E|^2 integrates |E_far|^2 over
    the directions."""
    t = np.pi * np.arange(ntheta) / (ntheta - 1)
    p = 2 * np.pi * np.arange(nphi) / (nphi - 1)
    th, ph = np.meshgrid(t, p, indexing="ij")
    wt = np.full(ntheta, np.pi / (ntheta - 1))
    wt[0] = wt[-1] = 0.5 * np.pi / (ntheta - 1)
    wp = np.full(nphi, 2 * np.pi / (nphi - 1))
    wp[-1] = 0.0
    W = (wt * np.sin(t))[:, None] * wp[None, :]
    f = lambda a: a.reshape(-1, order="F")
    return f(th), f(ph), f(W)


def edge_mask(problem, node_mask):
    """Global edges (of the master mesh) whose two nodes lie in node_mask."""
    en = problem.master.enodes
    return np.asarray(node_mask)[en[:, 0]] & np.asarray(node_mask)[en[:, 1]]


def _sh_forward(ctx, problem, dmeshes, omega, ri_list, shri_list, nls, chi2, ang, S1, S2, tol, th, ph, W, quad):
    from .gpu import far_fields

    N = problem.nedges
    S1.fill(dmeshes, ri_list, omega, quad)
    S1.factor()
    x1 = S1.solve_pw(dmeshes[0], omega, ri_list[0], ang[None, :], tol=tol, quad=quad)[:, 0]
    sh = solve_sh(ctx, problem, dmeshes, omega, ri_list, shri_list, nls, chi2, x1, tol=tol, quad=quad, system=S2)
    x2 = sh["x"]
    et, ep, _, _ = far_fields(ctx, dmeshes[0], N, 2.0 * omega, shri_list[0], x2, 1.0, th, ph, quad)
    P = float(np.sum(W * (np.abs(et) ** 2 + np.abs(ep) ** 2)))
    return P, x1, x2, et, ep


def sh_power(ctx, problem, dmeshes, omega, ri_list, shri_list, nls, chi2, angles, precision=capi.MOM_PREC_FP64,
             tol=1e-12, ntheta=37, nphi=72, quad=None):
    """The SH power radiated into the rcs_ext grid (power_grid weights) for
    the plane wave of `angles`: the forward part of
    sh_power_translation_gradient."""
    N = problem.nedges
    th, ph, W = power_grid(ntheta, nphi)
    S1 = System(ctx, N, precision)
    S2 = System(ctx, N, precision)
    try:
        return _sh_forward(ctx, problem, dmeshes, omega, ri_list, shri_list, nls, chi2,
                           np.asarray(angles, dtype=np.float64).reshape(3), S1, S2, tol, th, ph, W, quad)[0]
    finally:
        S1.close()
        S2.close()


def sh_power_translation_gradient(ctx, problem, dmeshes, omega, ri_list, shri_list, nls, chi2, angles, bodies,
                                  precision=capi.MOM_PREC_FP64, tol=1e-12, ntheta=37, nphi=72, quad=None,
                                  pivoting=0):
    """SH power radiated into the rcs_ext grid, P = sum_d W_d (|et|^2 +
    |ep|^2) at 2 omega in domain 1 (power_grid weights), and its derivative
    with respect to a parameter that translates bodies rigidly: `bodies` is
    a list of (node_mask over the master nodes, velocity dr/dp (3,) [m per
    unit of p]), e.g. the gap of a dimer: (left sphere, (-1/2, 0, 0)),
    (right sphere, (1/2, 0, 0)).

    By the adjoint of the chain Z1 x1 = q1, (src_coef, src_vec) =
    nlsurf_coef(x1) per nonlinear domain, Z2 x2 = rhs2(x1), P(x2):
        lambda2 = Z2^-T g2,  g2 = dP/dx2 (mom_far_fields_adjoint),
        mu = sum_d J_d^T (lambda2 restricted, -Z_d^T lambda2 restricted)
             (mom_nlsurf_coef_adjoint, mom_domain_block_apply),
        lambda1 = Z1^-T mu,
        dP/dp = dP/dp|_x2 + Re(lambda1^T dq1/dp) - Re(lambda1^T dZ1/dp x1)
                - Re(lambda2^T dZ2/dp x2)
    with the matrix terms from mom_pmchwt_vertex_adjoint of every domain
    and both frequencies, the source term from mom_cs_abs_vertex_explicit
    at x = 0, and dP/dp|_x2 from the far fields of each body's currents
    (a translation t multiplies them by exp(-i k dir . t)).  nlsurf_coef and
    the nonlinear domains' own blocks do not change under a rigid motion of
    their body, so the parameter must move every nonlinear domain rigidly
    (exact up to the quadrature table's 3e-8 barycentric defect).  The
    near/far classification is held at that of the unperturbed meshes.
    pivoting: of the complex64 factorisations in the mixed path
    (System.set_pivoting).  Returns dict(P, dP, terms, x1, x2, grid)."""
    from .gpu import far_fields

    Q, w, n = _quad(quad)
    N = problem.nedges
    L = capi.lib()
    ang = np.asarray(angles, dtype=np.float64).reshape(3)
    om2 = 2.0 * omega
    th, ph, W = power_grid(ntheta, nphi)
    S1 = System(ctx, N, precision)
    S2 = System(ctx, N, precision)
    S1.set_pivoting(pivoting)
    S2.set_pivoting(pivoting)
    try:
        P, x1, x2, et, ep = _sh_forward(ctx, problem, dmeshes, omega, ri_list, shri_list, nls, chi2, ang, S1, S2, tol,
                                        th, ph, W, quad)
        v = np.zeros((th.size, 4), np.complex128)
        v[:, 0] = 2 * W * np.conj(et)
        v[:, 1] = 2 * W * np.conj(ep)
        g2 = far_fields_adjoint(ctx, dmeshes[0], N, om2, shri_list[0], 1.0, th, ph, v, quad)
        lam2 = np.conj(S2.solve_adjoint(np.conj(g2), tol=tol)[:, 0])
        mu = np.zeros(2 * N, np.complex128)
        for d in range(len(dmeshes)):
            if not nls[d]:
                continue
            ind = problem.domains[d].eparent
            E = ind.size
            svb = np.concatenate([lam2[ind], lam2[ind + N]])
            zt = -domain_block_apply(ctx, dmeshes[d], N, om2, shri_list[d], lam2, trans=True, quad=quad)
            vb = np.concatenate([zt[ind], zt[ind + N]])  # adjoint of (srcJ, srcM)
            scb = np.concatenate([vb[E:], vb[:E]])  # src_coef = (srcM, srcJ)
            epsp = determine_epsp(problem, d, ri_list)
            mu += nlsurf_coef_adjoint(ctx, dmeshes[d], N, omega, ri_list[d], epsp, x1, chi2, scb, svb, quad)
        lam1 = np.conj(S1.solve_adjoint(np.conj(mu), tol=tol)[:, 0])
    finally:
        S1.close()
        S2.close()
    nv = problem.master.nnodes
    vel = np.zeros((nv, 3))
    for mask, t in bodies:
        vel[np.asarray(mask)] = np.asarray(t, dtype=np.float64)
    c128 = lambda a: np.ascontiguousarray(a, np.complex128).view(np.float64)

    def matrix_term(x, lam, om, ris):
        G = np.zeros((nv, 3))
        for d, (dm, mesh) in enumerate(zip(dmeshes, problem.domains)):
            r = complex(ris[d])
            dv = np.zeros(3 * mesh.nnodes)
            capi.check(L.mom_pmchwt_vertex_adjoint(ctx.h, dm.h, N, om, r.real, r.imag, Q, w, n, c128(x), c128(lam),
                                                   dv))
            np.add.at(G, mesh.node_parent, dv.reshape(-1, 3))
        return -float(np.sum(G * vel))

    t_z1 = matrix_term(x1, lam1, omega, ri_list)
    t_z2 = matrix_term(x2, lam2, 2.0 * omega, shri_list)
    mesh0 = problem.domains[0]
    ex = np.zeros(3 * mesh0.nnodes)
    r0 = complex(ri_list[0])
    capi.check(L.mom_cs_abs_vertex_explicit(ctx.h, dmeshes[0].h, N, omega, r0.real, r0.imag,
                                            np.zeros(4 * N), c128(lam1), float(ang[0]), float(ang[1]),
                                            float(ang[2]), Q, w, n, ex))
    Gq = np.zeros((nv, 3))
    np.add.at(Gq, mesh0.node_parent, ex.reshape(-1, 3))
    t_q1 = float(np.sum(Gq * vel))
    # explicit far-field term: each body's currents radiate with the phase of its translation
    k2 = complex(shri_list[0]) * 2.0 * omega / C0
    rhat = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    det = np.zeros_like(et)
    dep = np.zeros_like(ep)
    for mask, t in bodies:
        em = edge_mask(problem, mask)
        xb = np.where(np.concatenate([em, em]), x2, 0.0)
        etb, epb, _, _ = far_fields(ctx, dmeshes[0], N, om2, shri_list[0], xb, 1.0, th, ph, quad)
        fac = -1j * k2 * (rhat @ np.asarray(t, dtype=np.float64))
        det += fac * etb
        dep += fac * epb
    t_far = float(np.sum(W * 2 * np.real(np.conj(et) * det + np.conj(ep) * dep)))
    dP = t_far + t_q1 + t_z1 + t_z2
    return dict(P=P, dP=dP, terms=dict(far=t_far, source=t_q1, fundamental_matrix=t_z1, sh_matrix=t_z2),
                x1=x1, x2=x2, grid=(th, ph, W))
