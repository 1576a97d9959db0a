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
from .constants import EPS0
from .gpu import System, _domain_handles, _quad, _ri_array

__all__ = ["determine_epsp", "nlsurf_coef", "sysmat_pmchwt_nls", "solve_sh"]


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
