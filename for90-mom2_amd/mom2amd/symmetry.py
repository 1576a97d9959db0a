"""Symmetry decomposition: the `symm` command with the mirror planes x = 0,
y = 0, z = 0, the rotations about the z axis (rn) and their products, solved
one representation at a time on the device.

The reference builds the group (read_symm, interface.f90:160-210), fills one
matrix per representation (sysmat_pmchwt, sysmat.f90:18-105), projects the
source on the representations (srcvec with src_field_frags,
source.f90:136-315), drops the unknowns that the mirror planes force to zero
and solves each reduced system (solve_systems, solver.f90:327-357, with
edge_bc / reduce_system / expand_solution, bc.f90:28-232); cross-sections
and far fields sum the images of every representation's currents
(cs.f90:14-104, ffields.f90:16-122).  Here:

- mom_symm_group / mom_symm_edge_bc give the group table and each
  representation's removed unknowns;
- one device system per representation (mom_system_create_order) receives
  its reduced matrix from a single mom_system_fill_symm, which integrates
  every image once and writes all representations;
- the sources are projected on the device (mom_srcvec_pw_symm), reduced and
  solved per representation, and expanded with zeros at the removed
  unknowns;
- mom_cs_prtsrf_symm / mom_rcs_ext_symm post-process the expanded
  solutions.

An odd reduced order is padded by one decoupled unknown (unit diagonal,
zero source), which keeps every device matrix of even order.  Rotation
groups (group_rz, symmetry.f90:105-156) have complex field actions and are
for bodies that no rotation cut crosses, such as n copies of a particle
around the z axis: the reference sets no edge coupling for rz1/rz2 cuts.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .gpu import System, _domain_handles, _quad, _ri_array

__all__ = ["PLANES", "group", "edge_bc", "SymmetricSystems", "srcvec_symm", "cs_prtsrf_symm", "rcs_symm"]

# generators of the `symm` line (interface.f90:173-188): the mirror planes,
# and "rN" for group_rz(N) (code 100 + N)
PLANES = {"mxp": 1, "myp": 2, "mzp": 3}


def _code(g):
    if isinstance(g, str):
        if g in PLANES:
            return PLANES[g]
        if g[:1] == "r" and g[1:].isdigit():
            return 100 + int(g[1:])
        raise ValueError("unknown symmetry generator %r" % g)
    return int(g)


def _planes(gens):
    return np.ascontiguousarray([_code(g) for g in gens], dtype=np.int32)


def group(gens):
    """The group of the generators gens (e.g. ["mxp", "myp"] or ["r3",
    "mzp"]): dict(n, ef[n,n] complex (element, representation), detj[n],
    J[n,3,3], jdiag[n,3] (the diagonal of J), id[n])."""
    p = _planes(gens)
    n = C.c_int32()
    ef, detj, jm, gid = np.zeros(128), np.zeros(8), np.zeros(72), np.zeros(8, np.int32)
    capi.check(capi.lib().mom_symm_group_c(len(p), p, C.byref(n), ef, detj, jm, gid))
    n = n.value
    J = jm[:9 * n].reshape(n, 3, 3)
    return dict(n=n, ef=ef[:2 * n * n].view(np.complex128).reshape(n, n), detj=detj[:n], J=J,
                jdiag=np.array([np.diag(j) for j in J]), id=gid[:n])


def edge_bc(bnd, gens, rep):
    """(id[2N], dim): -1 marks the unknowns that representation rep removes."""
    p = _planes(gens)
    bnd = np.ascontiguousarray(bnd, dtype=np.int32)
    ids = np.zeros(2 * bnd.size, np.int32)
    dim = C.c_int64()
    capi.check(capi.lib().mom_symm_edge_bc(bnd.size, bnd, len(p), p, int(rep), ids, C.byref(dim)))
    return ids, dim.value


def srcvec_symm(ctx, dmesh, nedges_total, omega, ri, gens, angles, quad=None):
    """q[2N, nga, nsrc]: the plane waves projected on each representation."""
    Q, w, n = _quad(quad)
    p = _planes(gens)
    nga = group(gens)["n"]
    ang = np.ascontiguousarray(np.atleast_2d(np.asarray(angles, dtype=np.float64))).ravel()
    nsrc = ang.size // 3
    N2 = 2 * nedges_total
    q = np.zeros(2 * N2 * nga * nsrc)
    ri = complex(ri)
    capi.check(capi.lib().mom_srcvec_pw_symm(ctx.h, dmesh.h, nedges_total, omega, ri.real, ri.imag, len(p), p, nsrc,
                                             ang, Q, w, n, q, N2))
    return q.view(np.complex128).reshape(N2, nga, nsrc, order="F")


def cs_prtsrf_symm(ctx, dmesh, nedges_total, omega, ri, gens, x, theta, phi, psi, quad=None):
    """(csca, cabs) of the expanded solutions x[2N, nga] (cs_prtsrf with the group)."""
    Q, w, n = _quad(quad)
    p = _planes(gens)
    xv = np.ascontiguousarray(np.asfortranarray(x, dtype=np.complex128).reshape(-1, order="F").view(np.float64))
    cs, ca = C.c_double(), C.c_double()
    ri = complex(ri)
    capi.check(capi.lib().mom_cs_prtsrf_symm(ctx.h, dmesh.h, nedges_total, omega, ri.real, ri.imag, len(p), p, xv,
                                             theta, phi, psi, Q, w, n, C.byref(cs), C.byref(ca)))
    return cs.value, ca.value


def rcs_symm(ctx, dmesh, nedges_total, omega, ri, gens, x, ntheta, nphi, quad=None):
    """rcs[ntheta, nphi] of the expanded solutions x[2N, nga] (rcs with the group)."""
    Q, w, n = _quad(quad)
    p = _planes(gens)
    xv = np.ascontiguousarray(np.asfortranarray(x, dtype=np.complex128).reshape(-1, order="F").view(np.float64))
    out = np.zeros(ntheta * nphi)
    ri = complex(ri)
    capi.check(capi.lib().mom_rcs_ext_symm(ctx.h, dmesh.h, nedges_total, omega, ri.real, ri.imag, len(p), p, xv,
                                           ntheta, nphi, 0, None, Q, w, n, out))
    return out.reshape(ntheta, nphi, order="F")


class SymmetricSystems:
    """One device system per representation of the group `gens` for the
    problem whose master edges have boundary classes `bnd`."""

    def __init__(self, ctx, nedges_total, bnd, gens, precision=capi.MOM_PREC_MIXED):
        self.ctx = ctx
        self.N = nedges_total
        self.gens = list(gens)
        self.bnd = np.ascontiguousarray(bnd, dtype=np.int32)
        if self.bnd.size != nedges_total:
            raise ValueError("one boundary class per master edge is needed")
        self.G = group(gens)
        self.keep, self.dims, self.systems = [], [], []
        try:
            for r in range(self.G["n"]):
                ids, dim = edge_bc(self.bnd, gens, r)
                self.keep.append(np.flatnonzero(ids != -1))
                self.dims.append(dim)
                self.systems.append(System.of_order(ctx, dim + (dim & 1), precision))
        except Exception:
            self.close()
            raise

    @property
    def nga(self):
        return self.G["n"]

    def fill(self, dmeshes, ri_list, omega, quad=None):
        Q, w, n = _quad(quad)
        ptr, keep = _domain_handles(dmeshes)
        hs = (C.c_void_p * self.nga)(*[S.h.value for S in self.systems])
        p = _planes(self.gens)
        capi.check(capi.lib().mom_system_fill_symm(self.nga, C.cast(hs, C.POINTER(C.c_void_p)), len(dmeshes), ptr,
                                                   _ri_array(ri_list), self.N, self.bnd, len(p), p, omega, Q, w, n))

    def factor(self):
        for S in self.systems:
            S.factor()

    def reduced_matrix(self, r):
        """Representation r's reduced matrix (without the padding unknown)."""
        d = self.dims[r]
        return self.systems[r].get_matrix()[:d, :d]

    def solve(self, q, tol=1e-12, maxit=30):
        """x[2N, nga, nsrc] from the unreduced projected sources q[2N, nga,
        nsrc]: reduce, solve and expand per representation."""
        q = np.asarray(q, dtype=np.complex128)
        if q.ndim == 2:
            q = q[:, :, None]
        x = np.zeros_like(q)
        self.iters, self.berr = [], []
        for r, S in enumerate(self.systems):
            k, d = self.keep[r], self.dims[r]
            B = np.zeros((S.n, q.shape[2]), np.complex128)
            B[:d] = q[k, r, :]
            X = S.solve(B, tol, maxit)
            x[k, r, :] = X[:d]
            self.iters.append(S.iters)
            self.berr.append(S.berr)
        return x

    def stats(self):
        return [S.stats() for S in self.systems]

    def close(self):
        for S in getattr(self, "systems", []):
            S.close()
        self.systems = []
