"""Reference-named entry points of the hot path, running in libmomgpu.

  sysmat_pmchwt        sysmat.f90:18-105      PMCHWT matrix (host copy)
  computeD/H/K         sysmat.f90:272-519     operator matrices of one domain
  srcvec               source.f90:136-247     plane-wave right-hand sides
  solve_multi_linsys   linalg.f90:159-180     LU + solve (complex128)
  cs_prtsrf            cs.f90:14-104          cross-sections
  System               solver.f90:109-136     device-resident fill/factor/solve
Errors raise capi.MomError, the counterpart of the reference's WRITE + STOP.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref

import numpy as np

from . import capi
from .constants import tri_quad_data


def _quad(quad):
    w, nodes = tri_quad_data() if quad is None else quad
    w = np.ascontiguousarray(w, dtype=np.float64)
    # quad_data%nodes is (Q,3) column-major
    n = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64).T).ravel()
    return len(w), w, n


_live_contexts = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    # release device memory while the HIP runtime is still up, rather than
    # from finalizers that may run after it has been torn down
    for ctx in list(_live_contexts):
        try:
            ctx.close()
        except Exception:
            pass


class Context:
    """One HIP stream on one device (mom_ctx)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        self.device = device
        self._children = weakref.WeakSet()
        capi.check(capi.lib().mom_ctx_create(device, C.byref(self.h)))
        _live_contexts.add(self)

    def _adopt(self, child):
        self._children.add(child)

    def fill_tuning(self):
        """(source faces per wave chosen by the fill's autotuner, {candidate: ms})."""
        spw, ms = C.c_int32(), np.zeros(3)
        capi.check(capi.lib().mom_ctx_fill_tuning(self.h, C.byref(spw), ms))
        return spw.value, {2: ms[0], 4: ms[1], 8: ms[2]}

    def fill_profile(self, enable=-1):
        """Event timing of the fill kernels: returns the totals accumulated
        so far (pair_ms, pair_launches, pair_points, assemble_ms,
        assemble_launches, near_ms), then switches timing on (1) or off (0)
        and resets them; -1 only reads."""
        out = np.zeros(6)
        capi.check(capi.lib().mom_ctx_fill_profile(self.h, int(enable), out))
        keys = ("pair_ms", "pair_launches", "pair_points", "assemble_ms", "assemble_launches", "near_ms")
        return dict(zip(keys, out.tolist()))

    def close(self):
        """Destroy the context; meshes and systems created on it are
        destroyed first (they hold device memory and the context's stream)."""
        if self.h:
            for child in list(self._children):
                child.close()
            capi.lib().mom_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMesh:
    """A domain mesh uploaded to the device (mom_mesh)."""

    def __init__(self, ctx: Context, dm):
        self.ctx = ctx
        self.host = dm  # the host mesh, for copies on other contexts
        self.frozen = False  # near/far classification taken from another mesh (classify_like)
        self.nedges = dm.nedges
        self.nfaces = dm.nfaces
        self.h = C.c_void_p()
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).ravel()
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).ravel()
        capi.check(capi.lib().mom_mesh_create(
            ctx.h, dm.nnodes, f64(dm.p), dm.nfaces, i32(dm.fnodes), i32(dm.fedges), f64(dm.fn), f64(dm.fcp),
            f64(dm.fs), f64(dm.fm), f64(dm.farea), dm.nedges, i32(dm.efaces), i32(dm.ebn), f64(dm.elen),
            i32(dm.eparent), float(dm.avelen), C.byref(self.h)))
        ctx._adopt(self)

    @property
    def near_pairs(self):
        return int(capi.lib().mom_mesh_near_pairs(self.h))

    def classify_like(self, other: "DeviceMesh"):
        """Classify near face pairs by `other`'s centroids and average edge
        length (same face count), freezing the near/far split of a perturbed
        mesh to that of the unperturbed one (mom_mesh_classify_like)."""
        capi.check(capi.lib().mom_mesh_classify_like(self.h, other.h))
        self.frozen = True

    def close(self):
        if self.h:
            capi.lib().mom_mesh_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _operator(ctx, dmesh, op, omega, ri, quad):
    Q, w, n = _quad(quad)
    E = dmesh.nedges
    M = np.zeros(2 * E * E)
    capi.check(capi.lib().mom_compute_operator(ctx.h, dmesh.h, op.encode(), omega, complex(ri).real,
                                               complex(ri).imag, Q, w, n, M, E))
    return M.view(np.complex128).reshape(E, E, order="F")


def computeD(ctx, dmesh, omega, ri, quad=None):
    return _operator(ctx, dmesh, "D", omega, ri, quad)


def computeH(ctx, dmesh, omega, ri, quad=None):
    return _operator(ctx, dmesh, "H", omega, ri, quad)


def computeK(ctx, dmesh, omega, ri, quad=None):
    return _operator(ctx, dmesh, "K", omega, ri, quad)


def _domain_handles(dmeshes):
    arr = (C.c_void_p * len(dmeshes))(*[d.h.value for d in dmeshes])
    return C.cast(arr, C.POINTER(C.c_void_p)), arr


def _ri_array(ri_list):
    return np.ascontiguousarray(np.array([[complex(r).real, complex(r).imag] for r in ri_list]).ravel())


def sysmat_pmchwt(ctx, dmeshes, ri_list, nedges_total, omega, quad=None):
    """PMCHWT matrix A (2N x 2N complex128) of the domains (sysmat.f90:18-105)."""
    Q, w, n = _quad(quad)
    N2 = 2 * nedges_total
    A = np.zeros(2 * N2 * N2)
    ptr, keep = _domain_handles(dmeshes)
    capi.check(capi.lib().mom_sysmat_pmchwt(ctx.h, len(dmeshes), ptr, _ri_array(ri_list), nedges_total,
                                            omega, Q, w, n, A, N2))
    return A.view(np.complex128).reshape(N2, N2, order="F")


def srcvec(ctx, dmesh, nedges_total, omega, ri, angles, quad=None):
    """Plane-wave excitation vectors q (2N x nsrc) (source.f90:136-247).
    angles: (nsrc, 3) array of (theta, phi, psi) in radians."""
    Q, w, n = _quad(quad)
    ang = np.ascontiguousarray(np.atleast_2d(np.asarray(angles, dtype=np.float64))).ravel()
    nsrc = ang.size // 3
    N2 = 2 * nedges_total
    q = np.zeros(2 * N2 * nsrc)
    capi.check(capi.lib().mom_srcvec_pw(ctx.h, dmesh.h, nedges_total, omega, complex(ri).real, complex(ri).imag,
                                        nsrc, ang, Q, w, n, q, N2))
    return q.view(np.complex128).reshape(N2, nsrc, order="F")


def solve_multi_linsys(ctx, A, B):
    """LU-factor A in place and solve A X = B (linalg.f90:159-180).
    Returns (LU, X, ipiv) with ipiv 1-based as LAPACK's."""
    A = np.asfortranarray(A, dtype=np.complex128).copy(order="F")
    B2 = np.asfortranarray(np.atleast_2d(B.T).T, dtype=np.complex128).copy(order="F")
    n = A.shape[0]
    nrhs = B2.shape[1]
    ipiv = np.zeros(n, dtype=np.int32)
    Av = A.reshape(-1, order="F").view(np.float64)
    Bv = B2.reshape(-1, order="F").view(np.float64)
    st = capi.lib().mom_solve_multi_linsys(ctx.h, n, Av, n, nrhs, Bv, n, ipiv.ctypes.data)
    if st > 0:
        raise capi.MomError(st, "Linear system solution failed (singular matrix, INFO=%d)" % st)
    capi.check(st)
    LU = Av.view(np.complex128).reshape(n, n, order="F")
    X = Bv.view(np.complex128).reshape(n, nrhs, order="F")
    return LU, X, ipiv


def cs_prtsrf(ctx, dmesh, nedges_total, omega, ri, x, theta, phi, psi, quad=None):
    """(csca, cabs) of solution x for a plane wave (cs.f90:14-104)."""
    Q, w, n = _quad(quad)
    xv = np.ascontiguousarray(x, dtype=np.complex128).view(np.float64)
    cs, ca = C.c_double(), C.c_double()
    capi.check(capi.lib().mom_cs_prtsrf(ctx.h, dmesh.h, nedges_total, omega, complex(ri).real, complex(ri).imag,
                                        xv, theta, phi, psi, Q, w, n, C.byref(cs), C.byref(ca)))
    return cs.value, ca.value


class System:
    """Device-resident PMCHWT system (mom_system): fill, factor, solve."""

    def __init__(self, ctx, nedges_total, precision=capi.MOM_PREC_MIXED, order=None):
        self.ctx = ctx
        self.h = C.c_void_p()
        if order is None:
            self.N = nedges_total
            self.n = 2 * nedges_total
            capi.check(capi.lib().mom_system_create(ctx.h, nedges_total, precision, C.byref(self.h)))
        else:
            # one representation of a symmetry group (mom_system_create_order)
            self.N = 0
            self.n = int(order)
            capi.check(capi.lib().mom_system_create_order(ctx.h, self.n, precision, C.byref(self.h)))
        ctx._adopt(self)

    @classmethod
    def of_order(cls, ctx, order, precision=capi.MOM_PREC_MIXED):
        return cls(ctx, 0, precision, order=order)

    def fill(self, dmeshes, ri_list, omega, quad=None):
        Q, w, n = _quad(quad)
        ptr, keep = _domain_handles(dmeshes)
        capi.check(capi.lib().mom_system_fill(self.h, len(dmeshes), ptr, _ri_array(ri_list), omega, Q, w, n))

    def get_matrix(self):
        N2 = self.n
        A = np.zeros(2 * N2 * N2)
        capi.check(capi.lib().mom_system_get_matrix(self.h, A, N2))
        return A.view(np.complex128).reshape(N2, N2, order="F")

    def set_matrix(self, A):
        N2 = self.n
        Av = np.asfortranarray(A, dtype=np.complex128).reshape(-1, order="F").view(np.float64)
        capi.check(capi.lib().mom_system_set_matrix(self.h, np.ascontiguousarray(Av), N2))

    def set_precision(self, precision):
        """Precision of the next factorisation of a system created mixed:
        capi.MOM_PREC_FP64 or capi.MOM_PREC_MIXED (mom_system_set_precision)."""
        capi.check(capi.lib().mom_system_set_precision(self.h, int(precision)))

    def set_cooperative(self, allow):
        """Partial-pivoting panels by cooperative launches (1, default) or
        one ordinary launch per column (0; mom_system_set_cooperative)."""
        capi.check(capi.lib().mom_system_set_cooperative(self.h, int(allow)))

    def set_pivoting(self, mode):
        """Pivoting of the complex64 factorisation (mixed path): 0 partial as
        LAPACK (default), 1 tournament (mom_system_set_pivoting)."""
        capi.check(capi.lib().mom_system_set_pivoting(self.h, int(mode)))

    def factor(self):
        st = capi.lib().mom_system_factor(self.h)
        if st > 0:
            raise capi.MomError(st, "zero pivot in column %d" % st)
        capi.check(st)

    def solve(self, B, tol=1e-12, maxit=30):
        N2 = self.n
        B2 = np.asfortranarray(np.atleast_2d(np.asarray(B).T).T, dtype=np.complex128).copy(order="F")
        nrhs = B2.shape[1]
        Bv = np.ascontiguousarray(B2.reshape(-1, order="F").view(np.float64))
        it, be = C.c_int32(), C.c_double()
        st = capi.lib().mom_system_solve(self.h, nrhs, Bv, N2, tol, maxit, C.byref(it), C.byref(be))
        self.iters, self.berr = it.value, be.value
        capi.check(st)
        return Bv.view(np.complex128).reshape(N2, nrhs, order="F")

    def solve_adjoint(self, B, tol=1e-12, maxit=30):
        """Z^H X = B on the same factorisation (mom_system_solve_adjoint)."""
        N2 = self.n
        B2 = np.asfortranarray(np.atleast_2d(np.asarray(B).T).T, dtype=np.complex128).copy(order="F")
        nrhs = B2.shape[1]
        Bv = np.ascontiguousarray(B2.reshape(-1, order="F").view(np.float64))
        it, be = C.c_int32(), C.c_double()
        st = capi.lib().mom_system_solve_adjoint(self.h, nrhs, Bv, N2, tol, maxit, C.byref(it), C.byref(be))
        self.iters, self.berr = it.value, be.value
        capi.check(st)
        return Bv.view(np.complex128).reshape(N2, nrhs, order="F")

    def solve_pw(self, dmesh, omega, ri, angles, tol=1e-12, maxit=30, quad=None):
        Q, w, n = _quad(quad)
        ang = np.ascontiguousarray(np.atleast_2d(np.asarray(angles, dtype=np.float64))).ravel()
        nsrc = ang.size // 3
        N2 = 2 * self.N
        X = np.zeros(2 * N2 * nsrc)
        it, be = C.c_int32(), C.c_double()
        st = capi.lib().mom_system_solve_pw(self.h, dmesh.h, omega, complex(ri).real, complex(ri).imag, nsrc, ang,
                                            Q, w, n, tol, maxit, X, N2, C.byref(it), C.byref(be))
        self.iters, self.berr = it.value, be.value
        capi.check(st)
        return X.view(np.complex128).reshape(N2, nsrc, order="F")

    def stats(self):
        out = np.zeros(23)
        capi.check(capi.lib().mom_system_stats(self.h, out))
        return dict(fill_s=out[0], factor_s=out[1], solve_s=out[2], gemm_s=out[3], gemm_launches=out[4],
                    gemm_flops=out[5], panel_s=out[6], trsm_s=out[7], laswp_s=out[8], n_minv=int(out[9]),
                    n_matvec=int(out[10]), berr_history=[float(v) for v in out[11:19] if v >= 0],
                    gemm_narrow_s=out[19], gemm_narrow_launches=int(out[20]), gemm_narrow_flops=out[21],
                    fell_back_to_fp64=bool(out[22]))

    def close(self):
        if self.h:
            capi.lib().mom_system_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def far_fields(ctx, dmesh, nedges_total, omega, ri, x, r, theta, phi, quad=None):
    """far_fields (ffields.f90:16-122) for arrays of directions theta, phi
    [rad] at distance r: returns (et, ep, ht, hp), complex arrays of the
    directions' shape."""
    Q, w, n = _quad(quad)
    th = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
    ph = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
    if th.shape != ph.shape:
        raise ValueError("theta and phi must have the same number of directions")
    xv = np.ascontiguousarray(x, dtype=np.complex128).view(np.float64)
    out = np.zeros(8 * th.size)
    ri = complex(ri)
    capi.check(capi.lib().mom_far_fields(ctx.h, dmesh.h, nedges_total, omega, ri.real, ri.imag, xv, r, th.size,
                                         th, ph, Q, w, n, out))
    f = out.view(np.complex128).reshape(-1, 4)
    shape = np.shape(theta)
    return tuple(f[:, c].reshape(shape) for c in range(4))


def rcs_ext(ctx, dmesh, nedges_total, omega, ri, x, ntheta, nphi, filter=False, fdir=(0.0, 0.0, 0.0), quad=None):
    """rcs_ext (ffields.f90:203-249): ntheta x nphi array of |E_far|^2 at
    r = 1, optionally through the polarisation filter fdir."""
    Q, w, n = _quad(quad)
    xv = np.ascontiguousarray(x, dtype=np.complex128).view(np.float64)
    out = np.zeros(ntheta * nphi)
    fd = np.ascontiguousarray(fdir, dtype=np.float64)
    ri = complex(ri)
    capi.check(capi.lib().mom_rcs_ext(ctx.h, dmesh.h, nedges_total, omega, ri.real, ri.imag, xv, ntheta, nphi,
                                      1 if filter else 0, fd.ctypes.data, Q, w, n, out))
    return out.reshape(ntheta, nphi, order="F")


def rcs(ctx, dmesh, nedges_total, omega, ri, x, ntheta, nphi, quad=None):
    """rcs (ffields.f90:251-265): rcs_ext without the filter."""
    return rcs_ext(ctx, dmesh, nedges_total, omega, ri, x, ntheta, nphi, quad=quad)
