"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        L.orc_problem_load.restype = C.c_void_p
        L.orc_problem_load.argtypes = [C.c_char_p, C.c_double, C.c_int, _ip, _ip]
        L.orc_mesh_sizes.argtypes = [C.c_void_p, C.c_int, _ip]
        L.orc_mesh_avelen.restype = C.c_double
        L.orc_mesh_avelen.argtypes = [C.c_void_p, C.c_int]
        L.orc_mesh_export.argtypes = [C.c_void_p, C.c_int] + [_dp, _ip, _ip, _dp, _dp, _dp, _dp, _dp,
                                                              _ip, _ip, _ip, _dp, _ip]
        L.orc_quad_gl13.argtypes = [_dp, _dp]
        L.orc_constants.argtypes = [_dp]
        L.orc_compute_op.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_int, _dp, _dp, _dp]
        L.orc_sysmat.argtypes = [C.c_void_p, C.c_double, _dp, C.c_int, _dp, _dp, _dp]
        L.orc_srcvec.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_int, _dp, _dp, _dp]
        L.orc_cs.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, _dp, C.c_double,
                             C.c_double, C.c_double, C.c_int, _dp, _dp, _dp]
        L.orc_rcs.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, _dp, C.c_int,
                              C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_fill_rows_seconds.restype = C.c_double
        L.orc_fill_rows_seconds.argtypes = [C.c_void_p, C.c_double, _dp, C.c_int, _dp, _dp, C.c_int, C.c_int]
        L.orc_int_scalar.restype = C.c_double
        L.orc_int_scalar.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_int_vector.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_nls_stage1.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp,
                                     C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_rwg_moments.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_epsp.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.orc_nls_stage2.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        _LIB = L
    return _LIB


ETA0 = 3.767303134622e2


def quad_gl13():
    w = np.zeros(13)
    n = np.zeros(39)
    lib().orc_quad_gl13(w, n)
    return w, n.reshape(3, 13).T.copy()  # nodes (13,3)


def constants():
    out = np.zeros(6)
    lib().orc_constants(out)
    return dict(zip(["pi", "eps0", "mu0", "c0", "eta0", "degtorad"], out.tolist()))


class OracleProblem:
    """A problem built by the oracle from a Gmsh file (mesh/sdom/fdom)."""

    def __init__(self, mshfile: str, scale: float, domains):
        nsurf = np.array([len(d) for d in domains], dtype=np.int32)
        ids = np.array([s for d in domains for s in d], dtype=np.int32)
        self.h = lib().orc_problem_load(mshfile.encode(), scale, len(domains), nsurf, ids)
        if not self.h:
            raise RuntimeError("oracle failed to load " + mshfile)
        self.ndom = len(domains)
        self.w, self.nodes = quad_gl13()

    def sizes(self, dom):
        out = np.zeros(3, dtype=np.int32)
        lib().orc_mesh_sizes(self.h, dom, out)
        return tuple(int(x) for x in out)

    @property
    def nedges(self):
        return self.sizes(-1)[2]

    def mesh(self, dom):
        V, F, E = self.sizes(dom)
        a = dict(p=np.zeros(3 * V), fnodes=np.zeros(3 * F, np.int32), fedges=np.zeros(3 * F, np.int32),
                 fn=np.zeros(3 * F), fcp=np.zeros(3 * F), fs=np.zeros(9 * F), fm=np.zeros(9 * F),
                 farea=np.zeros(F), enodes=np.zeros(2 * E, np.int32), ebn=np.zeros(2 * E, np.int32),
                 efaces=np.zeros(2 * E, np.int32), elen=np.zeros(E), eparent=np.zeros(E, np.int32))
        lib().orc_mesh_export(self.h, dom, *a.values())
        shp = dict(p=(V, 3), fnodes=(F, 3), fedges=(F, 3), fn=(F, 3), fcp=(F, 3), fs=(F, 3, 3),
                   fm=(F, 3, 3), enodes=(E, 2), ebn=(E, 2), efaces=(E, 2))
        out = {k: (v.reshape(shp[k]) if k in shp else v) for k, v in a.items()}
        out["avelen"] = lib().orc_mesh_avelen(self.h, dom)
        return out

    def _q(self, quad):
        if quad is None:
            return 13, self.w, np.ascontiguousarray(self.nodes.T).ravel()
        w, nodes = quad
        return len(w), np.ascontiguousarray(w, dtype=np.float64), \
            np.ascontiguousarray(np.asarray(nodes, dtype=np.float64).T).ravel()

    def operator(self, dom, op, omega, ri, quad=None):
        """computeD/H/K on domain `dom` (op 'D','H','K'): E x E complex."""
        E = self.sizes(dom)[2]
        M = np.zeros(2 * E * E)
        Q, w, n = self._q(quad)
        lib().orc_compute_op(self.h, dom, "DHK".index(op), omega, ri.real, ri.imag, Q, w, n, M)
        return M.view(np.complex128).reshape(E, E, order="F")

    def sysmat(self, omega, ri_list, quad=None):
        N = 2 * self.nedges
        A = np.zeros(2 * N * N)
        ri = np.array([[r.real, r.imag] for r in ri_list]).ravel()
        Q, w, n = self._q(quad)
        lib().orc_sysmat(self.h, omega, ri, Q, w, n, A)
        return A.view(np.complex128).reshape(N, N, order="F")

    def srcvec(self, omega, ri, theta, phi, psi, dom=0, quad=None):
        q = np.zeros(4 * self.nedges)
        Q, w, n = self._q(quad)
        lib().orc_srcvec(self.h, dom, omega, ri.real, ri.imag, theta, phi, psi, Q, w, n, q)
        return q.view(np.complex128)

    def cross_sections(self, omega, ri, x, theta, phi, psi, dom=0, quad=None):
        x = np.ascontiguousarray(x, dtype=np.complex128).view(np.float64)
        out = np.zeros(2)
        Q, w, n = self._q(quad)
        lib().orc_cs(self.h, dom, omega, ri.real, ri.imag, x, theta, phi, psi, Q, w, n, out)
        return float(out[0]), float(out[1])  # (csca, cabs)

    def rcs(self, omega, ri, x, ntheta, nphi, dom=0, quad=None):
        x = np.ascontiguousarray(x, dtype=np.complex128).view(np.float64)
        out = np.zeros(ntheta * nphi)
        Q, w, n = self._q(quad)
        lib().orc_rcs(self.h, dom, omega, ri.real, ri.imag, x, ntheta, nphi, Q, w, n, out)
        return out.reshape(ntheta, nphi, order="F")

    # ---- second harmonic (surface chi2), identity group --------------------
    def epsp(self, dom, ri_list):
        """determine_epsp (solver.f90:393-433) for domain dom."""
        F = self.sizes(dom)[1]
        out = np.zeros(2 * F)
        ri = np.array([[complex(r).real, complex(r).imag] for r in ri_list]).ravel()
        if lib().orc_epsp(self.h, dom, ri, out) != 0:
            raise RuntimeError("Could not find epsp!")
        return out.view(np.complex128)

    def rwg_moments(self, dom, quad=None):
        """rwg_moments (rwgf.f90:148-221): E x E Gram matrix <f_m, f_n>."""
        E = self.sizes(dom)[2]
        F = np.zeros(E * E)
        Q, w, n = self._q(quad)
        lib().orc_rwg_moments(self.h, dom, Q, w, n, F)
        return F.reshape(E, E, order="F")

    def nlsurf_coef(self, dom, omega, riff, epsp, x, chi2, quad=None):
        """nlsurf_coef (nlsurf.f90:137-338), identity group: returns
        (src_coef[2E], src_vec[2E]).  The two Gram solves use LAPACK
        (scipy), as the reference's solve_multi_linsys/solve_linsys do."""
        import scipy.linalg as sl

        E = self.sizes(dom)[2]
        Q, w, n = self._q(quad)
        coef, srcJ, sv = np.zeros(2 * E), np.zeros(2 * E), np.zeros(4 * E)
        riff = complex(riff)
        lib().orc_nls_stage1(self.h, dom, omega, riff.real, riff.imag,
                             np.ascontiguousarray(epsp, np.complex128).view(np.float64),
                             np.ascontiguousarray(x, np.complex128).view(np.float64),
                             np.ascontiguousarray(chi2, np.complex128).view(np.float64), Q, w, n, coef, srcJ, sv)
        lu = sl.lu_factor(self.rwg_moments(dom, quad).astype(np.complex128))
        sol = sl.lu_solve(lu, np.stack([coef.view(np.complex128), srcJ.view(np.complex128)], axis=1))
        srcM = np.zeros(2 * E)
        lib().orc_nls_stage2(self.h, dom, Q, w, n, np.ascontiguousarray(sol[:, 0]).view(np.float64), srcM)
        srcM = sl.lu_solve(lu, srcM.view(np.complex128))
        return np.concatenate([srcM, sol[:, 1]]), sv.view(np.complex128).copy()

    def sh_chain(self, omega, ri_list, shri_list, nls, chi2, x, quad=None):
        """solve_batch's centrosymmetric SH branch (solver.f90:143-231):
        nlsurf_coef per nonlinear domain, the SH matrix of sysmat_pmchwt_nls
        (sysmat.f90:112-229) at 2 omega with the SH indices and its products
        with src_coef subtracted from the source, then the SH solution."""
        import scipy.linalg as sl

        Nt = self.nedges
        rhs = np.zeros(2 * Nt, np.complex128)
        out = {}
        for d in range(self.ndom):
            if not nls[d]:
                continue
            ep = self.epsp(d, ri_list)
            sc, sv = self.nlsurf_coef(d, omega, ri_list[d], ep, x, chi2, quad)
            ind = self.mesh(d)["eparent"]
            E = ind.size
            rhs[ind] += sv[:E]
            rhs[ind + Nt] += sv[E:]
            out["epsp_%d" % d], out["src_coef_%d" % d], out["src_vec_%d" % d] = ep, sc, sv
        out["rhs0"] = rhs.copy()
        om2 = 2.0 * omega
        A = self.sysmat(om2, shri_list, quad)
        for d in range(self.ndom):
            if not nls[d]:
                continue
            sc = out["src_coef_%d" % d]
            ind = self.mesh(d)["eparent"]
            E = ind.size
            scM, scJ = sc[:E], sc[E:]
            zsi = complex(shri_list[d]) ** 2 / ETA0 ** 2
            D = self.operator(d, "D", om2, complex(shri_list[d]), quad)
            H = self.operator(d, "H", om2, complex(shri_list[d]), quad)
            K = self.operator(d, "K", om2, complex(shri_list[d]), quad)
            for M in (D, H):
                rhs[ind] -= M @ scJ
                rhs[ind + Nt] -= (M @ scM) * zsi
            rhs[ind] += K @ scM
            rhs[ind + Nt] -= K @ scJ
        out["rhs"] = rhs.copy()
        out["A"] = A
        out["x"] = sl.solve(A, rhs)
        return out

    def fill_rows_seconds(self, omega, ri_list, m0, m1, quad=None):
        """Wall time of the oracle fill (3 operators, every domain) for test
        faces [m0, m1) -- a bounded CPU timing sample."""
        ri = np.array([[complex(r).real, complex(r).imag] for r in ri_list]).ravel()
        Q, w, n = self._q(quad)
        return lib().orc_fill_rows_seconds(self.h, omega, ri, Q, w, n, m0, m1)

    def int_scalar(self, dom, which, f, e, r):
        return lib().orc_int_scalar(self.h, dom, which, f, e, np.ascontiguousarray(r, dtype=np.float64))

    def int_vector(self, dom, which, f, e, r):
        out = np.zeros(3)
        lib().orc_int_vector(self.h, dom, which, f, e, np.ascontiguousarray(r, dtype=np.float64), out)
        return out
