"""Batched spectrum with many plane waves: the wavelength loop of
solve_batch (solver.f90:75-140) -- per wavelength the fill (109), the
plane-wave sources (123-127) and solve_systems on all of them (136) -- with
one device-resident System reused across the wavelengths and every source
of a wavelength solved as one block on the reused factors.

Also the workload's inputs: wavelengths as `wlrg` spaces them
(read_wlrange, interface.f90:368-389, linterp aux.f90:58-63), the interior
index from a Drude permittivity, and the seeded set of plane waves.
"""
from __future__ import annotations

import threading
import time

import numpy as np

from . import capi
from .constants import DEGTORAD, omega_of
from .gpu import Context, DeviceMesh, System

HC_EVNM = 1239.841984  # h c [eV nm]


def wavelengths(nwl, wl1=400e-9, wl2=900e-9):
    """`wlrg nwl wl1 wl2`: wl1 + (n-1)/(nwl-1) (wl2 - wl1), n = 1..nwl."""
    if nwl == 1:
        return np.array([wl1])
    t = np.arange(nwl, dtype=np.float64) / float(nwl - 1)
    return wl1 + t * (wl2 - wl1)


def drude_index(wl, eps_inf=9.5, wp_ev=9.0, gamma_ev=0.07):
    """Refractive index sqrt(eps) with Im >= 0 of eps = eps_inf - wp^2 / (w (w + i gamma)),
    w = h c / wl in eV (one Drude pole)."""
    w = HC_EVNM / (wl * 1e9)
    eps = eps_inf - wp_ev ** 2 / (w * (w + 1j * gamma_ev))
    n = np.sqrt(complex(eps))
    return n if n.imag >= 0 else -n


def plane_waves_deg(ndir=128, seed=20251203, psis=(0.0, 90.0)):
    """(theta, phi, psi) in degrees, psi-major: every psi over ndir directions
    uniform on the sphere (cos theta ~ U(-1, 1), phi ~ U(0, 2 pi) from
    numpy's default_rng(seed)).  As `ssrc ... pw theta phi psi` takes them."""
    rng = np.random.default_rng(seed)
    ct = rng.uniform(-1.0, 1.0, ndir)
    ph = rng.uniform(0.0, 2 * np.pi, ndir)
    th = np.degrees(np.arccos(ct))
    ph = np.degrees(ph)
    return np.array([[t, p, ps] for ps in psis for t, p in zip(th, ph)])


def solve_spectrum(ctx, dmeshes, nedges_total, wls, ri_of, angles_deg, precision=capi.MOM_PREC_MIXED, tol=1e-12,
                   maxit=30, keep=None, pivoting=0, lanes=1):
    """For every wavelength wl in wls: fill with ri_of(wl) (one index per
    domain), factor, and solve all plane waves (angles in degrees, converted
    with the reference's degtorad) in one block in domain 1's medium.
    Returns (X, info): X[i] the 2N x nsrc solutions of wavelength i (kept
    only for i in `keep` when given), info per wavelength (in wls order):
    fill/factor/solve seconds, refinement steps, backward error, whether the
    block solve fell back to FP64 factors, the precision the wavelength was
    factored in and the lane that solved it.  pivoting: of the complex64
    factorisation in the mixed path (0 partial, 1 tournament;
    System.set_pivoting).

    lanes > 1 solves that many wavelengths at once: each lane is a host
    thread with its own context (stream pair, scratch) and system on the same
    device, taking the next wavelength in order; the library's calls release
    the interpreter lock, so one lane's fill or refinement runs beside
    another's factorisation (whose late steps leave most of the device
    idle).  The extra contexts and their mesh copies live for the call.
    Lane systems factor with ordinary launches only (System.set_cooperative):
    the complex128 factorisations a lane may need run their partial-pivoting
    panels one launch per column.

    Mixed path: a wavelength whose nearest finished shorter wavelength (among
    the last `lanes`) fell back to FP64 factors is factored directly in FP64
    (System.set_precision) -- the complex64 factors stop converging over a
    band of wavelengths where the interior permittivity is near zero, and an
    attempt there costs a complex64 factorisation and some refinement before
    the fallback.  The wavelength after a direct one tries complex64 again."""
    ang = np.asarray(angles_deg, dtype=np.float64) * DEGTORAD
    wls = list(wls)
    nl = max(1, min(int(lanes), len(wls)))
    X, info = {}, [None] * len(wls)
    lock = threading.Lock()
    state = {"next": 0, "error": None}
    status = {}  # wavelength index -> "mixed", "fell_back" or "fp64"

    def direct_fp64(i):
        if precision != capi.MOM_PREC_MIXED:
            return False
        done = [j for j in status if i - nl <= j < i]
        return bool(done) and status[max(done)] == "fell_back"

    def lane(ctx_l, dms_l, li):
        S = System(ctx_l, nedges_total, precision)
        try:
            S.set_pivoting(pivoting)
            if nl > 1:
                # lanes run in worker threads: no cooperative launches from them
                S.set_cooperative(0)
            while True:
                with lock:
                    i = state["next"]
                    if i >= len(wls) or state["error"] is not None:
                        return
                    state["next"] = i + 1
                    fp64 = direct_fp64(i)
                wl = wls[i]
                om = omega_of(wl)
                ri = ri_of(wl)
                if precision == capi.MOM_PREC_MIXED:
                    S.set_precision(capi.MOM_PREC_FP64 if fp64 else capi.MOM_PREC_MIXED)
                t0 = time.perf_counter()
                S.fill(dms_l, ri, om)
                t1 = time.perf_counter()
                S.factor()
                t2 = time.perf_counter()
                Xi = S.solve_pw(dms_l[0], om, ri[0], ang, tol=tol, maxit=maxit)
                t3 = time.perf_counter()
                st = S.stats()
                with lock:
                    if keep is None or i in keep:
                        X[i] = Xi
                    status[i] = "fp64" if fp64 else ("fell_back" if st["fell_back_to_fp64"] else "mixed")
                    info[i] = dict(fill=t1 - t0, factor=t2 - t1, solve=t3 - t2, iters=S.iters, berr=S.berr,
                                   fell_back=st["fell_back_to_fp64"],
                                   precision="fp64" if (fp64 or precision == capi.MOM_PREC_FP64) else "mixed",
                                   lane=li)
        except BaseException as e:  # noqa: BLE001 -- handed to the caller's thread
            with lock:
                if state["error"] is None:
                    state["error"] = e
        finally:
            S.close()

    if nl == 1:
        lane(ctx, dmeshes, 0)
    else:
        if any(dm.frozen for dm in dmeshes):
            raise ValueError("lanes > 1 copies the meshes to new contexts, which cannot carry a classification "
                             "frozen by classify_like")
        extra = [Context(ctx.device) for _ in range(nl - 1)]
        try:
            meshes = [dmeshes] + [[DeviceMesh(c, dm.host) for dm in dmeshes] for c in extra]
            threads = [threading.Thread(target=lane, args=(c, m, li))
                       for li, (c, m) in enumerate(zip([ctx] + extra, meshes))]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
        finally:
            for c in extra:
                c.close()
    if state["error"] is not None:
        raise state["error"]
    return X, info
