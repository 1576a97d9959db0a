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

import time

import numpy as np

from . import capi
from .constants import DEGTORAD, omega_of
from .gpu import System

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
                   maxit=30, keep=None, pivoting=0):
    """For every wavelength wl in wls: fill with ri_of(wl) (one index per
    domain), factor, and solve all plane waves (angles in degrees, converted
    with the reference's degtorad) in one block in domain 1's medium.
    Returns (X, info): X[i] the 2N x nsrc solutions of wavelength i (kept
    only for i in `keep` when given), info per wavelength: fill/factor/solve
    seconds, refinement steps, backward error and whether the block solve
    fell back to FP64 factors.  pivoting: of the complex64 factorisation in
    the mixed path (0 partial, 1 tournament; System.set_pivoting)."""
    ang = np.asarray(angles_deg, dtype=np.float64) * DEGTORAD
    S = System(ctx, nedges_total, precision)
    S.set_pivoting(pivoting)
    X, info = {}, []
    try:
        for i, wl in enumerate(wls):
            om = omega_of(wl)
            ri = ri_of(wl)
            t0 = time.perf_counter()
            S.fill(dmeshes, ri, om)
            t1 = time.perf_counter()
            S.factor()
            t2 = time.perf_counter()
            Xi = S.solve_pw(dmeshes[0], om, ri[0], ang, tol=tol, maxit=maxit)
            t3 = time.perf_counter()
            st = S.stats()
            if keep is None or i in keep:
                X[i] = Xi
            info.append(dict(fill=t1 - t0, factor=t2 - t1, solve=t3 - t2, iters=S.iters, berr=S.berr,
                             fell_back=st["fell_back_to_fp64"]))
    finally:
        S.close()
    return X, info
