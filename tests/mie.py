"""Mie series for a homogeneous sphere (Bohren & Huffman, ch. 4).

Closed-form check of the whole PMCHWT chain: the cross-sections of a sphere
must approach these values as the mesh is refined.  Uses the Riccati-Bessel
functions psi_n(z) = z j_n(z), xi_n(z) = z h1_n(z) from scipy.special with
complex arguments, e^{-i omega t} convention (Im n > 0 is lossy).
"""
import numpy as np
from scipy.special import spherical_jn, spherical_yn


def mie_cross_sections(radius, k_out, n_out, n_in, nmax=None):
    """Return (sigma_sca, sigma_abs, sigma_ext) [m^2] for a sphere of `radius`
    in a medium of real index n_out, wavenumber k_out in that medium."""
    x = k_out * radius
    m = n_in / n_out
    if nmax is None:
        nmax = int(np.ceil(abs(x) + 4.05 * abs(x) ** (1 / 3) + 10))
    n = np.arange(1, nmax + 1)
    mx = m * x

    def psi(z):
        j = spherical_jn(n, z)
        dj = spherical_jn(n, z, derivative=True)
        return z * j, j + z * dj

    def xi(z):
        h = spherical_jn(n, z) + 1j * spherical_yn(n, z)
        dh = spherical_jn(n, z, derivative=True) + 1j * spherical_yn(n, z, derivative=True)
        return z * h, h + z * dh

    px, dpx = psi(x)
    pmx, dpmx = psi(mx)
    xx, dxx = xi(x)
    a = (m * pmx * dpx - px * dpmx) / (m * pmx * dxx - xx * dpmx)
    b = (pmx * dpx - m * px * dpmx) / (pmx * dxx - m * xx * dpmx)
    qsca = 2.0 / x**2 * np.sum((2 * n + 1) * (abs(a) ** 2 + abs(b) ** 2))
    qext = 2.0 / x**2 * np.sum((2 * n + 1) * np.real(a + b))
    g = np.pi * radius**2
    return qsca * g, (qext - qsca) * g, qext * g
