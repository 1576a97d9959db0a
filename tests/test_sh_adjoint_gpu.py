"""Adjoint of the second-harmonic chain (no counterpart in the reference):
the transposes it is built from, each against its forward operator by the
dot-product identity Re(g^T x) = Re(v^T L x) on seeded random vectors, and
the derivative of the SH power radiated into the 37 x 72 rcs_ext grid
with respect to the gap of the dimer (two R = 40 nm spheres, 10 nm gap,
chi2_nnn = 1, 1060 nm / 530 nm) against central differences of the whole
FP64 chain (set-up, both fills, LU, solves, nlsurf_coef, far field) at
gap +- h, h = 1e-3 nm, with the near/far classification of the perturbed
meshes frozen to that of the unperturbed ones.
Tolerances: dot-product identities 1e-11 (far field, block products) and
1e-9 (nlsurf_coef, through two conjugate-gradient Gram solves of relative
residual 1e-15 each); the gap derivative 2e-6 relative."""
import numpy as np
import pytest

from mom2amd import DEGTORAD, DeviceMesh, build_problem, determine_epsp, far_fields, meshgen, nlsurf_coef, omega_of
from mom2amd import sysmat_pmchwt
from mom2amd.capi import MOM_PREC_FP64
from mom2amd.nlsurf import (domain_block_apply, far_fields_adjoint, nlsurf_coef_adjoint, power_grid, sh_power,
                            sh_power_translation_gradient)

pytestmark = pytest.mark.gpu

OMEGA = omega_of(1060e-9)
RI = [1.0 + 0j, complex(np.sqrt(-40 + 2.5j)), complex(np.sqrt(-40 + 2.5j))]
SHRI = [1.0 + 0j, complex(np.sqrt(-9.4 + 1.5j)), complex(np.sqrt(-9.4 + 1.5j))]
NLS, CHI2 = [0, 1, 1], (1.0 + 0j, 0j, 0j)
ANG = np.array([0.0, 0.0, 0.0]) * DEGTORAD
DOMS = [[15, 16], [-15], [-16]]


def _dimer(s, gap):
    nodes, faces, ids = meshgen.dimer(s, 40.0, gap)
    return build_problem(nodes, faces, ids, 1e-9, DOMS)


def _crand(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


@pytest.fixture(scope="module")
def dimer1(gpu_ctx):
    P = _dimer(1, 10.0)
    dms = [DeviceMesh(gpu_ctx, d) for d in P.domains]
    yield P, dms
    for d in dms:
        d.close()


def test_far_fields_adjoint_is_the_transpose(gpu_ctx, dimer1):
    P, dms = dimer1
    N = P.nedges
    rng = np.random.default_rng(11)
    th, ph, _ = power_grid(9, 13)
    x, v = _crand(rng, 2 * N), _crand(rng, 4 * th.size).reshape(-1, 4)
    for om, ri in ((OMEGA, RI[0]), (2 * OMEGA, SHRI[0])):
        out = np.stack(far_fields(gpu_ctx, dms[0], N, om, ri, x, 1.0, th, ph), axis=1)
        g = far_fields_adjoint(gpu_ctx, dms[0], N, om, ri, 1.0, th, ph, v)
        a, b = np.sum(g * x), np.sum(v * out)
        print("far field: g^T x %.12e%+.12ej, v^T L x %.12e%+.12ej" % (a.real, a.imag, b.real, b.imag), flush=True)
        assert abs(a - b) <= 1e-11 * np.abs(g).sum() * np.abs(x).max()


def test_domain_block_apply_matches_the_matrix(gpu_ctx, dimer1):
    P, dms = dimer1
    N = P.nedges
    rng = np.random.default_rng(12)
    x = _crand(rng, 2 * N)
    for d in range(3):
        Z = sysmat_pmchwt(gpu_ctx, [dms[d]], [SHRI[d]], N, 2 * OMEGA)
        y0 = domain_block_apply(gpu_ctx, dms[d], N, 2 * OMEGA, SHRI[d], x)
        y1 = domain_block_apply(gpu_ctx, dms[d], N, 2 * OMEGA, SHRI[d], x, trans=True)
        e0 = np.abs(y0 - Z @ x).max() / np.abs(Z @ x).max()
        e1 = np.abs(y1 - Z.T @ x).max() / np.abs(Z.T @ x).max()
        print("domain %d: Z x %.2e, Z^T x %.2e" % (d, e0, e1), flush=True)
        assert e0 < 1e-11 and e1 < 1e-11


def test_nlsurf_coef_adjoint_is_the_transpose(gpu_ctx, dimer1):
    """nlsurf_coef is quadratic in x, so (N(x + dx) - N(x - dx)) / 2 is its
    Jacobian at x applied to dx, up to rounding."""
    P, dms = dimer1
    N = P.nedges
    rng = np.random.default_rng(13)
    x, dx = _crand(rng, 2 * N), _crand(rng, 2 * N)
    chi = (1.0 + 0.3j, 0.2 - 0.1j, 0.4 + 0.2j)  # every chi2 component, to reach every term
    for d in (1, 2):
        E = dms[d].nedges
        ep = determine_epsp(P, d, RI)
        sp, vp = nlsurf_coef(gpu_ctx, dms[d], N, OMEGA, RI[d], ep, x + dx, chi)
        sm, vm = nlsurf_coef(gpu_ctx, dms[d], N, OMEGA, RI[d], ep, x - dx, chi)
        js, jv = (sp - sm) / 2, (vp - vm) / 2
        scb, svb = _crand(rng, 2 * E), _crand(rng, 2 * E)
        xb = nlsurf_coef_adjoint(gpu_ctx, dms[d], N, OMEGA, RI[d], ep, x, chi, scb, svb)
        a, b = np.sum(xb * dx), np.sum(scb * js) + np.sum(svb * jv)
        scale = np.abs(scb * js).sum() + np.abs(svb * jv).sum()
        print("domain %d: xb^T dx %.12e%+.12ej, (J^T b)^T dx %.12e%+.12ej, rel %.1e"
              % (d, a.real, a.imag, b.real, b.imag, abs(a - b) / scale), flush=True)
        assert abs(a - b) <= 1e-9 * scale


@pytest.mark.parametrize("s", [1, 2])
def test_sh_power_gap_gradient_matches_frozen_differences(gpu_ctx, s):
    base = _dimer(s, 10.0)
    bdms = [DeviceMesh(gpu_ctx, d) for d in base.domains]
    try:
        px = base.master.p[:, 0]
        bodies = [(px < 0, (-0.5e-9, 0.0, 0.0)), (px > 0, (0.5e-9, 0.0, 0.0))]  # per nm of gap
        r = sh_power_translation_gradient(gpu_ctx, base, bdms, OMEGA, RI, SHRI, NLS, CHI2, ANG, bodies,
                                          precision=MOM_PREC_FP64)
        h = 1e-3
        vals = []
        for gap in (10.0 + h, 10.0 - h):
            Pp = _dimer(s, gap)
            dms = [DeviceMesh(gpu_ctx, d) for d in Pp.domains]
            try:
                for dm, bm in zip(dms, bdms):
                    dm.classify_like(bm)
                vals.append(sh_power(gpu_ctx, Pp, dms, OMEGA, RI, SHRI, NLS, CHI2, ANG, precision=MOM_PREC_FP64))
            finally:
                for d in dms:
                    d.close()
        fd = (vals[0] - vals[1]) / (2 * h)
        print("dimer s=%d: P %.10e, dP/dgap adjoint %.10e per nm, central differences %.10e, rel %.2e; terms %s"
              % (s, r["P"], r["dP"], fd, abs(r["dP"] - fd) / abs(fd), r["terms"]), flush=True)
        assert abs(r["dP"] - fd) <= 2e-6 * abs(fd)
    finally:
        for d in bdms:
            d.close()
