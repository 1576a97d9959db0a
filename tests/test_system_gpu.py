"""End-to-end PMCHWT on the GPU: device-resident fill, complex128 LU or
complex64 LU + GMRES-IR, plane-wave RHS, cross-sections; against the
reference's golden solution and cross-sections (sigma rel. err <= 1e-8)."""
import numpy as np
import pytest

from cases import dimer, sphere
from conftest import golden
from mom2amd import DEGTORAD, DeviceMesh, System, cs_prtsrf
from mom2amd.capi import MOM_PREC_FP64, MOM_PREC_MIXED

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [MOM_PREC_FP64, MOM_PREC_MIXED])
@pytest.mark.parametrize("name", ["sphere_s1", "sphere_s2", "dimer_s1"])
def test_solution_and_cross_sections(gpu_ctx, msh_dir, name, prec):
    g = golden(name)
    case = {"sphere_s1": lambda: sphere(msh_dir, 1), "sphere_s2": lambda: sphere(msh_dir, 2),
            "dimer_s1": lambda: dimer(msh_dir, 1)}[name]()
    N = case["problem"].nedges
    dms = [DeviceMesh(gpu_ctx, d) for d in case["problem"].domains]
    S = System(gpu_ctx, N, prec)
    S.fill(dms, case["ri"], case["omega"])
    S.factor()
    ang = np.array(g["angles_deg"]) * DEGTORAD
    x = S.solve_pw(dms[0], case["omega"], case["ri"][0], ang, tol=1e-13, maxit=30)[:, 0]
    xerr = np.abs(x - g["x"]).max() / np.abs(g["x"]).max()
    cs, ca = cs_prtsrf(gpu_ctx, dms[0], N, case["omega"], case["ri"][0], x, *ang)
    es, ea = abs(cs / g["csca"] - 1), abs(ca / g["cabs"] - 1)
    print("%s prec %d: x err %.3e csca err %.3e cabs err %.3e (IR iters %d berr %.2e)" %
          (name, prec, xerr, es, ea, S.iters, S.berr), flush=True)
    assert xerr < 1e-8
    assert es < 1e-8 and ea < 1e-8
    S.close()


def test_fill_matches_sysmat(gpu_ctx, msh_dir):
    case = sphere(msh_dir, 1)
    g = golden("sphere_s1")
    dms = [DeviceMesh(gpu_ctx, d) for d in case["problem"].domains]
    S = System(gpu_ctx, case["problem"].nedges, MOM_PREC_MIXED)
    S.fill(dms, case["ri"], case["omega"])
    A = S.get_matrix()
    assert np.abs(A - g["A"]).max() / np.abs(g["A"]).max() < 1e-10
