"""PyTorch harness around the adjoint entry points: sigma_abs(eps_in) of a
sphere as a torch.autograd.Function whose backward is the library's adjoint
gradient (one Z^H solve and one derivative fill), checked by
torch.autograd.gradcheck (central differences of the whole solve).
Units: sigma in nm^2, so that gradcheck's absolute tolerances are
meaningful."""
import numpy as np
import pytest

from cases import sphere
from mom2amd import DeviceMesh, System, sigma_abs_eps_gradient
from mom2amd.capi import MOM_PREC_FP64, MOM_PREC_MIXED

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def make_function(ctx, dms, N, omega, prec):
    ang = np.array([0.0, 90.0, 0.0]) * np.pi / 180

    class SigmaAbs(torch.autograd.Function):
        @staticmethod
        def forward(fctx, eps):
            e = complex(eps[0].item(), eps[1].item())
            ri = [1.0 + 0j, np.sqrt(e)]
            S = System(ctx, N, prec)
            try:
                S.fill(dms, ri, omega)
                S.factor()
                x = S.solve_pw(dms[0], omega, 1.0, ang)[:, 0]
                sab, dre, dim = sigma_abs_eps_gradient(ctx, S, x, dms, 1, omega, ri)
            finally:
                S.close()
            fctx.save_for_backward(torch.tensor([dre, dim], dtype=torch.float64) * 1e18)
            return torch.tensor(sab * 1e18, dtype=torch.float64)

        @staticmethod
        def backward(fctx, gout):
            (grad,) = fctx.saved_tensors
            return gout * grad

    return SigmaAbs


@pytest.mark.parametrize("prec", [MOM_PREC_FP64, MOM_PREC_MIXED], ids=["fp64", "mixed"])
def test_sigma_abs_autograd_gradcheck(gpu_ctx, msh_dir, prec):
    case = sphere(msh_dir, 2)
    dms = [DeviceMesh(gpu_ctx, d) for d in case["problem"].domains]
    try:
        fn = make_function(gpu_ctx, dms, case["problem"].nedges, case["omega"], prec).apply
        e = complex(case["ri"][1]) ** 2
        eps = torch.tensor([e.real, e.imag], dtype=torch.float64, requires_grad=True)
        out = fn(eps)
        out.backward()
        print("sigma_abs %.6f nm^2, grad %s" % (out.item(), eps.grad.tolist()), flush=True)
        # mixed: the forward values carry the refinement's ~1e-11 relative
        # error, which the central differences divide by 2h
        tol = 1e-6 if prec == MOM_PREC_FP64 else 1e-4
        assert torch.autograd.gradcheck(fn, (eps,), eps=1e-3, atol=tol * abs(eps.grad).max().item(), rtol=tol)
    finally:
        for d in dms:
            d.close()


def make_vertex_function(ctx, faces, omega, prec):
    """sigma_abs(node coordinates [nm]) of the sphere as a torch.autograd.Function
    whose backward is the node-coordinate adjoint (sigma_abs_vertex_gradient)."""
    from mom2amd import DEGTORAD, build_problem, cs_prtsrf, sigma_abs_vertex_gradient
    from cases import N_IN_SPHERE

    ang = np.array([40.0, 25.0, 30.0]) * DEGTORAD
    ri = [1.0 + 0j, complex(N_IN_SPHERE)]

    class SigmaAbsNodes(torch.autograd.Function):
        @staticmethod
        def forward(fctx, p):
            nodes = p.detach().cpu().numpy().reshape(-1, 3)
            P = build_problem(nodes, faces, np.full(faces.shape[0], 15), 1e-9, [[15], [-15]])
            dms = [DeviceMesh(ctx, d) for d in P.domains]
            S = System(ctx, P.nedges, prec)
            try:
                S.fill(dms, ri, omega)
                S.factor()
                x = S.solve_pw(dms[0], omega, ri[0], ang)[:, 0]
                if fctx is None:
                    sab, g = cs_prtsrf(ctx, dms[0], P.nedges, omega, ri[0], x, *ang)[1], None
                else:
                    sab, g = sigma_abs_vertex_gradient(ctx, S, x, dms, P, omega, ri, ang)
            finally:
                S.close()
                for d in dms:
                    d.close()
            if g is not None:
                # sigma in nm^2, coordinates in nm: d sigma / d p [nm] = g [m^2/m] * 1e18 * 1e-9
                fctx.save_for_backward(torch.tensor(g.reshape(p.shape) * 1e9, dtype=torch.float64))
            return torch.tensor(sab * 1e18, dtype=torch.float64)

        @staticmethod
        def backward(fctx, gout):
            (grad,) = fctx.saved_tensors
            return gout * grad

    return SigmaAbsNodes


def test_sigma_abs_node_gradient_gradcheck(gpu_ctx):
    """gradcheck of the node-coordinate adjoint on the s=1 icosphere (42
    nodes, 126 coordinates): central differences of the whole FP64 solve with
    eps = 1e-3 nm, rtol 1e-5."""
    from mom2amd import meshgen, omega_of

    nodes, faces = meshgen.icosphere(1, 50.0)
    fn = make_vertex_function(gpu_ctx, faces, omega_of(600e-9), MOM_PREC_FP64).apply
    p = torch.tensor(nodes, dtype=torch.float64, requires_grad=True)
    out = fn(p)
    out.backward()
    print("sigma_abs %.6f nm^2, |grad| max %.3e nm" % (out.item(), p.grad.abs().max().item()), flush=True)
    assert torch.autograd.gradcheck(fn, (p,), eps=1e-3, atol=1e-5 * p.grad.abs().max().item(), rtol=1e-5)
