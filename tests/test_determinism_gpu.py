"""Run-to-run determinism of the fill, the right-hand sides and the mixed
solve: the same problem filled, sourced and solved twice on fresh systems
gives bitwise-identical Z, q, solutions and refinement history.  The
reference accumulates serially (sysmat.f90:326-343, source.f90:198-244);
here every sum has a fixed order (no floating-point atomics)."""
import numpy as np
import pytest
import torch  # loaded before the library, as the process's HIP runtime

from cases import dimer, sphere
from mom2amd import DEGTORAD, DeviceMesh, System, srcvec
from mom2amd.capi import MOM_PREC_MIXED

pytestmark = pytest.mark.gpu

ANGLES = np.array([[0.0, 90.0, 0.0], [30.0, 45.0, 90.0], [120.0, 200.0, 0.0], [75.0, 10.0, 45.0],
                   [10.0, 300.0, 90.0], [160.0, 100.0, 30.0]]) * DEGTORAD


def _run(ctx, case, dms):
    N = case["problem"].nedges
    S = System(ctx, N, MOM_PREC_MIXED)
    try:
        S.fill(dms, case["ri"], case["omega"])
        Z = S.get_matrix()
        q = srcvec(ctx, dms[0], N, case["omega"], case["ri"][0], ANGLES)
        S.factor()
        x1 = S.solve_pw(dms[0], case["omega"], case["ri"][0], ANGLES[:1])
        h1 = S.stats()["berr_history"]
        xb = S.solve_pw(dms[0], case["omega"], case["ri"][0], ANGLES)
        hb = S.stats()["berr_history"]
    finally:
        S.close()
    return Z, q, x1, h1, xb, hb


@pytest.mark.parametrize("name", ["sphere_s3", "dimer_s1"])
def test_fill_sources_and_solve_are_bitwise_reproducible(gpu_ctx, msh_dir, name):
    case = sphere(msh_dir, 3) if name == "sphere_s3" else dimer(msh_dir, 1)
    dms = [DeviceMesh(gpu_ctx, d) for d in case["problem"].domains]
    a = _run(gpu_ctx, case, dms)
    b = _run(gpu_ctx, case, dms)
    labels = ["Z", "q", "x (1 source)", "refinement history (1 source)", "X (6 sources)",
              "refinement history (6 sources)"]
    for lab, u, v in zip(labels, a, b):
        same = np.array_equal(np.asarray(u), np.asarray(v))
        print("%s %s bitwise equal: %s" % (name, lab, same), flush=True)
        assert same, lab
    for d in dms:
        d.close()


def test_lossless_pair_kernel_gives_the_same_bits(gpu_ctx, msh_dir):
    """The exterior domain (vacuum, Im k = 0) fills with the pair kernel's
    lossless variant, which skips the exponential of e^{ikR} (exactly 1
    there): the matrix is bitwise the one of the general kernel."""
    from mom2amd import capi
    case = sphere(msh_dir, 3)
    dms = [DeviceMesh(gpu_ctx, d) for d in case["problem"].domains]
    N = case["problem"].nedges
    L = capi.lib()
    Z = []
    try:
        for enable in (1, 0):
            capi.check(L.mom_test_fill_lossless(enable))
            S = System(gpu_ctx, N, MOM_PREC_MIXED)
            try:
                S.fill(dms, case["ri"], case["omega"])
                Z.append(S.get_matrix())
            finally:
                S.close()
    finally:
        capi.check(L.mom_test_fill_lossless(1))
        for d in dms:
            d.close()
    assert case["ri"][0].imag == 0.0
    same = np.array_equal(Z[0], Z[1])
    print("lossless variant vs general kernel: bitwise equal %s" % same, flush=True)
    assert same


@pytest.mark.parametrize("name", ["sphere_s3", "dimer_s1"])
def test_every_pair_kernel_variant_gives_the_same_bits(gpu_ctx, msh_dir, name):
    """The fill's autotuner picks 2, 4 or 8 source faces per wave per
    context, and a lossless medium takes the variant without the
    exponential: all six fills give bitwise the same matrix."""
    from mom2amd import capi
    case = sphere(msh_dir, 3) if name == "sphere_s3" else dimer(msh_dir, 1)
    dms = [DeviceMesh(gpu_ctx, d) for d in case["problem"].domains]
    N = case["problem"].nedges
    L = capi.lib()
    ref, same = None, {}
    try:
        for ll in (1, 0):
            capi.check(L.mom_test_fill_lossless(ll))
            for spw in (2, 4, 8):
                capi.check(L.mom_test_fill_spw(gpu_ctx.h, spw))
                S = System(gpu_ctx, N, MOM_PREC_MIXED)
                try:
                    S.fill(dms, case["ri"], case["omega"])
                    Z = S.get_matrix()
                finally:
                    S.close()
                if ref is None:
                    ref = Z
                same[(ll, spw)] = bool(np.array_equal(Z, ref))
                if not same[(ll, spw)]:
                    d = np.abs(Z - ref)
                    print("%s lossless %d, %d sources per wave: max diff %.3e at %s (|Z| %.3e)"
                          % (name, ll, spw, d.max(), np.unravel_index(d.argmax(), d.shape), np.abs(ref).max()),
                          flush=True)
    finally:
        capi.check(L.mom_test_fill_lossless(1))
        capi.check(L.mom_test_fill_spw(gpu_ctx.h, 0))
        for d in dms:
            d.close()
    print("%s bitwise equal by (lossless, sources per wave): %s" % (name, same), flush=True)
    assert all(same.values())


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_single_rhs_solve_is_bitwise_stable_under_contention(gpu_ctx, msh_dir, prec):
    """The one-column solves (block substitution with the diagonal-block
    inverses, whose product writes into a separate block buffer that the
    next launch copies back) give the same bits while another stream keeps
    the device busy and their workgroups start at scattered times."""
    from mom2amd.capi import MOM_PREC_FP64
    case = sphere(msh_dir, 3)
    dms = [DeviceMesh(gpu_ctx, d) for d in case["problem"].domains]
    N = case["problem"].nedges
    S = System(gpu_ctx, N, MOM_PREC_MIXED if prec == "mixed" else MOM_PREC_FP64)
    try:
        S.fill(dms, case["ri"], case["omega"])
        S.factor()
        ref = S.solve_pw(dms[0], case["omega"], case["ri"][0], ANGLES[:1])
        href = S.stats()["berr_history"]
        side = torch.cuda.Stream()
        a = torch.randn(4096, 4096, device="cuda")
        for rep in range(4):
            with torch.cuda.stream(side):
                for _ in range(12):
                    a = torch.tanh(a @ a * 1e-3)
            x = S.solve_pw(dms[0], case["omega"], case["ri"][0], ANGLES[:1])
            h = S.stats()["berr_history"]
            torch.cuda.synchronize()
            same = np.array_equal(x, ref) and np.array_equal(np.asarray(h), np.asarray(href))
            print("%s repeat %d under contention: bitwise equal %s" % (prec, rep, same), flush=True)
            assert same
    finally:
        S.close()
        for d in dms:
            d.close()
