"""GPU check of the allocation-free far-field entry point
(mom_far_fields_noalloc): every array on the device, work areas in the
library's fixed per-GPU scratch array.  Its output must equal that of
mom_far_fields on the same inputs (the same kernels run on the same data,
so the values are compared bit for bit), on two calls in a row, the second
with a smaller mesh and fewer directions; a request beyond the scratch size
is rejected with MOM_ERR_ARG before anything runs."""
import ctypes as C

import numpy as np
import pytest
import torch

from cases import sphere
from conftest import golden
from mom2amd import DeviceMesh, capi, far_fields
from mom2amd.gpu import _quad

pytestmark = pytest.mark.gpu


def _noalloc(ctx, dm, N, omega, ri, x_d, r, th_d, ph_d, out_d):
    Q, w, n = _quad(None)
    ri = complex(ri)
    return capi.lib().mom_far_fields_noalloc(ctx.h, dm.h, N, omega, ri.real, ri.imag, C.c_void_p(x_d.data_ptr()),
                                             r, th_d.numel(), C.c_void_p(th_d.data_ptr()),
                                             C.c_void_p(ph_d.data_ptr()), Q, w, n, C.c_void_p(out_d.data_ptr()))


def _case(ctx, msh_dir, s, th, ph, r):
    g = golden("sphere_s%d" % s)
    case = sphere(msh_dir, s)
    N = case["problem"].nedges
    dm = DeviceMesh(ctx, case["problem"].domains[0])
    om, ri = case["omega"], case["ri"][0]
    ref = np.stack(far_fields(ctx, dm, N, om, ri, g["x"], r, th, ph), axis=-1)
    dev = torch.device("cuda:0")
    x_d = torch.from_numpy(np.ascontiguousarray(g["x"]).view(np.float64).copy()).to(dev)
    th_d = torch.from_numpy(np.ascontiguousarray(th, dtype=np.float64)).to(dev)
    ph_d = torch.from_numpy(np.ascontiguousarray(ph, dtype=np.float64)).to(dev)
    out_d = torch.full((8 * th.size,), np.nan, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    print("no-alloc far field: sphere s=%d, %d directions, r=%g" % (s, th.size, r), flush=True)
    capi.check(_noalloc(ctx, dm, N, om, ri, x_d, r, th_d, ph_d, out_d))
    got = out_d.cpu().numpy().view(np.complex128).reshape(-1, 4)
    return dm, N, om, ri, x_d, got, ref


def test_far_fields_noalloc_matches_far_fields(gpu_ctx, msh_dir):
    # 1) the reference's 37 x 72 rcs grid on the s = 2 sphere
    t = np.pi * np.arange(37) / 36
    p = 2 * np.pi * np.arange(72) / 71
    T, P = np.meshgrid(t, p, indexing="ij")
    dm, N, om, ri, x_d, got, ref = _case(gpu_ctx, msh_dir, 2, T.ravel(), P.ravel(), 1.0)
    assert np.isfinite(got).all()
    assert np.array_equal(got, ref), np.abs(got - ref).max() / np.abs(ref).max()
    # 2) smaller: the s = 1 sphere, 101 seeded directions at r = 3
    rng = np.random.default_rng(7)
    th, ph = np.arccos(rng.uniform(-1, 1, 101)), rng.uniform(0, 2 * np.pi, 101)
    dm1, N1, om1, ri1, x1, got1, ref1 = _case(gpu_ctx, msh_dir, 1, th, ph, 3.0)
    assert np.isfinite(got1).all()
    assert np.array_equal(got1, ref1), np.abs(got1 - ref1).max() / np.abs(ref1).max()
    # a request beyond the scratch array is refused before anything runs
    big = 400000
    dev = torch.device("cuda:0")
    th_b = torch.zeros(big, dtype=torch.float64, device=dev)
    out_b = torch.zeros(8 * big, dtype=torch.float64, device=dev)
    st = _noalloc(gpu_ctx, dm1, N1, om1, ri1, x1, 1.0, th_b, th_b, out_b)
    assert st == -1, st
    assert "scratch" in capi.lib().mom_last_error().decode()
    dm.close()
    dm1.close()
