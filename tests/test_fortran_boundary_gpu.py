"""The C-ABI exercised from Fortran: tests/fortran/boundary_test (built from
tests/fortran/boundary_test.f90 against the ISO_C_BINDING module
for90-mom2_amd/fortran/mom_gpu.f90 by `make -C tests/fortran`) passes the
reference's own flattened domain meshes of the s = 1 sphere to
mom_mesh_create and runs sysmat_pmchwt, srcvec, solve_multi_linsys and
cs_prtsrf through the binding on COMPLEX(dp) arrays it owns.  Its A, q, x
and sigma are checked against the reference's golden vectors (A 1e-10 per
block, q 1e-12, x and sigma 1e-8), which pins the binding's kinds, VALUE
attributes and argument order."""
import os
import subprocess

import numpy as np
import pytest

from cases import blockwise_relerr
from conftest import ROOT, golden
from mom2amd import DEGTORAD, omega_of, tri_quad_data

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "fortran", "_build", "boundary_test")


def _write_input(path, g):
    f64 = lambda a: np.ascontiguousarray(a, dtype="<f8").tobytes()
    i32 = lambda a: np.ascontiguousarray(a, dtype="<i4").tobytes()
    i64 = lambda a: np.ascontiguousarray(a, dtype="<i8").tobytes()
    ndom = 2
    N = g["mesh0_enodes"].shape[0]
    w, nodes = tri_quad_data()
    with open(path, "wb") as fh:
        fh.write(i32([ndom]) + i64([N]))
        for d in (1, 2):
            m = lambda k: g["mesh%d_%s" % (d, k)]
            fh.write(i64([m("p").shape[0], m("fnodes").shape[0], m("enodes").shape[0]]))
            fh.write(f64(m("p")) + i32(m("fnodes")) + i32(m("fedges")) + f64(m("fn")) + f64(m("fcp")) +
                     f64(m("fs")) + f64(m("fm")) + f64(m("farea")) + i32(m("efaces")) + i32(m("ebn")) +
                     f64(m("elen")) + i32(m("eparent")) + f64([float(m("avelen"))]))
        ri = np.array([[complex(r).real, complex(r).imag] for r in g["ri"]]).ravel()
        fh.write(f64([omega_of(float(g["wl"]))]) + f64(ri) + i32([len(w)]))
        fh.write(f64(w) + f64(np.asarray(nodes, dtype=np.float64).T.ravel()) + f64(g["angles_deg"] * DEGTORAD))
    return N


def test_fortran_binding_matches_reference(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/fortran/_build/boundary_test is not built (make -C tests/fortran)")
    g = golden("sphere_s1")
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    N = _write_input(fin, g)
    print("Fortran caller: sphere s=1, N=%d" % (2 * N), flush=True)
    subprocess.run([EXE, fin, fout], check=True, timeout=120)
    n2 = 2 * N
    raw = open(fout, "rb").read()
    c = np.frombuffer(raw, dtype="<c16", count=n2 * n2 + 2 * n2)
    A = c[:n2 * n2].reshape(n2, n2, order="F")
    q, x = c[n2 * n2:n2 * n2 + n2], c[n2 * n2 + n2:]
    off = 16 * (n2 * n2 + 2 * n2)
    ipiv = np.frombuffer(raw, dtype="<i4", count=n2, offset=off)
    csca, cabs = np.frombuffer(raw, dtype="<f8", count=2, offset=off + 4 * n2)
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()
    errs = dict(A=blockwise_relerr(A, g["A"]), q=rel(q, g["q"]), x=rel(x, g["x"]),
                csca=abs(csca / g["csca"] - 1), cabs=abs(cabs / g["cabs"] - 1))
    print("Fortran caller vs reference: " + ", ".join("%s %.2e" % kv for kv in errs.items()), flush=True)
    assert errs["A"] < 1e-10
    assert errs["q"] < 1e-12
    assert errs["x"] < 1e-8 and errs["csca"] < 1e-8 and errs["cabs"] < 1e-8
    assert ipiv.min() >= 1 and ipiv.max() <= n2
