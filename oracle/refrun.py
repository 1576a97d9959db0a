"""Run the reference golden-vector driver (oracle/_ref/mom_ref) and read its dumps.

TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py, the
reference-vs-oracle tests (skipped where oracle/_ref is not built) and
bench.py's cpu_baseline leg, which times the reference's own fill and LU.
"""
import os
import resource
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "mom_ref")


def available():
    return os.path.exists(REF_BIN)


def _unlimited_stack():
    # the reference keeps large automatic arrays (srcvec's field arrays,
    # sysmat_pmchwt's vector-subscript temporaries) on the stack
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def _env(threads):
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    env.setdefault("OMP_STACKSIZE", "256M")
    return env


def read_mesh(path):
    with open(path, "rb") as fh:
        V, F, E = np.fromfile(fh, np.int32, 3)
        r = lambda n, dt: np.fromfile(fh, dt, n)
        out = dict(fnodes=r(3 * F, np.int32).reshape(F, 3) - 1,
                   fedges=r(3 * F, np.int32).reshape(F, 3) - 1,
                   efaces=r(2 * E, np.int32).reshape(E, 2) - 1,
                   ebn=r(2 * E, np.int32).reshape(E, 2) - 1,
                   enodes=r(2 * E, np.int32).reshape(E, 2) - 1,
                   eparent=r(E, np.int32) - 1,
                   p=r(3 * V, np.float64).reshape(V, 3),
                   fn=r(3 * F, np.float64).reshape(F, 3),
                   fcp=r(3 * F, np.float64).reshape(F, 3),
                   fs=r(9 * F, np.float64).reshape(F, 3, 3),
                   fm=r(9 * F, np.float64).reshape(F, 3, 3),
                   farea=r(F, np.float64), elen=r(E, np.float64),
                   avelen=float(r(1, np.float64)[0]))
    return out


def run(mshfile, scale, domains, wl, ri, angles_deg, dump_ops=True, dump_sys=True, x=None,
        ntheta=5, nphi=7, outdir=None, threads=None, sh=None, solve=False, write_A=True):
    """Run the reference set-up + hot path; returns a dict of arrays.

    angles_deg: one (theta, phi, psi) triple or a list of them (one plane wave
    per triple, all solved together as solve_batch does).  dump_sys fills A
    (written unless write_A is False, for timing); solve runs the
    reference's own solve_systems (ZGETRF + ZGETRS) on A and returns its
    solution as "xref", which also feeds cs_prtsrf and rcs when x is None.
    sh (needs a solution, single source): dict(shri=[per domain], nls=[per
    domain 0/1], chi2=(nnn, ntt, ttn)) runs the second-harmonic chain of
    solve_batch."""
    outdir = outdir or tempfile.mkdtemp(prefix="momref_")
    ang = np.atleast_2d(np.asarray(angles_deg, dtype=np.float64))
    nsrc = ang.shape[0]
    if x is not None:
        np.asfortranarray(np.asarray(x, dtype=np.complex128).reshape(-1, nsrc, order="F")).T.ravel().tofile(
            os.path.join(outdir, "x.bin"))
    lines = [mshfile, repr(float(scale)), str(len(domains))]
    lines += [" ".join(str(v) for v in [len(d)] + list(d)) for d in domains]
    lines += [repr(float(wl))]
    lines += ["%r %r" % (complex(r).real, complex(r).imag) for r in ri]
    lines += [str(nsrc)] + ["%r %r %r" % tuple(float(a) for a in row) for row in ang]
    fill = 0 if not dump_sys else (1 if write_A else 2)
    lines += ["%d %d %d" % (int(dump_ops), fill, int(bool(solve) and fill > 0)), "%d %d" % (ntheta, nphi), outdir]
    if sh is not None:
        lines += ["1"] + ["%r %r" % (complex(r).real, complex(r).imag) for r in sh["shri"]]
        lines += [str(int(f)) for f in sh["nls"]]
        lines += [" ".join("%r %r" % (complex(c).real, complex(c).imag) for c in sh["chi2"])]
    subprocess.run([REF_BIN], input="\n".join(lines) + "\n", text=True, check=True,
                   stdout=subprocess.DEVNULL, env=_env(threads), preexec_fn=_unlimited_stack)
    res = {"mesh": [read_mesh(os.path.join(outdir, "mesh_%d.bin" % d)) for d in range(len(domains) + 1)]}
    N2 = 2 * res["mesh"][0]["enodes"].shape[0]
    q = np.fromfile(os.path.join(outdir, "q.bin"), np.complex128).reshape(N2, nsrc, order="F")
    res["q"] = q[:, 0] if nsrc == 1 else q
    if dump_sys and write_A:
        res["A"] = np.fromfile(os.path.join(outdir, "A.bin"), np.complex128).reshape(N2, N2, order="F")
    if dump_sys and solve:
        xr = np.fromfile(os.path.join(outdir, "xref.bin"), np.complex128).reshape(N2, nsrc, order="F")
        res["xref"] = xr[:, 0] if nsrc == 1 else xr
    if dump_ops:
        for d in range(len(domains)):
            E = res["mesh"][d + 1]["enodes"].shape[0]
            for op in "DHK":
                res["op_%d_%s" % (d, op)] = np.fromfile(os.path.join(outdir, "op_%d_%s.bin" % (d + 1, op)),
                                                        np.complex128).reshape(E, E, order="F")
    if (x is not None or (dump_sys and solve)) and ntheta > 0:
        post = np.fromfile(os.path.join(outdir, "post.bin"), np.float64)
        cs = post[:2 * nsrc].reshape(nsrc, 2)
        rcs = post[2 * nsrc:].reshape(ntheta, nphi, nsrc, order="F")
        if nsrc == 1:
            res["csca"], res["cabs"], res["rcs"] = cs[0, 0], cs[0, 1], rcs[:, :, 0]
        else:
            res["csca"], res["cabs"], res["rcs"] = cs[:, 0], cs[:, 1], rcs
    if sh is not None and (x is not None or solve):
        N2 = 2 * res["mesh"][0]["enodes"].shape[0]
        with open(os.path.join(outdir, "sh.bin"), "rb") as fh:
            for d, f in enumerate(sh["nls"]):
                if not f:
                    continue
                F = res["mesh"][d + 1]["fnodes"].shape[0]
                E = res["mesh"][d + 1]["enodes"].shape[0]
                res["sh_epsp_%d" % d] = np.fromfile(fh, np.complex128, F)
                res["sh_src_coef_%d" % d] = np.fromfile(fh, np.complex128, 2 * E)
                res["sh_src_vec_%d" % d] = np.fromfile(fh, np.complex128, 2 * E)
            res["sh_rhs0"] = np.fromfile(fh, np.complex128, N2)
            res["sh_rhs"] = np.fromfile(fh, np.complex128, N2)
            res["sh_x"] = np.fromfile(fh, np.complex128, N2)
            res["sh_rcs"] = np.fromfile(fh, np.float64, ntheta * nphi).reshape(ntheta, nphi, order="F")
    with open(os.path.join(outdir, "times.txt")) as fh:
        for ln in fh:
            k, v = ln.split()
            res["time_" + k] = float(v)
    return res


def run_lu(n, nrhs=1, threads=None, outdir=None):
    """Seconds of the reference's solve_multi_linsys (linalg.f90:159-180:
    ZGETRF + ZGETRS) on a seeded random complex n x n system with nrhs
    right-hand sides."""
    outdir = outdir or tempfile.mkdtemp(prefix="momreflu_")
    lines = ["lu", "%d %d" % (n, nrhs), outdir]
    subprocess.run([REF_BIN], input="\n".join(lines) + "\n", text=True, check=True, stdout=subprocess.DEVNULL,
                   stderr=subprocess.DEVNULL, env=_env(threads), preexec_fn=_unlimited_stack)
    with open(os.path.join(outdir, "times.txt")) as fh:
        t = dict((k, float(v)) for k, v in (ln.split() for ln in fh))
    return t["solve_multi_linsys"], int(t["threads"])


REF_SYMM_BIN = os.path.join(ROOT, "oracle", "_ref", "mom_ref_symm")


def run_symm(mshfile, scale, gens, sbnd, domains, wl, ri, angles_deg, ntheta=5, nphi=7, outdir=None, threads=None):
    """Run the reference's symmetric set-up and linear solve
    (oracle/_ref/mom_ref_symm): gens e.g. ["mxp", "myp", "mzp"] or ["r3"], sbnd a list
    of (line id, "xplane"|"yplane"|"zplane").  Returns the meshes with their
    edge classes, A[2N,2N,nga] before reduction, q[2N,nga], the expanded
    solutions x[2N,nga], csca, cabs and rcs."""
    outdir = outdir or tempfile.mkdtemp(prefix="momrefs_")
    lines = [mshfile, repr(float(scale)), " ".join([str(len(gens))] + list(gens)), str(len(sbnd))]
    lines += ["%d %s" % (int(i), name) for i, name in sbnd]
    lines += [str(len(domains))]
    lines += [" ".join(str(v) for v in [len(d)] + list(d)) for d in domains]
    lines += [repr(float(wl))]
    lines += ["%r %r" % (complex(r).real, complex(r).imag) for r in ri]
    lines += ["%r %r %r" % tuple(float(a) for a in angles_deg), "%d %d" % (ntheta, nphi), outdir]
    subprocess.run([REF_SYMM_BIN], input="\n".join(lines) + "\n", text=True, check=True,
                   stdout=subprocess.DEVNULL, env=_env(threads), preexec_fn=_unlimited_stack)
    res = {"mesh": [read_mesh(os.path.join(outdir, "mesh_%d.bin" % d)) for d in range(len(domains) + 1)]}
    for d in range(len(domains) + 1):
        res["mesh"][d]["bnd"] = np.fromfile(os.path.join(outdir, "bnd_%d.bin" % d), np.int32)
    N2 = 2 * res["mesh"][0]["enodes"].shape[0]
    nga = 1
    for g in gens:
        nga *= int(g[1:]) if g.startswith("r") else 2
    res["A"] = np.fromfile(os.path.join(outdir, "A.bin"), np.complex128).reshape(N2, N2, nga, order="F")
    res["q"] = np.fromfile(os.path.join(outdir, "q.bin"), np.complex128).reshape(N2, nga, order="F")
    res["x"] = np.fromfile(os.path.join(outdir, "x.bin"), np.complex128).reshape(N2, nga, order="F")
    post = np.fromfile(os.path.join(outdir, "post.bin"), np.float64)
    res["csca"], res["cabs"] = post[0], post[1]
    res["rcs"] = post[2:].reshape(ntheta, nphi, order="F")
    return res
