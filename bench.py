#!/usr/bin/env python3
"""Benchmark of the PMCHWT hot path on one MI355X.

Workload (default): a sphere of radius 50 nm meshed as an icosahedron
subdivided 5 times and projected to the sphere (20480 triangles, 30720 RWG
edges, N = 61440 unknowns J and M), exterior index 1, interior
eps = -9.4+1.5i, vacuum wavelength 600 nm, x-polarised plane wave along +z.

One step = the whole solve on the device: PMCHWT fill (complex128), FP32
matrix-core LU of the equilibrated complex64 copy (tournament pivoting;
--pmchwt-pivoting partial for LAPACK's pivots), plane-wave RHS and GMRES
iterative refinement against the complex128 matrix to backward error
<= 1e-12.  `value` is the mean wall time of a step [s].  After the timed
steps the all-FP64 GPU path (complex128 fill + LU + solve) is timed once on
the same problem, and the cross-sections of both solutions are compared.

Prints one JSON line.  --gpus > 1 is not supported (one GPU).
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "for90-mom2_amd"))

import numpy as np  # noqa: E402

METRIC = "fill+LU+solve s @N=61440: FP32-MFMA+IR vs FP64 GPU path, σ rel err≤1e-8"
PEAK_FP32_MFMA = 157.3  # TFLOP/s, MI355X dense FP32 matrix
PEAK_FP64_MFMA = 78.6
# HBM bytes per launch of the depth-256 trailing update at N = 61440 (mean over its 477 launches per
# solve, next-column and remaining-columns launches together): rocprofv3 FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0, the factors read off a 1 GiB elementwise copy.  Recorded in
# profiles/gemm3m_dma_kernel32_N61440_hbm_traffic.md; counters cannot be read in-process.
TRAFFIC_S5_BYTES = 13.06e9
PEAK_FP64_VALU = 78.6  # TFLOP/s, MI355X FP64 vector
PEAK_HBM = 8000.0  # GB/s
# Fill: SURVEY's algorithmic count of 60 flops per (test point, source point) pair.  The FP64 VALU
# flops the pair kernel issues per point pair (SQ_INSTS_VALU_FLOPS_FP64 over its launches of one
# N = 61440 fill divided by the point pairs they evaluate) and the HBM bytes of the assembly kernels
# per fill (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0) are recorded in
# profiles/pair_plane_kernel_N61440_fp64_counters.md.
FILL_FLOPS_PER_POINT_PAIR = 60
FILL_ISSUED_FLOPS_PER_POINT_PAIR = 159.8
ASSEMBLY_TRAFFIC_S5_BYTES = 6.586e11


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def build_case(subdiv):
    from mom2amd import load_problem, meshgen, omega_of

    nodes, faces = meshgen.icosphere(subdiv, 50.0)
    d = tempfile.mkdtemp(prefix="mom_bench_")
    path = os.path.join(d, "ico%d.msh" % subdiv)
    meshgen.write_gmsh(path, nodes, faces, 15)
    doms = [[15], [-15]]
    prob = load_problem(path, 1e-9, doms)
    return dict(path=path, domains=doms, problem=prob, omega=omega_of(600e-9),
                ri=[1.0 + 0j, complex(np.sqrt(-9.4 + 1.5j))])


def run_path(ctx, dms, case, prec, profile=0, tol=1e-12, pivoting=0):
    from mom2amd import DEGTORAD, System, cs_prtsrf, rcs
    from mom2amd import capi

    N = case["problem"].nedges
    ang = np.array([0.0, 90.0, 0.0]) * DEGTORAD
    S = System(ctx, N, prec)
    S.set_pivoting(pivoting)
    if profile:
        capi.check(capi.lib().mom_system_profile(S.h, profile))
        ctx.fill_profile(1)  # events around the fill's pair and assembly launches
    t0 = time.perf_counter()
    S.fill(dms, case["ri"], case["omega"])
    fill_prof = ctx.fill_profile(0) if profile else None
    S.factor()
    x = S.solve_pw(dms[0], case["omega"], case["ri"][0], ang, tol=tol, maxit=30)[:, 0]
    wall = time.perf_counter() - t0
    st = S.stats()
    st.update(wall_s=wall, iters=S.iters, berr=S.berr, fill_prof=fill_prof)
    cs = cs_prtsrf(ctx, dms[0], N, case["omega"], case["ri"][0], x, *ang)
    # bistatic radar cross-section on the reference's 37 x 72 grid, outside the timed region
    st["rcs"] = rcs(ctx, dms[0], N, case["omega"], case["ri"][0], x, 37, 72)
    S.close()
    return st, cs


def library_zgemm_tflops(m=16384, k=256, reps=3):
    """Yardstick for the FP64 trailing update: the vendor library's
    complex128 GEMM (torch.addmm on the GPU, which calls hipBLAS/rocBLAS) on
    C -= A B with M = N = m and K = k, TFLOP/s at 8MNK.  Not part of the
    library; None where torch cannot run it."""
    try:
        import torch
        if not torch.cuda.is_available():
            log("library GEMM yardstick skipped: torch sees no GPU")
            return None
        dev = torch.device("cuda:0")
        g = torch.Generator(device=dev).manual_seed(0)
        A = torch.randn(m, k, dtype=torch.complex128, device=dev, generator=g)
        B = torch.randn(k, m, dtype=torch.complex128, device=dev, generator=g)
        C = torch.randn(m, m, dtype=torch.complex128, device=dev, generator=g)
        C.addmm_(A, B, beta=1.0, alpha=-1.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            C.addmm_(A, B, beta=1.0, alpha=-1.0)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        del A, B, C
        torch.cuda.empty_cache()
        return 8.0 * m * m * k / dt / 1e12
    except Exception as e:  # noqa: BLE001
        log("library GEMM yardstick skipped: %s" % e)
        return None


def cpu_baseline(case, subdiv, threads):
    """The reference itself (oracle/_ref/mom_ref: its Fortran modules built
    with flang -O2 -fopenmp, LAPACK from scipy's OpenBLAS), timed on a bounded
    sample and scaled to the workload: sysmat_pmchwt and the srcvec loop on
    the icospheres s = 2 and 3 of the same sphere (fill time fitted as
    a F^2 + b F in the faces per domain F: far pairs and near pairs), and its
    solve_multi_linsys (ZGETRF + ZGETRS) on a random system of order 6144,
    scaled by n^3.  Falls back to the C restatement (oracle/liboracle.so) on
    a sample where the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    from mom2amd import meshgen

    F = case["problem"].domains[0].nfaces
    n = 2 * case["problem"].nedges
    if refrun.available():
        d = tempfile.mkdtemp(prefix="mom_cpuref_")
        pts = []
        for s in (2, 3):
            nodes, faces = meshgen.icosphere(s, 50.0)
            path = os.path.join(d, "ico%d.msh" % s)
            meshgen.write_gmsh(path, nodes, faces, 15)
            R = refrun.run(path, 1e-9, case["domains"], 600e-9, case["ri"], (0.0, 90.0, 0.0), dump_ops=False,
                           dump_sys=True, write_A=False, solve=False, ntheta=0, nphi=0, threads=threads,
                           outdir=tempfile.mkdtemp(dir=d))
            pts.append((faces.shape[0], R["time_sysmat_pmchwt"], R["time_srcvec"]))
        (f2, t2, s2), (f3, t3, s3) = pts
        a = (t3 - t2 * f3 / f2) / (f3 * f3 - f2 * f3)
        b = (t2 - a * f2 * f2) / f2
        fill_est = a * F * F + b * F
        src_est = s3 * F / f3
        ns = 6144
        t_lu, used = refrun.run_lu(ns, 1, threads=threads)
        lu_est = t_lu * (n / ns) ** 3
        sample = ("reference Fortran (flang -O2 -fopenmp, %d OpenMP threads, scipy OpenBLAS): sysmat_pmchwt on the "
                  "icospheres s=2 (%d faces/domain, %.2f s) and s=3 (%d, %.2f s), fitted a F^2 + b F and evaluated at "
                  "F=%d; srcvec scaled by F; solve_multi_linsys (ZGETRF+ZGETRS) at n=%d (%.2f s) scaled by n^3 to n=%d"
                  % (used, f2, t2, f3, t3, F, ns, t_lu, n))
        return dict(value=fill_est + src_est + lu_est, unit="s", cores=used, kind="reference", sample=sample,
                    fill_s=fill_est, lu_s=lu_est)
    # no reference build: the C restatement on a sample
    import oracle
    import scipy.linalg as sl

    op = oracle.OracleProblem(case["path"], 1e-9, case["domains"])
    rows = 32
    t_rows = op.fill_rows_seconds(case["omega"], case["ri"], 0, rows)
    fill_est = t_rows * F / rows
    ns = 4096
    rng = np.random.default_rng(0)
    A = rng.standard_normal((ns, ns)) + 1j * rng.standard_normal((ns, ns))
    b = rng.standard_normal(ns) + 0j
    t0 = time.perf_counter()
    lu = sl.lu_factor(A, overwrite_a=True, check_finite=False)
    sl.lu_solve(lu, b, check_finite=False)
    t_lu = time.perf_counter() - t0
    lu_est = t_lu * (n / ns) ** 3
    sample = ("oracle C fill (OpenMP) over %d of %d test faces per domain (%.2f s) scaled by faces; "
              "LAPACK zgetrf+zgetrs (scipy OpenBLAS) at n=%d (%.2f s) scaled by n^3 to n=%d"
              % (rows, F, t_rows, ns, t_lu, n))
    return dict(value=fill_est + lu_est, unit="s", cores=threads, kind="port", sample=sample,
                fill_s=fill_est, lu_s=lu_est)


def bench_sh(args):
    """Second-harmonic workload: dimer of two R = 40 nm spheres (icosphere
    s each, 10 nm gap along x), 3 domains, eps_in = -40+2.5i at 1060 nm and
    -9.4+1.5i at the SH, surface chi2 nnn = 1 on both spheres, x-polarised
    plane wave along +z.  One step = fundamental fill + LU + solve,
    nlsurf_coef of both spheres, SH fill (sysmat_pmchwt_nls) + LU + solve and
    the SH radar cross-section on the 37 x 72 grid, all in the mixed path.
    After the timed steps: the derivative of the SH power on that grid with
    respect to the gap width by the adjoint, and its central differences."""
    from mom2amd import (DEGTORAD, Context, DeviceMesh, System, build_problem, load_problem, meshgen, omega_of, rcs,
                         sh_power, sh_power_translation_gradient, solve_sh)
    from mom2amd.capi import MOM_PREC_FP64, MOM_PREC_MIXED

    s = args.subdiv if args.subdiv != 5 else 4
    nodes, faces, ids = meshgen.dimer(s, 40.0, 10.0)
    d = tempfile.mkdtemp(prefix="mom_bench_")
    path = os.path.join(d, "dimer%d.msh" % s)
    meshgen.write_gmsh(path, nodes, faces, ids)
    doms = [[15, 16], [-15], [-16]]
    t0 = time.perf_counter()
    prob = load_problem(path, 1e-9, doms)
    N = prob.nedges
    log("dimer s=%d: %d faces, N=%d unknowns (set-up %.2f s)" % (s, prob.master.nfaces, 2 * N,
                                                                  time.perf_counter() - t0))
    omega = omega_of(1060e-9)
    ri = [1.0 + 0j, complex(np.sqrt(-40 + 2.5j)), complex(np.sqrt(-40 + 2.5j))]
    shri = [1.0 + 0j, complex(np.sqrt(-9.4 + 1.5j)), complex(np.sqrt(-9.4 + 1.5j))]
    nls, chi2 = [0, 1, 1], (1.0 + 0j, 0j, 0j)
    ang = np.array([0.0, 0.0, 0.0]) * DEGTORAD
    ctx = Context(0)
    dms = [DeviceMesh(ctx, dm) for dm in prob.domains]

    def step(prec):
        S = System(ctx, N, prec)
        S.set_pivoting(1 if args.pivoting == "tournament" else 0)
        try:
            t = [time.perf_counter()]
            S.fill(dms, ri, omega)
            S.factor()
            x = S.solve_pw(dms[0], omega, ri[0], ang)[:, 0]
            t.append(time.perf_counter())
            sh = solve_sh(ctx, prob, dms, omega, ri, shri, nls, chi2, x, precision=prec, system=S)
            t.append(time.perf_counter())
            r = rcs(ctx, dms[0], N, 2 * omega, shri[0], sh["x"], 37, 72)
            t.append(time.perf_counter())
        finally:
            S.close()
        return dict(wall=t[-1] - t[0], fund=t[1] - t[0], sh=t[2] - t[1], rcs=t[3] - t[2], rcs_grid=r,
                    iters=sh["iters"], berr=sh["berr"])

    for i in range(args.warmup):
        w = step(MOM_PREC_MIXED)
        log("warmup %d: %.2f s" % (i, w["wall"]))
    runs = []
    for i in range(args.steps):
        r = step(MOM_PREC_MIXED)
        runs.append(r)
        log("step %d: %.2f s (fundamental fill+LU+solve %.2f, SH source+fill+LU+solve %.2f, SH rcs %.3f; "
            "SH IR iters %d berr %.1e)" % (i, r["wall"], r["fund"], r["sh"], r["rcs"], r["iters"], r["berr"]))
    mean = float(np.mean([r["wall"] for r in runs]))
    out = {"metric": "SH chain s @N=%d: fundamental + SH fill/LU/solve, nlsurf_coef, SH rcs 37x72 (mixed path)"
                     % (2 * N),
           "value": round(mean, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(mean * 1e3, 2), "higher_is_better": False, "scaling": "strong",
           "vs_baseline": None, "dtype": "fp32-mfma LU + fp64 fill/residual", "data": "synthetic icosphere dimer",
           "config": {"workload": "SH dimer: two R=40nm icospheres s=%d, 10 nm gap, %d faces, N=%d unknowns, "
                                  "1060 nm / 530 nm, chi2_nnn=1" % (s, prob.master.nfaces, 2 * N),
                      "n_unknowns": 2 * N, "complex64_pivoting": args.pivoting},
           "phases_s": {k: round(float(np.mean([r[k] for r in runs])), 4) for k in ("fund", "sh", "rcs")}}
    if not args.no_fp64:
        r64 = step(MOM_PREC_FP64)
        log("fp64 path: %.2f s (fundamental %.2f, SH %.2f)" % (r64["wall"], r64["fund"], r64["sh"]))
        out["fp64_path_s"] = round(r64["wall"], 4)
        out["speedup_vs_fp64_path"] = round(r64["wall"] / mean, 3)
        out["sh_rcs_relerr_vs_fp64_path"] = float(np.abs(runs[-1]["rcs_grid"] - r64["rcs_grid"]).max()
                                                  / np.abs(r64["rcs_grid"]).max())
    # derivative of the SH power radiated into the 37 x 72 grid with respect to the gap, by the adjoint
    # (mixed path), checked against central differences of the FP64 chain (h = 1e-3 nm, near/far
    # classification of the perturbed meshes frozen to the unperturbed ones)
    px = prob.master.p[:, 0]
    bodies = [(px < 0, (-0.5e-9, 0.0, 0.0)), (px > 0, (0.5e-9, 0.0, 0.0))]  # per nm of gap
    t0 = time.perf_counter()
    gr = sh_power_translation_gradient(ctx, prob, dms, omega, ri, shri, nls, chi2, ang, bodies,
                                       precision=MOM_PREC_MIXED, pivoting=1 if args.pivoting == "tournament" else 0)
    tg = time.perf_counter() - t0
    out["gap_gradient"] = {"P_sh_grid": gr["P"], "dP_dgap_per_nm": gr["dP"], "seconds": round(tg, 3),
                           "terms": gr["terms"],
                           "note": "P = sum over the 37 x 72 rcs_ext grid of W (|et|^2 + |ep|^2) at 2 omega, W the "
                                   "trapezoid solid-angle weights; forward chain + adjoint (far-field transpose, Z2^T "
                                   "and Z1^T solves, nlsurf_coef transpose, vertex adjoint of every domain's blocks)"}
    log("SH power gap gradient by the adjoint: P %.6e, dP/dgap %.6e per nm (%.2f s incl. the forward chain)"
        % (gr["P"], gr["dP"], tg))
    if not args.no_fp64:
        vals = []
        for gap in (10.0 + 1e-3, 10.0 - 1e-3):
            n2, f2, i2 = meshgen.dimer(s, 40.0, gap)
            p2 = build_problem(n2, f2, i2, 1e-9, doms)
            dms2 = [DeviceMesh(ctx, dm) for dm in p2.domains]
            try:
                for dm2, dm in zip(dms2, dms):
                    dm2.classify_like(dm)
                vals.append(sh_power(ctx, p2, dms2, omega, ri, shri, nls, chi2, ang, precision=MOM_PREC_FP64))
            finally:
                for dm in dms2:
                    dm.close()
        fd = (vals[0] - vals[1]) / 2e-3
        out["gap_gradient"]["fd_fp64_per_nm"] = fd
        out["gap_gradient"]["relerr_vs_fd"] = abs(gr["dP"] - fd) / abs(fd)
        log("central differences of the FP64 chain: %.6e per nm (rel err of the adjoint %.1e)"
            % (fd, abs(gr["dP"] - fd) / abs(fd)))
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_spectrum(args):
    """Batched spectrum with many right-hand sides: a prolate ellipsoid (the
    s-subdivided unit icosphere scaled to semi-axes 80/40/40 nm), 64 vacuum
    wavelengths of `wlrg 64 400d-9 900d-9`, interior eps from the Drude model
    eps_inf - wp^2/(w(w + i gamma)) (eps_inf 9.5, wp 9.0 eV, gamma 0.07 eV; a
    single pole), 256 plane waves per wavelength (psi 0 and 90 deg times 128
    directions uniform on the sphere, numpy seed 20251203).  One step = for
    every wavelength the fill, the LU and the 256 solves on the reused
    factors (block refinement to backward error 1e-12), --lanes wavelengths
    at a time, each lane a device-resident system reused across its
    wavelengths (mom2amd.spectrum.solve_spectrum, the path the spectrum
    parity test runs).  The all-FP64 path is timed one wavelength at a time
    on eight wavelengths spread over the range, and in lanes over all of
    them."""
    from mom2amd import DEGTORAD, Context, DeviceMesh, cs_prtsrf, load_problem, meshgen, omega_of
    from mom2amd.capi import MOM_PREC_FP64, MOM_PREC_MIXED
    from mom2amd.spectrum import drude_index, plane_waves_deg, solve_spectrum, wavelengths

    s = args.subdiv if args.subdiv != 5 else 4
    nodes, faces = meshgen.icosphere(s, 1.0)
    nodes = nodes * np.array([80.0, 40.0, 40.0])
    d = tempfile.mkdtemp(prefix="mom_bench_")
    path = os.path.join(d, "ellipsoid%d.msh" % s)
    meshgen.write_gmsh(path, nodes, faces, 15)
    prob = load_problem(path, 1e-9, [[15], [-15]])
    N = prob.nedges
    nwl = args.nwl
    wls = wavelengths(nwl)
    ang = plane_waves_deg(128)
    ri_of = lambda wl: [1.0 + 0j, drude_index(wl)]  # noqa: E731
    log("ellipsoid s=%d: %d faces, N=%d unknowns, %d wavelengths x %d sources, %d lanes"
        % (s, prob.master.nfaces, 2 * N, nwl, ang.shape[0], args.lanes))
    ctx = Context(0)
    dms = [DeviceMesh(ctx, dm) for dm in prob.domains]
    piv = 1 if args.pivoting == "tournament" else 0

    def sweep(prec, idx, keep=None, lanes=args.lanes):
        t0 = time.perf_counter()
        X, info = solve_spectrum(ctx, dms, N, wls[list(idx)], ri_of, ang, precision=prec, keep=keep, pivoting=piv,
                                 lanes=lanes)
        t = {k: float(sum(r[k] for r in info)) for k in ("fill", "factor", "solve")}
        t["clock"] = time.perf_counter() - t0
        t["iters"] = max(r["iters"] for r in info)
        t["berr"] = max(r["berr"] for r in info)
        t["fell_back"] = sum(r["fell_back"] for r in info)
        t["fp64_direct"] = sum(r["precision"] == "fp64" for r in info) if prec == MOM_PREC_MIXED else 0
        return t, X

    for i in range(args.warmup):
        w = sweep(MOM_PREC_MIXED, range(min(nwl, 2 * args.lanes)))[0]
        log("warmup %d (%d wavelengths): %.2f s" % (i, min(nwl, 2 * args.lanes), w["clock"]))
    runs = []
    check = [0, nwl - 1]
    for i in range(args.steps):
        r, X = sweep(MOM_PREC_MIXED, range(nwl), keep=check)
        runs.append(r)
        log("step %d: %.2f s for %d wavelengths (lane sums: fill %.2f factor %.2f solve %.2f; refinement steps <= %d, "
            "berr <= %.1e, %d fell back to FP64 factors, %d factored directly in FP64)"
            % (i, r["clock"], nwl, r["fill"], r["factor"], r["solve"], r["iters"], r["berr"], r["fell_back"],
               r["fp64_direct"]))
    mean = float(np.mean([r["clock"] for r in runs]))
    nrhs = ang.shape[0]
    out = {"metric": "spectrum s @N=%d: %d wavelengths x %d plane waves, fill+LU+solves (mixed path)"
                     % (2 * N, nwl, nrhs),
           "value": round(mean, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(mean * 1e3, 2), "higher_is_better": False, "scaling": "strong",
           "vs_baseline": None, "dtype": "fp32-mfma LU + fp64 fill/residual", "data": "synthetic icosphere ellipsoid",
           "config": {"workload": "ellipsoid 80/40/40 nm from icosphere s=%d, %d faces, N=%d, %d wavelengths "
                                  "400-900 nm, Drude eps_in, %d plane waves per wavelength"
                                  % (s, prob.master.nfaces, 2 * N, nwl, nrhs), "n_unknowns": 2 * N,
                      "complex64_pivoting": args.pivoting, "lanes": args.lanes},
           "lane_sums_s": {k: round(float(np.mean([r[k] for r in runs])), 4) for k in ("fill", "factor", "solve")},
           "fell_back_to_fp64": runs[-1]["fell_back"], "factored_directly_in_fp64": runs[-1]["fp64_direct"],
           "per_wavelength_ms": round(mean / nwl * 1e3, 2)}
    if not args.no_fp64:
        # the all-FP64 path one wavelength at a time (cooperative panels) on eight wavelengths spread over the
        # range (both ends included; the cross-sections of 16 sources at the two ends compared with the mixed
        # solutions), and in lanes (panels one launch per column) over all the wavelengths, as the mixed sweep
        # runs (the lanes' set-up spread over as many wavelengths); the speed-up is stated against the faster
        sub = sorted(set(int(round(v)) for v in np.linspace(0, nwl - 1, min(nwl, 8))))
        t64_1, X64 = sweep(MOM_PREC_FP64, sub, keep=[0, len(sub) - 1], lanes=1)
        w64_1 = t64_1["clock"] / len(sub)
        w64_l = sweep(MOM_PREC_FP64, range(nwl), lanes=args.lanes)[0]["clock"] / nwl if args.lanes > 1 else w64_1
        log("fp64 path: %.3f s per wavelength one at a time (%d wavelengths), %.3f s in %d lanes (%d wavelengths)"
            % (w64_1, len(sub), w64_l, args.lanes, nwl))
        w64 = min(w64_1, w64_l)
        out["fp64_path_per_wavelength_s_by_lanes"] = {"1": round(w64_1, 4), str(args.lanes): round(w64_l, 4)}
        err = 0.0
        for j, i in zip([0, len(sub) - 1], check):
            om = omega_of(wls[i])
            for k in range(0, nrhs, nrhs // 8):
                a = cs_prtsrf(ctx, dms[0], N, om, 1.0, X[i][:, k], *(ang[k] * DEGTORAD))
                b = cs_prtsrf(ctx, dms[0], N, om, 1.0, X64[j][:, k], *(ang[k] * DEGTORAD))
                err = max(err, abs(a[0] / b[0] - 1), abs(a[1] / b[1] - 1))
        out["fp64_path_per_wavelength_s"] = round(w64, 4)
        out["fp64_path_wavelengths"] = {"1": len(sub), str(args.lanes): nwl}
        out["speedup_vs_fp64_path"] = round(w64 / (mean / nwl), 3)
        out["sigma_relerr_vs_fp64_path"] = err
        log("sigma rel err mixed vs fp64 (16 checks) %.1e" % err)
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_adjoint(args):
    """Adjoint sensitivity: sphere of radius 60 nm (icosphere s=4, 5120
    faces, N=15360) with a seeded radial perturbation (l <= 6, 3 nm), eps_in
    = -9.4+1.5i at 600 nm.  One step = forward fill + LU + solve, sigma_abs,
    and its gradient with respect to Re and Im of eps_in by the adjoint (one
    Z^H solve on the same factors and one derivative fill of the interior
    domain), in the mixed path; central differences of the FP64 solve check
    the gradient once."""
    from mom2amd import (Context, DeviceMesh, System, build_problem, cs_prtsrf, load_problem, meshgen, omega_of,
                         sigma_abs_eps_gradient, sigma_abs_vertex_gradient)
    from mom2amd.capi import MOM_PREC_FP64, MOM_PREC_MIXED

    s = args.subdiv if args.subdiv != 5 else 4
    nodes, faces = meshgen.perturbed_sphere(s, 60.0, 3.0)
    d = tempfile.mkdtemp(prefix="mom_bench_")
    path = os.path.join(d, "pert%d.msh" % s)
    meshgen.write_gmsh(path, nodes, faces, 15)
    prob = load_problem(path, 1e-9, [[15], [-15]])
    N = prob.nedges
    omega = omega_of(600e-9)
    eps = -9.4 + 1.5j
    ang = np.array([0.0, 90.0, 0.0]) * np.pi / 180
    log("perturbed sphere s=%d: %d faces, N=%d unknowns" % (s, prob.master.nfaces, 2 * N))
    ctx = Context(0)
    dms = [DeviceMesh(ctx, dm) for dm in prob.domains]

    def forward(e, prec):
        S = System(ctx, N, prec)
        S.set_pivoting(1 if args.pivoting == "tournament" else 0)
        ri = [1.0 + 0j, np.sqrt(e)]
        S.fill(dms, ri, omega)
        S.factor()
        x = S.solve_pw(dms[0], omega, 1.0, ang)[:, 0]
        return S, x, ri

    def step(prec):
        t0 = time.perf_counter()
        S, x, ri = forward(eps, prec)
        t1 = time.perf_counter()
        try:
            sab, dre, dim = sigma_abs_eps_gradient(ctx, S, x, dms, 1, omega, ri)
        finally:
            S.close()
        t2 = time.perf_counter()
        return dict(wall=t2 - t0, forward=t1 - t0, adjoint=t2 - t1, sab=sab, grad=(dre, dim))

    for i in range(args.warmup):
        w = step(MOM_PREC_MIXED)
        log("warmup %d: %.2f s" % (i, w["wall"]))
    runs = []
    for i in range(args.steps):
        r = step(MOM_PREC_MIXED)
        runs.append(r)
        log("step %d: %.3f s (forward fill+LU+solve %.3f, adjoint gradient %.3f); sigma_abs %.6e, "
            "dsigma/deps %.6e %+.6ei" % (i, r["wall"], r["forward"], r["adjoint"], r["sab"], *r["grad"]))
    mean = float(np.mean([r["wall"] for r in runs]))
    out = {"metric": "sigma_abs + adjoint eps gradient s @N=%d (mixed path)" % (2 * N),
           "value": round(mean, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(mean * 1e3, 2), "higher_is_better": False, "scaling": "strong",
           "vs_baseline": None, "dtype": "fp32-mfma LU + fp64 fill/residual", "data": "synthetic perturbed icosphere",
           "config": {"workload": "sphere R=60nm icosphere s=%d with radial l<=6 perturbation of 3 nm, %d faces, "
                                  "N=%d, 600 nm, eps_in=-9.4+1.5i; forward solve + d sigma_abs / d eps_in by the "
                                  "adjoint" % (s, prob.master.nfaces, 2 * N), "n_unknowns": 2 * N,
                      "complex64_pivoting": args.pivoting},
           "phases_s": {k: round(float(np.mean([r[k] for r in runs])), 4) for k in ("forward", "adjoint")},
           "sigma_abs_m2": runs[-1]["sab"], "dsigma_deps": list(runs[-1]["grad"])}
    # the node-coordinate gradient of sigma_abs (all 3V coordinates), reported beside the value
    S, x, ri = forward(eps, MOM_PREC_MIXED)
    try:
        t0 = time.perf_counter()
        sab_v, gv = sigma_abs_vertex_gradient(ctx, S, x, dms, prob, omega, ri, ang)
        tv = time.perf_counter() - t0
    finally:
        S.close()
    out["vertex_gradient"] = {"seconds": round(tv, 3), "n_coordinates": int(gv.size),
                              "max_abs_m": float(np.abs(gv).max())}
    log("vertex gradient (%d node coordinates, one adjoint solve + pair kernel): %.3f s" % (gv.size, tv))
    if not args.no_fp64:
        # central differences (h = 1e-3 nm) of the whole FP64 solve in all three coordinates of
        # 32 seeded nodes, each perturbed mesh classified like the unperturbed one (the gradient
        # holds the near/far split fixed)
        t0 = time.perf_counter()
        rng = np.random.default_rng(20251204)
        verts = rng.choice(nodes.shape[0], 32, replace=False)
        scale = float(np.abs(gv).max())
        fdv, worst = [], 0.0
        for v in verts:
            for c in range(3):
                vals = []
                for sgn in (1.0, -1.0):
                    pn = nodes.copy()
                    pn[v, c] += sgn * 1e-3
                    p2 = build_problem(pn, faces, np.full(faces.shape[0], 15), 1e-9, [[15], [-15]])
                    dms2 = [DeviceMesh(ctx, dm) for dm in p2.domains]
                    S2 = System(ctx, N, MOM_PREC_FP64)
                    try:
                        for dm2, dm in zip(dms2, dms):
                            dm2.classify_like(dm)
                        S2.fill(dms2, [1.0 + 0j, np.sqrt(eps)], omega)
                        S2.factor()
                        x2 = S2.solve_pw(dms2[0], omega, 1.0, ang)[:, 0]
                        vals.append(cs_prtsrf(ctx, dms2[0], N, omega, 1.0, x2, *ang)[1])
                    finally:
                        S2.close()
                        for dm in dms2:
                            dm.close()
                fd_v = (vals[0] - vals[1]) / (2e-3 * 1e-9)
                err = abs(gv[v, c] - fd_v) / scale
                worst = max(worst, err)
                fdv.append([int(v), c, float(gv[v, c]), fd_v])
            log("node %d: adjoint %s fd %s" % (v, gv[v], np.array([f[3] for f in fdv[-3:]])))
        out["vertex_gradient"]["fd_check"] = {
            "nodes": 32, "coordinates": 96, "h_nm": 1e-3, "classification": "frozen to the unperturbed mesh",
            "max_abs_err_over_max_grad": worst, "seconds": round(time.perf_counter() - t0, 1),
            "node_coord_adjoint_fd": fdv}
        log("vertex gradient vs central differences (32 nodes x 3): max |adjoint - fd| / max |grad| = %.2e" % worst)
        fd = []
        for hdir in (1.0, 1j):
            h = 1e-4 * abs(eps) * hdir
            vals = []
            for e in (eps + h, eps - h):
                S, x, _ = forward(e, MOM_PREC_FP64)
                S.close()
                vals.append(cs_prtsrf(ctx, dms[0], N, omega, 1.0, x, *ang)[1])
            fd.append((vals[0] - vals[1]) / (2 * abs(h)))
        g = runs[-1]["grad"]
        out["fd_check"] = {"dsigma_deps_fd": fd,
                           "relerr": max(abs(g[0] - fd[0]), abs(g[1] - fd[1])) / max(abs(fd[0]), abs(fd[1]))}
        log("central differences of the FP64 solve: %.6e %+.6ei (rel err of the adjoint %.1e)"
            % (fd[0], fd[1], out["fd_check"]["relerr"]))
    ctx.close()
    print(json.dumps(out), flush=True)


def bench_symmetric(args):
    """Symmetry decomposition, --group octant (default): the sphere (R = 50
    nm, eps_in = -9.4+1.5i, 600 nm) meshed as an octasphere of 8 n^2 faces,
    n = 50 (2E = 60000 unknowns for the closed surface), solved on its
    octant cell with the group mxp myp mzp: 8 representations of about 7500
    unknowns each.  --group ring: four such spheres (n = 25, 5000 faces
    each) centred 120 nm from the z axis at 0, 90, 180 and 270 degrees,
    solved on the upper half of one (2500 faces) with the group r4 mzp: 8
    representations with complex characters, 60000 unknowns for the whole
    ring.  One step = the fill of all representations (each image
    integrated once), their LU factorisations, the plane-wave sources
    projected on the representations, the solves and cs_prtsrf over the
    images, in the mixed path.  The whole structure solved without symmetry
    (mixed path, once, not in value; each sphere of the ring its own
    interior domain) gives the comparison time and the sigma agreement."""
    from mom2amd import (Context, DeviceMesh, DEGTORAD, SymmetricSystems, System, build_problem, cs_prtsrf,
                         cs_prtsrf_symm, load_problem, meshgen, omega_of, srcvec_symm, symm_group)
    from mom2amd.capi import MOM_PREC_MIXED

    ring = args.group == "ring"
    n = args.cell_n if args.cell_n > 0 else (25 if ring else 50)
    radius, shift = 50.0, (120.0 if ring else 0.0)
    nodes, faces, lines, lids = meshgen.sphere_cell(n, radius, (0, 1) if ring else ())
    nodes = nodes + np.array([shift, 0.0, 0.0])
    d = tempfile.mkdtemp(prefix="mom_bench_")
    path = os.path.join(d, "cell%d.msh" % n)
    meshgen.write_gmsh(path, nodes, faces, 15, lines, lids)
    planes = ("zplane",) if ring else ("xplane", "yplane", "zplane")
    sbnd = [(meshgen.OCTANT_LINE_IDS[c], c) for c in planes]
    prob = load_problem(path, 1e-9, [[15], [-15]], sbnd=sbnd)
    N = prob.nedges
    omega = omega_of(600e-9)
    ri = [1.0 + 0j, np.sqrt(-9.4 + 1.5j)]
    ang = np.array([40.0, 25.0, 30.0]) * DEGTORAD
    gens = ["r4", "mzp"] if ring else ["mxp", "myp", "mzp"]
    nwhole = (4 if ring else 1) * 24 * n * n
    piv = 1 if args.pivoting == "tournament" else 0
    ctx = Context(0)
    dms = [DeviceMesh(ctx, dm) for dm in prob.domains]
    log("%s cell n=%d: %d faces, %d edges (%d on the planes), group %s; whole structure %d faces, %d unknowns"
        % ("ring (upper half sphere)" if ring else "octant", n, prob.master.nfaces, N,
           int((prob.master.bnd > 1).sum()), " ".join(gens), nwhole // 3, nwhole))

    def step():
        t0 = time.perf_counter()
        SS = SymmetricSystems(ctx, N, prob.master.bnd, gens, MOM_PREC_MIXED)
        try:
            for sy in SS.systems:
                sy.set_pivoting(piv)
            SS.fill(dms, ri, omega)
            t1 = time.perf_counter()
            SS.factor()
            t2 = time.perf_counter()
            q = srcvec_symm(ctx, dms[0], N, omega, ri[0], gens, ang)[:, :, 0]
            x = SS.solve(q)[:, :, 0]
            cs = cs_prtsrf_symm(ctx, dms[0], N, omega, ri[0], gens, x, *ang)
            t3 = time.perf_counter()
            return dict(wall=t3 - t0, fill=t1 - t0, factor=t2 - t1, solve=t3 - t2, cs=cs, dims=list(SS.dims),
                        berr=max(SS.berr))
        finally:
            SS.close()

    for i in range(args.warmup):
        w = step()
        log("warmup %d: %.3f s" % (i, w["wall"]))
    runs = []
    for i in range(args.steps):
        r = step()
        runs.append(r)
        log("step %d: %.3f s (fill %.3f, factor %.3f, sources+solve+cs %.3f); csca %.9e cabs %.9e, berr %.1e"
            % (i, r["wall"], r["fill"], r["factor"], r["solve"], r["cs"][0], r["cs"][1], r["berr"]))
    mean = float(np.mean([r["wall"] for r in runs]))
    if ring:
        metric = "fill+LU+solve s, ring of four spheres (%d unknowns) by r4 x mzp symmetry (mixed path)" % nwhole
        wl = ("four spheres R=50nm (octaspheres n=%d, %d faces each) centred 120 nm from the z axis at 0/90/180/270 "
              "deg, upper half cell %d faces, group r4 mzp (8 representations of %s unknowns), 600 nm, "
              "eps_in=-9.4+1.5i" % (n, 8 * n * n, prob.master.nfaces, runs[-1]["dims"]))
    else:
        metric = "fill+LU+solve s, octant of a %d-unknown sphere by mirror symmetry (mixed path)" % nwhole
        wl = ("sphere R=50nm as octasphere n=%d (%d faces), octant cell %d faces, group mxp myp mzp (8 "
              "representations of %s unknowns), 600 nm, eps_in=-9.4+1.5i"
              % (n, 8 * n * n, prob.master.nfaces, runs[-1]["dims"]))
    out = {"metric": metric,
           "value": round(mean, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(mean * 1e3, 2), "higher_is_better": False, "scaling": "strong",
           "vs_baseline": None, "dtype": "fp32-mfma LU + fp64 fill/residual", "data": "synthetic octasphere",
           "config": {"workload": wl, "group": " ".join(gens), "n_unknowns_closed": nwhole,
                      "complex64_pivoting": args.pivoting},
           "phases_s": {k: round(float(np.mean([r[k] for r in runs])), 4) for k in ("fill", "factor", "solve")},
           "csca_m2": runs[-1]["cs"][0], "cabs_m2": runs[-1]["cs"][1]}
    if not args.no_fp64:
        if ring:
            # the four spheres, each its own interior domain (the decomposition takes the rotation images in the
            # exterior domain alone)
            sn, sf, _, _ = meshgen.sphere_cell(n, radius, (0, 1, 2))
            sn = sn + np.array([shift, 0.0, 0.0])
            Js = symm_group(["r4"])["J"]
            fn = np.concatenate([sn @ J.T for J in Js])
            ff = np.concatenate([sf + k * sn.shape[0] for k in range(4)])
            fid = np.repeat(15 + np.arange(4), sf.shape[0])
            P = build_problem(fn, ff, fid, 1e-9, [[15, 16, 17, 18], [-15], [-16], [-17], [-18]])
            riw = [ri[0]] + [ri[1]] * 4
        else:
            fn, ff = meshgen.octasphere(n, 50.0)
            P = build_problem(fn, ff, np.full(ff.shape[0], 15), 1e-9, [[15], [-15]])
            riw = ri
        fdms = [DeviceMesh(ctx, dm) for dm in P.domains]
        t0 = time.perf_counter()
        S = System(ctx, P.nedges, MOM_PREC_MIXED)
        try:
            S.set_pivoting(piv)
            S.fill(fdms, riw, omega)
            S.factor()
            xf = S.solve_pw(fdms[0], omega, ri[0], ang)[:, 0]
            csf = cs_prtsrf(ctx, fdms[0], P.nedges, omega, ri[0], xf, *ang)
        finally:
            S.close()
        tf = time.perf_counter() - t0
        for dm in fdms:
            dm.close()
        cs = runs[-1]["cs"]
        out["closed_sphere"] = {"seconds": round(tf, 4), "n_unknowns": 2 * P.nedges, "speedup": round(tf / mean, 2),
                                "csca_relerr": abs(cs[0] / csf[0] - 1), "cabs_relerr": abs(cs[1] / csf[1] - 1)}
        log("whole %s (N=%d, mixed path): %.3f s, %.1fx the symmetric solve; sigma rel diff %.1e / %.1e"
            % ("ring" if ring else "octasphere", 2 * P.nedges, tf, tf / mean, out["closed_sphere"]["csca_relerr"],
               out["closed_sphere"]["cabs_relerr"]))
    ctx.close()
    print(json.dumps(out), flush=True)


def roofline_fill(fp, steps, n):
    """Roofline block of the fill's pair kernel (FP64 VALU) and its assembly
    passes (HBM), from the event times of the timed steps."""
    if not fp or fp.get("pair_launches", 0) == 0:
        return None
    pair_s = fp["pair_ms"] / 1e3
    points = fp["pair_points"]
    ach = FILL_FLOPS_PER_POINT_PAIR * points / pair_s / 1e12
    out = {"kernel": "pair_plane_kernel<13> (far and smoothed near face pairs of one slab of 1024 source faces "
                     "against the test faces m >= the slab start; each unordered pair once); achieved = 60 "
                     "algorithmic flops per evaluated (test point, source point) pair / kernel time (HIP events "
                     "around each launch)",
           "bound": "valu-fp64", "achieved": round(ach, 3), "peak": PEAK_FP64_VALU, "unit": "TFLOP/s",
           "frac": round(ach / PEAK_FP64_VALU, 4), "launches": int(fp["pair_launches"]),
           "avg_launch_ms": round(fp["pair_ms"] / fp["pair_launches"], 4),
           "point_pairs_per_solve": points / steps,
           "pair_kernel_s_per_solve": round(pair_s / steps, 4)}
    if FILL_ISSUED_FLOPS_PER_POINT_PAIR:
        iss = FILL_ISSUED_FLOPS_PER_POINT_PAIR * points / pair_s / 1e12
        out["issued_fp64_flops_per_point_pair"] = FILL_ISSUED_FLOPS_PER_POINT_PAIR
        out["issued_tflops"] = round(iss, 3)
        out["issued_frac"] = round(iss / PEAK_FP64_VALU, 4)
    asm_s = fp["assemble_ms"] / 1e3 / steps
    floor = 16.0 * n * n  # complex128 Z written once
    out["assembly"] = {"kernels": "assemble_kernel + assemble_transposed_kernel (face-side planes of a slab "
                                  "gathered into Z, one read-modify-write per entry and slab)",
                       "s_per_solve": round(asm_s, 4), "launch_groups_per_solve": int(fp["assemble_launches"] / steps),
                       "write_once_floor_bytes": floor,
                       "floor_s_at_peak": round(floor / (PEAK_HBM * 1e9), 5)}
    if ASSEMBLY_TRAFFIC_S5_BYTES and n == 61440:
        out["assembly"]["traffic_bytes_per_solve"] = ASSEMBLY_TRAFFIC_S5_BYTES
        out["assembly"]["achieved_GBs"] = round(ASSEMBLY_TRAFFIC_S5_BYTES / asm_s / 1e9, 1)
        out["assembly"]["traffic_over_floor"] = round(ASSEMBLY_TRAFFIC_S5_BYTES / floor, 2)
    out["near_part_s_per_solve"] = round(fp["near_ms"] / 1e3 / steps, 4)
    return out


def bench_fill(args):
    """The complex128 PMCHWT fill of the sphere workload alone (no LU, no
    solve), timed per step, with the pair kernel's and assembly's event
    times: the workload for counter runs of the fill kernels."""
    from mom2amd import Context, DeviceMesh, System
    from mom2amd.capi import MOM_PREC_FP64

    case = build_case(args.subdiv)
    N = case["problem"].nedges
    log("fill workload s=%d: N=%d unknowns" % (args.subdiv, 2 * N))
    ctx = Context(0)
    dms = [DeviceMesh(ctx, d) for d in case["problem"].domains]
    S = System(ctx, N, MOM_PREC_FP64)
    for _ in range(args.warmup):
        S.fill(dms, case["ri"], case["omega"])
    ctx.fill_profile(1)
    t0 = time.perf_counter()
    for i in range(args.steps):
        S.fill(dms, case["ri"], case["omega"])
        log("fill step %d done" % i)
    wall = (time.perf_counter() - t0) / args.steps
    fp = ctx.fill_profile(0)
    S.close()
    ctx.close()
    print(json.dumps({"metric": "PMCHWT fill s @N=%d" % (2 * N), "value": round(wall, 4), "unit": "s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 2),
                      "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
                      "data": "synthetic icosphere mesh",
                      "config": {"workload": "complex128 PMCHWT fill, sphere R=50nm icosphere s=%d, N=%d, 600 nm"
                                             % (args.subdiv, 2 * N), "n_unknowns": 2 * N},
                      "roofline_fill": roofline_fill(fp, args.steps, 2 * N)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--subdiv", type=int, default=5, help="icosphere subdivision (5: N = 61440)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the all-FP64 comparison path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--workload", choices=("pmchwt", "sh", "spectrum", "adjoint", "symmetric", "fill"),
                    default="pmchwt",
                    help="pmchwt: the sphere solve (default); sh: the second-harmonic dimer chain; spectrum: the "
                         "ellipsoid over many wavelengths with 256 plane waves each; adjoint: sigma_abs and its "
                         "permittivity gradient on a perturbed sphere (sh, spectrum and adjoint use s=4 unless "
                         "--subdiv is given); symmetric: the octant of an octasphere by mirror symmetry; fill: the "
                         "complex128 PMCHWT fill of the default sphere alone (for counter runs)")
    ap.add_argument("--cell-n", type=int, default=0,
                    help="octasphere divisions of the symmetric workload (default 50 for octant, 25 for ring)")
    ap.add_argument("--group", choices=("octant", "ring"), default="octant",
                    help="symmetric workload: the octant of one sphere (mxp myp mzp) or a ring of four spheres "
                         "(r4 mzp)")
    ap.add_argument("--nwl", type=int, default=64, help="wavelengths of the spectrum workload")
    ap.add_argument("--lanes", type=int, default=4, help="wavelengths solved at once in the spectrum workload")
    ap.add_argument("--pmchwt-pivoting", choices=("partial", "tournament"), default="tournament",
                    help="complex64 panels of the pmchwt workload's mixed path (the FP64 path always pivots as LAPACK)")
    ap.add_argument("--pivoting", choices=("tournament", "partial"), default="tournament",
                    help="complex64 panels of the mixed path in the spectrum, sh and adjoint workloads (the library "
                         "default is partial pivoting)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "sh":
        bench_sh(args)
        return
    if args.workload == "spectrum":
        bench_spectrum(args)
        return
    if args.workload == "adjoint":
        bench_adjoint(args)
        return
    if args.workload == "symmetric":
        bench_symmetric(args)
        return
    if args.workload == "fill":
        bench_fill(args)
        return

    from mom2amd import Context, DeviceMesh
    from mom2amd.capi import MOM_PREC_FP64, MOM_PREC_MIXED

    # the vendor library's complex128 GEMM, timed before this library takes the GPU
    lib_tf = library_zgemm_tflops() if not args.no_fp64 else None
    t0 = time.perf_counter()
    case = build_case(args.subdiv)
    N = case["problem"].nedges
    log("mesh s=%d: %d faces, N=%d unknowns (set-up %.1f s)" % (args.subdiv, case["problem"].domains[0].nfaces,
                                                                 2 * N, time.perf_counter() - t0))
    ctx = Context(0)
    dms = [DeviceMesh(ctx, d) for d in case["problem"].domains]
    piv = 1 if args.pmchwt_pivoting == "tournament" else 0
    for i in range(args.warmup):
        st, _ = run_path(ctx, dms, case, MOM_PREC_MIXED, pivoting=piv)
        log("warmup %d: %.2f s (fill %.2f factor %.2f solve %.2f)" % (i, st["wall_s"], st["fill_s"], st["factor_s"],
                                                                     st["solve_s"]))
    times, gemm_s, gemm_flops, gemm_launches, last = [], 0.0, 0.0, 0, None
    fprof = {}
    for i in range(args.steps):
        # timed steps: HIP events around the depth-256 trailing updates only
        st, cs_mixed = run_path(ctx, dms, case, MOM_PREC_MIXED, profile=1, pivoting=piv)
        times.append(st["wall_s"])
        gemm_s += st["gemm_s"]
        gemm_flops += st["gemm_flops"]
        gemm_launches += int(st["gemm_launches"])
        for k, v in st["fill_prof"].items():
            fprof[k] = fprof.get(k, 0.0) + v
        last = st
        log("step %d: %.2f s (fill %.2f factor %.2f solve %.2f, IR iters %d (%d factor solves, %d FP64 products), "
            "berr %s); trailing-update gemm %.2f"
            % (i, st["wall_s"], st["fill_s"], st["factor_s"], st["solve_s"], st["iters"], st["n_minv"],
               st["n_matvec"], " ".join("%.1e" % v for v in st["berr_history"]), st["gemm_s"]))
    mixed_s = float(np.mean(times))
    # one more, untimed step with every kernel group timed, for the factor breakdown
    bd, _ = run_path(ctx, dms, case, MOM_PREC_MIXED, profile=2, pivoting=piv)
    log("breakdown step (all groups timed, not in value): %.2f s; factor %.2f: gemm %.2f (narrow %.2f) panel %.2f "
        "trsm %.2f laswp %.2f" % (bd["wall_s"], bd["factor_s"], bd["gemm_s"], bd["gemm_narrow_s"], bd["panel_s"],
                                  bd["trsm_s"], bd["laswp_s"]))

    fp64 = None
    if not args.no_fp64:
        st64, cs64 = run_path(ctx, dms, case, MOM_PREC_FP64, profile=1)
        fp64 = st64
        log("fp64 path: %.2f s (fill %.2f factor %.2f solve %.2f); trailing-update gemm %.2f"
            % (st64["wall_s"], st64["fill_s"], st64["factor_s"], st64["solve_s"], st64["gemm_s"]))
    achieved = gemm_flops / gemm_s / 1e12 if gemm_s > 0 else None
    out = {
        "metric": METRIC,
        "value": round(mixed_s, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed_s * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "fp32-mfma LU + fp64 fill/residual",
        "data": "synthetic icosphere mesh",
        "config": {"workload": "PMCHWT sphere R=50nm icosphere s=%d (%d faces), N=%d unknowns, 600 nm, "
                               "eps_in=-9.4+1.5i, 1 plane wave; fill+LU+solve, GMRES-IR to backward err 1e-12"
                               % (args.subdiv, case["problem"].domains[0].nfaces, 2 * N),
                   "n_unknowns": 2 * N, "complex64_pivoting": args.pmchwt_pivoting},
        "phases_s": {"fill": round(last["fill_s"], 4), "factor": round(last["factor_s"], 4),
                     "solve_ir": round(last["solve_s"], 4), "ir_iters": last["iters"], "ir_factor_solves": last["n_minv"],
                     "ir_fp64_products": last["n_matvec"], "berr": last["berr"]},
        "factor_breakdown_s": dict({k: round(bd[k + "_s"], 4) for k in ("factor", "gemm", "gemm_narrow", "panel",
                                                                         "trsm", "laswp")},
                                   note="an extra untimed step with every kernel group event-timed"),
        "fill_tuning": {"pair_kernel_sources_per_wave": ctx.fill_tuning()[0],
                        "first_slab_ms_by_candidate": {str(k): round(v, 3) for k, v in ctx.fill_tuning()[1].items()}},
        "gemm_narrow": {"launches": bd["gemm_narrow_launches"],
                        "tflops": round(bd["gemm_narrow_flops"] / max(bd["gemm_narrow_s"], 1e-12) / 1e12, 2)},
        "roofline": None if achieved is None else {
            "kernel": "gemm3m_dma_kernel<32> (LU trailing update A22 -= L21 U12 of depth 256: the next block "
                      "column on the look-ahead stream and the remaining columns on the main stream; complex64 as "
                      "three real products on v_mfma_f32_32x32x2f32); achieved = 8MNK algorithmic flops / kernel "
                      "time (HIP events around each launch)",
            "bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP32_MFMA, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_MFMA, 4),
            "frac_note": "the 3M product issues 6MNK real flops for the 8MNK of a complex product, so the 8MNK rate "
                         "may exceed the FP32 matrix peak; matrix-core utilisation is mfma_issued_frac",
            "traffic": TRAFFIC_S5_BYTES if args.subdiv == 5 else None, "traffic_unit": "bytes/launch",
            "mfma_issued_frac": round(achieved * 6 / 8 / PEAK_FP32_MFMA, 4),
            "launches": gemm_launches, "avg_launch_ms": round(gemm_s / max(gemm_launches, 1) * 1e3, 4)},
    }
    out["roofline_fill"] = roofline_fill(fprof, args.steps, 2 * N)
    if fp64 is not None:
        ach64 = fp64["gemm_flops"] / fp64["gemm_s"] / 1e12 if fp64["gemm_s"] > 0 else None
        out["fp64_path"] = {"value_s": round(fp64["wall_s"], 4), "fill": round(fp64["fill_s"], 4),
                            "factor": round(fp64["factor_s"], 4), "solve": round(fp64["solve_s"], 4),
                            "trailing_update_gemm_s": round(fp64["gemm_s"], 4),
                            "gemm_tflops": None if ach64 is None else round(ach64, 3),
                            "gemm_frac_of_fp64_peak": None if ach64 is None else round(ach64 / PEAK_FP64_MFMA, 4)}
        out["speedup_vs_fp64_path"] = round(fp64["wall_s"] / mixed_s, 3)
        if lib_tf is not None and ach64:
            # the FP64 path with its trailing updates at the library GEMM's rate instead of this kernel's,
            # whichever is faster, so that the speed-up is stated against the faster FP64 path
            alt = fp64["wall_s"] - fp64["gemm_s"] + fp64["gemm_s"] * ach64 / lib_tf
            out["fp64_gemm_yardstick"] = {
                "library": "torch.addmm complex128 (hipBLAS/rocBLAS), C -= A B, M=N=16384, K=256",
                "tflops": round(lib_tf, 3), "this_kernel_tflops_in_factorisation": round(ach64, 3),
                "fp64_path_s_at_faster_gemm": round(min(fp64["wall_s"], alt), 4),
                "speedup_vs_faster_fp64_path": round(min(fp64["wall_s"], alt) / mixed_s, 3)}
            log("library zgemm yardstick: %.1f TFLOP/s (this kernel %.1f in the factorisation)" % (lib_tf, ach64))
        out["sigma_relerr_vs_fp64_path"] = {"sca": abs(cs_mixed[0] / cs64[0] - 1), "abs": abs(cs_mixed[1] / cs64[1] - 1)}
        out["rcs_relerr_vs_fp64_path"] = float(np.abs(last["rcs"] - st64["rcs"]).max() / np.abs(st64["rcs"]).max())
        out["sigma"] = {"sca_m2": cs_mixed[0], "abs_m2": cs_mixed[1]}
    ctx.close()
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        log("cpu baseline (bounded sample, %d threads)" % threads)
        cb = cpu_baseline(case, args.subdiv, threads)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["cpu_baseline"]["value"] = round(cb["value"], 1)
        out["cpu_baseline"]["fill_s"] = round(cb["fill_s"], 1)
        out["cpu_baseline"]["lu_s"] = round(cb["lu_s"], 1)
        out["speedup_vs_cpu"] = round(cb["value"] / mixed_s, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
