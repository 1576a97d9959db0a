"""Generate the golden vectors under tests/golden/ from the reference Fortran.

Runs oracle/_ref/mom_ref (the reference's own modules, built by
oracle/build_ref.sh) on small synthetic problems and stores inputs and
outputs as .npz fixtures.  The solution vector fed to the reference's
post-processing (cs_prtsrf, rcs) is computed here with LAPACK (scipy), which
is what the reference's solve_multi_linsys calls (linalg.f90:159-180).

Fixtures (all complex128 / float64 at full precision):
  sphere_s1.npz  icosphere s=1, R=50 nm, eps_in=-9.4+1.5i, 600 nm:
                 full A (240x240), q, x, csca, cabs, rcs 37x72, mesh arrays
  sphere_s2.npz  icosphere s=2: sampled A entries, block norms, q, x, csca,
                 cabs, rcs 37x72, mesh index arrays
  dimer_s1.npz   two s=1 spheres (ids 15, 16), 3 domains, eps_in=-40+2.5i at
                 1060 nm: sampled A entries, q, x, csca, cabs, mesh arrays;
                 second harmonic (surface chi2 nnn = 1 on both spheres, SH
                 index sqrt(-9.4+1.5i)): nlsurf_coef's src_coef/src_vec and
                 epsp per sphere, the SH source before and after
                 sysmat_pmchwt_nls, the SH solution and its rcs 37x72
  ops_s1.npz     computeD, computeH and computeK of both domains of the
                 icosphere s=1 (E = 120 each), as the reference's driver dumps
                 them one by one (dump_ops)
  spectrum_s2.npz  the batched-spectrum workload at small size: icosphere s=2
                 scaled to the 80/40/40 nm ellipsoid, 3 of the 64 wavelengths
                 of `wlrg 64 400d-9 900d-9` (n = 1, 32, 64) with the Drude
                 index, 8 plane waves (4 seeded directions x psi 0, 90 deg),
                 all solved together by the reference's own solve_systems:
                 per wavelength q, x, csca, cabs, the rcs 37x72 of source 1,
                 sampled A entries and block norms
  trimer_r3.npz, ring_r4_mzp.npz  rotation groups (main_rot)
Usage: python tests/golden/make_golden.py [symm|rot|spectrum|ops]   (only those fixtures)
"""
import os
import sys

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "for90-mom2_amd"))

import refrun  # noqa: E402
from mom2amd import meshgen  # noqa: E402
from mom2amd.spectrum import drude_index, plane_waves_deg, wavelengths  # noqa: E402

PI_F32 = float(np.float32(3.141592653589))
C0 = 2.997924580003e8


def mesh_fields(m, full):
    keys = ["fnodes", "fedges", "efaces", "ebn", "enodes", "eparent"]
    if full:
        keys += ["p", "fn", "fcp", "fs", "fm", "farea", "elen"]
    out = {k: m[k] for k in keys}
    out["avelen"] = np.array(m["avelen"])
    return out


def case(name, mshfile, domains, wl, ri, angles, full_A, nsample=3000, seed=1, sh=None):
    R = refrun.run(mshfile, 1e-9, domains, wl, ri, angles, dump_ops=False, dump_sys=True)
    A, q = R["A"], R["q"]
    x = sl.solve(A, q)
    P = refrun.run(mshfile, 1e-9, domains, wl, ri, angles, dump_ops=False, dump_sys=False, x=x,
                   ntheta=37, nphi=72, sh=sh)
    N = A.shape[0]
    data = dict(wl=np.array(wl), ri=np.array(ri, dtype=np.complex128), angles_deg=np.array(angles, float),
                domains=np.array([d + [0] * (4 - len(d)) for d in domains]), q=q, x=x,
                csca=np.array(P["csca"]), cabs=np.array(P["cabs"]), rcs=P["rcs"])
    for d, m in enumerate(R["mesh"]):
        for k, v in mesh_fields(m, full_A).items():
            data["mesh%d_%s" % (d, k)] = v
    if sh is not None:
        data["sh_shri"] = np.array(sh["shri"], dtype=np.complex128)
        data["sh_nls"] = np.array(sh["nls"], dtype=np.int32)
        data["sh_chi2"] = np.array(sh["chi2"], dtype=np.complex128)
        for k, v in P.items():
            if k.startswith("sh_"):
                data[k] = v
    E = N // 2
    data["block_fro"] = np.array([[np.linalg.norm(A[i * E:(i + 1) * E, j * E:(j + 1) * E]) for j in range(2)]
                                  for i in range(2)])
    if full_A:
        data["A"] = A
    else:
        rng = np.random.default_rng(seed)
        ii = rng.integers(0, N, nsample)
        jj = rng.integers(0, N, nsample)
        ii = np.concatenate([ii, np.arange(N)])
        jj = np.concatenate([jj, np.arange(N)])
        data["A_i"], data["A_j"], data["A_v"] = ii.astype(np.int32), jj.astype(np.int32), A[ii, jj]
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **data)
    print(name, "N=%d" % N, "csca=%.6e cabs=%.6e" % (P["csca"], P["cabs"]))


def symm_case(name, n, reflect, gens, angles, with_A, radius=50.0, shift=0.0):
    """Symmetry cell (meshgen.sphere_cell of radius `radius`, moved by
    `shift` along x) solved by the reference per representation
    (oracle/_ref/mom_ref_symm)."""
    tmp = os.path.join(HERE, "_msh")
    nodes, faces, lines, lids = meshgen.sphere_cell(n, radius, reflect)
    nodes = nodes + np.array([shift, 0.0, 0.0])
    fn = os.path.join(tmp, name + ".msh")
    meshgen.write_gmsh(fn, nodes, faces, 15, lines, lids)
    sbnd = [(i, c) for c, i in meshgen.OCTANT_LINE_IDS.items() if i in set(int(v) for v in lids)]
    ri = [1.0 + 0j, np.sqrt(-9.4 + 1.5j)]
    R = refrun.run_symm(fn, 1e-9, gens, sbnd, [[15], [-15]], 600e-9, ri, angles, ntheta=37, nphi=72)
    data = dict(n=np.array(n), reflect=np.array(list(reflect) + [-1] * (3 - len(reflect))),
                radius=np.array(radius), shift=np.array(shift),
                gens=np.array(gens), wl=np.array(600e-9), ri=np.array(ri, dtype=np.complex128),
                angles_deg=np.array(angles, float), q=R["q"], x=R["x"], csca=np.array(R["csca"]),
                cabs=np.array(R["cabs"]), rcs=R["rcs"])
    for d, m in enumerate(R["mesh"]):
        for k, v in mesh_fields(m, False).items():
            data["mesh%d_%s" % (d, k)] = v
        data["mesh%d_bnd" % d] = m["bnd"]
    if with_A:
        data["A"] = R["A"]
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **data)
    print(name, "N=%d nga=%d" % (R["q"].shape[0] // 2, R["q"].shape[1]),
          "csca=%.6e cabs=%.6e" % (R["csca"], R["cabs"]))


def main_symm():
    """Fixtures of the mirror-symmetric solve:
      octant_n4.npz  octant (x, y, z >= 0) of the n=4 octasphere, R=50 nm,
                     group mxp myp mzp (8 representations), incidence
                     (40, 25, 30) deg: A[60,60,8] before reduction, q, the
                     expanded solutions x, csca, cabs, rcs 37x72, mesh
                     arrays and edge classes
      quarter_n4.npz the n=4 cell x, y >= 0 (both z halves), group myp mxp
                     (4 representations, generators in that order): q, x,
                     csca, cabs, rcs, mesh arrays and edge classes"""
    os.makedirs(os.path.join(HERE, "_msh"), exist_ok=True)
    symm_case("octant_n4", 4, (), ["mxp", "myp", "mzp"], (40.0, 25.0, 30.0), True)
    symm_case("quarter_n4", 4, (2,), ["myp", "mxp"], (40.0, 25.0, 30.0), False)
    main_rot()


def main_rot():
    """Fixtures of the rotation groups (group_rz; bodies that no rotation cut
    crosses):
      trimer_r3.npz  the n=2 octasphere of radius 20 nm centred at (35, 0, 0)
                     nm, group r3 (three copies around z, 3 complex
                     representations), incidence (40, 25, 30) deg: A before
                     reduction, q, the expanded solutions x, csca, cabs, rcs
                     37x72, mesh arrays and edge classes
      ring_r4_mzp.npz the upper half (z >= 0) of the same sphere, group r4 mzp
                     (four spheres around z from eight half cells, 8
                     representations): the same outputs"""
    os.makedirs(os.path.join(HERE, "_msh"), exist_ok=True)
    symm_case("trimer_r3", 2, (0, 1, 2), ["r3"], (40.0, 25.0, 30.0), True, radius=20.0, shift=35.0)
    symm_case("ring_r4_mzp", 2, (0, 1), ["r4", "mzp"], (40.0, 25.0, 30.0), True, radius=20.0, shift=35.0)


def main_ops():
    """computeD/H/K of each domain of the s = 1 sphere (sysmat.f90:272-519)."""
    tmp = os.path.join(HERE, "_msh")
    os.makedirs(tmp, exist_ok=True)
    nodes, faces = meshgen.icosphere(1, 50.0)
    fn = os.path.join(tmp, "ico1.msh")
    meshgen.write_gmsh(fn, nodes, faces, 15)
    ri = [1.0 + 0j, np.sqrt(-9.4 + 1.5j)]
    R = refrun.run(fn, 1e-9, [[15], [-15]], 600e-9, ri, (0.0, 90.0, 0.0), dump_ops=True, dump_sys=False)
    data = dict(wl=np.array(600e-9), ri=np.array(ri, dtype=np.complex128))
    for d in range(2):
        for op in "DHK":
            data["op_%d_%s" % (d, op)] = R["op_%d_%s" % (d, op)]
    np.savez_compressed(os.path.join(HERE, "ops_s1.npz"), **data)
    print("ops_s1", {k: v.shape for k, v in data.items() if k.startswith("op_")})


SPECTRUM_WL_INDEX = (0, 31, 63)


def main_spectrum(nsample=2000, seed=3):
    """The spectrum fixture: one reference run per wavelength, every source
    solved at once (solver.f90:123-136) and post-processed per source."""
    tmp = os.path.join(HERE, "_msh")
    os.makedirs(tmp, exist_ok=True)
    nodes, faces = meshgen.icosphere(2, 1.0)
    nodes = nodes * np.array([80.0, 40.0, 40.0])
    fn = os.path.join(tmp, "ellipsoid2.msh")
    meshgen.write_gmsh(fn, nodes, faces, 15)
    wls = wavelengths(64)[list(SPECTRUM_WL_INDEX)]
    ang = plane_waves_deg(128)
    ang = np.concatenate([ang[:4], ang[128:132]])
    data = dict(wl=wls, wl_index=np.array(SPECTRUM_WL_INDEX), angles_deg=ang)
    rng = np.random.default_rng(seed)
    for i, wl in enumerate(wls):
        ri = [1.0 + 0j, drude_index(wl)]
        R = refrun.run(fn, 1e-9, [[15], [-15]], wl, ri, ang, dump_ops=False, dump_sys=True, solve=True,
                       ntheta=37, nphi=72)
        A = R["A"]
        N = A.shape[0]
        E = N // 2
        data["ri_%d" % i] = np.array(ri, dtype=np.complex128)
        data["q_%d" % i] = R["q"]
        data["x_%d" % i] = R["xref"]
        data["csca_%d" % i], data["cabs_%d" % i] = R["csca"], R["cabs"]
        data["rcs0_%d" % i] = R["rcs"][:, :, 0]
        data["block_fro_%d" % i] = np.array([[np.linalg.norm(A[a * E:(a + 1) * E, b * E:(b + 1) * E])
                                              for b in range(2)] for a in range(2)])
        ii = np.concatenate([rng.integers(0, N, nsample), np.arange(N)])
        jj = np.concatenate([rng.integers(0, N, nsample), np.arange(N)])
        data["A_i_%d" % i], data["A_j_%d" % i], data["A_v_%d" % i] = ii.astype(np.int32), jj.astype(np.int32), A[ii, jj]
        print("spectrum wl=%.1f nm n_in=%s N=%d csca[0]=%.6e" % (wl * 1e9, ri[1], N, R["csca"][0]))
    for d, m in enumerate(R["mesh"]):
        for k, v in mesh_fields(m, False).items():
            data["mesh%d_%s" % (d, k)] = v
    np.savez_compressed(os.path.join(HERE, "spectrum_s2.npz"), **data)


def main():
    if not refrun.available():
        raise SystemExit("oracle/_ref/mom_ref is not built (run oracle/build_ref.sh)")
    tmp = os.path.join(HERE, "_msh")
    os.makedirs(tmp, exist_ok=True)
    n_in = np.sqrt(-9.4 + 1.5j)
    for s, full in ((1, True), (2, False)):
        nodes, faces = meshgen.icosphere(s, 50.0)
        fn = os.path.join(tmp, "ico%d.msh" % s)
        meshgen.write_gmsh(fn, nodes, faces, 15)
        case("sphere_s%d" % s, fn, [[15], [-15]], 600e-9, [1.0 + 0j, n_in], (0.0, 90.0, 0.0), full)
    nodes, faces, ids = meshgen.dimer(1, 40.0, 10.0)
    fn = os.path.join(tmp, "dimer1.msh")
    meshgen.write_gmsh(fn, nodes, faces, ids)
    n_sh = np.sqrt(-9.4 + 1.5j)
    case("dimer_s1", fn, [[15, 16], [-15], [-16]], 1060e-9, [1.0 + 0j, np.sqrt(-40 + 2.5j), np.sqrt(-40 + 2.5j)],
         (0.0, 90.0, 0.0), False,
         sh=dict(shri=[1.0 + 0j, n_sh, n_sh], nls=[0, 1, 1], chi2=(1.0 + 0j, 0j, 0j)))


if __name__ == "__main__":
    if sys.argv[1:] == ["symm"]:
        main_symm()
    elif sys.argv[1:] == ["rot"]:
        main_rot()
    elif sys.argv[1:] == ["spectrum"]:
        main_spectrum()
    elif sys.argv[1:] == ["ops"]:
        main_ops()
    else:
        main()
        main_symm()
        main_spectrum()
        main_ops()
