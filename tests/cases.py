"""Problem set-ups shared by the tests (meshes written once per session)."""
import os

import numpy as np

from mom2amd import meshgen, build_problem, load_problem, omega_of

N_IN_SPHERE = np.sqrt(-9.4 + 1.5j)
N_IN_DIMER = np.sqrt(-40 + 2.5j)


def sphere(msh_dir, s, radius=50.0):
    nodes, faces = meshgen.icosphere(s, radius)
    path = os.path.join(str(msh_dir), "ico%d_r%g.msh" % (s, radius))
    if not os.path.exists(path):
        meshgen.write_gmsh(path, nodes, faces, 15)
    doms = [[15], [-15]]
    return dict(path=path, domains=doms, problem=build_problem(*meshgen.read_gmsh(path), 1e-9, doms),
                wl=600e-9, omega=omega_of(600e-9), ri=[1.0 + 0j, complex(N_IN_SPHERE)])


def dimer(msh_dir, s=1):
    nodes, faces, ids = meshgen.dimer(s, 40.0, 10.0)
    path = os.path.join(str(msh_dir), "dimer%d.msh" % s)
    if not os.path.exists(path):
        meshgen.write_gmsh(path, nodes, faces, ids)
    doms = [[15, 16], [-15], [-16]]
    return dict(path=path, domains=doms, problem=build_problem(*meshgen.read_gmsh(path), 1e-9, doms),
                wl=1060e-9, omega=omega_of(1060e-9),
                ri=[1.0 + 0j, complex(N_IN_DIMER), complex(N_IN_DIMER)])


def blockwise_relerr(A, B, nblk=2):
    """max over the 2x2 PMCHWT blocks of max|A-B| / max|B| in that block."""
    n = A.shape[0] // nblk
    err = 0.0
    for i in range(nblk):
        for j in range(nblk):
            a = A[i * n:(i + 1) * n, j * n:(j + 1) * n]
            b = B[i * n:(i + 1) * n, j * n:(j + 1) * n]
            err = max(err, np.abs(a - b).max() / np.abs(b).max())
    return err


def sphere_cell(msh_dir, n, reflect=(), radius=50.0, shift=0.0):
    """Symmetry cell of the octasphere (meshgen.sphere_cell, moved by `shift`
    along x) with its boundary edges classified (`sbnd`), at the sphere
    case's wavelength and permittivity."""
    nodes, faces, lines, lids = meshgen.sphere_cell(n, radius, reflect)
    nodes = nodes + np.array([shift, 0.0, 0.0])
    tag = "".join(str(a) for a in reflect)
    path = os.path.join(str(msh_dir), "cell%d_%s_r%g_x%g.msh" % (n, tag, radius, shift))
    if not os.path.exists(path):
        meshgen.write_gmsh(path, nodes, faces, 15, lines, lids)
    sbnd = [(i, c) for c, i in meshgen.OCTANT_LINE_IDS.items() if i in set(int(v) for v in lids)]
    doms = [[15], [-15]]
    return dict(path=path, domains=doms, sbnd=sbnd, problem=load_problem(path, 1e-9, doms, sbnd=sbnd),
                wl=600e-9, omega=omega_of(600e-9), ri=[1.0 + 0j, complex(N_IN_SPHERE)])
