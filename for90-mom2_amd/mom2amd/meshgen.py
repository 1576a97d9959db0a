"""Synthetic surface meshes for the PMCHWT workloads, written as Gmsh 2.2 ASCII.

The reference reads Gmsh 2.1/2.2 triangle meshes (mesh.f90:610-736) whose
surfaces carry a physical id (first tag) and whose normals point into the
surrounding space (README.md:356).  The generators here build icospheres by
midpoint subdivision of the icosahedron with vertices projected to the
sphere, in a fixed, seed-free order, so that every consumer of the same file
sees the same node and face numbering.

Sizes: an icosphere subdivided s times has F = 20*4**s faces,
E = 1.5*F edges and V = E - F + 2 vertices.
"""
from __future__ import annotations

import numpy as np

__all__ = ["icosphere", "ellipsoid", "perturbed_sphere", "dimer", "octant_sphere", "sphere_cell", "octasphere",
           "write_gmsh", "read_gmsh"]


def _icosahedron():
    t = (1.0 + np.sqrt(5.0)) / 2.0
    v = np.array([
        [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
        [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
        [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1],
    ], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1)[:, None]
    f = np.array([
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
    ], dtype=np.int64)
    return v, f


def icosphere(subdiv: int, radius: float = 1.0):
    """Unit icosahedron subdivided `subdiv` times, projected to `radius`.

    Returns (nodes[V,3] float64, faces[F,3] int64 0-based) with outward
    normals ((p2-p1)x(p3-p1) points away from the centre).
    """
    v, f = _icosahedron()
    verts = [tuple(x) for x in v]
    faces = f
    for _ in range(subdiv):
        cache = {}
        new_faces = np.empty((faces.shape[0] * 4, 3), dtype=np.int64)

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            idx = cache.get(key)
            if idx is None:
                pa, pb = np.array(verts[a]), np.array(verts[b])
                pm = 0.5 * (pa + pb)
                pm /= np.linalg.norm(pm)
                idx = len(verts)
                verts.append(tuple(pm))
                cache[key] = idx
            return idx

        for i, (a, b, c) in enumerate(faces):
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_faces[4 * i + 0] = (a, ab, ca)
            new_faces[4 * i + 1] = (b, bc, ab)
            new_faces[4 * i + 2] = (c, ca, bc)
            new_faces[4 * i + 3] = (ab, bc, ca)
        faces = new_faces
    nodes = np.array(verts, dtype=np.float64) * radius
    # orientation check: outward normals
    p1, p2, p3 = nodes[faces[:, 0]], nodes[faces[:, 1]], nodes[faces[:, 2]]
    nrm = np.cross(p2 - p1, p3 - p1)
    cp = (p1 + p2 + p3) / 3.0
    flip = np.einsum("ij,ij->i", nrm, cp) < 0
    faces[flip] = faces[flip][:, ::-1]
    return nodes, faces


def ellipsoid(subdiv: int, semi_axes=(80.0, 40.0, 40.0)):
    """Icosphere scaled per axis (prolate ellipsoid for semi-axes a > b = c)."""
    nodes, faces = icosphere(subdiv, 1.0)
    return nodes * np.asarray(semi_axes, dtype=np.float64)[None, :], faces


def _real_sph_harm_basis(lmax: int, xyz: np.ndarray) -> np.ndarray:
    from scipy.special import sph_harm_y  # noqa: F401  (scipy >= 1.15)

    r = np.linalg.norm(xyz, axis=1)
    theta = np.arccos(np.clip(xyz[:, 2] / r, -1.0, 1.0))
    phi = np.arctan2(xyz[:, 1], xyz[:, 0])
    cols = []
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), theta, phi)
            if m < 0:
                cols.append(np.sqrt(2.0) * (-1) ** m * y.imag)
            elif m == 0:
                cols.append(y.real)
            else:
                cols.append(np.sqrt(2.0) * (-1) ** m * y.real)
    return np.stack(cols, axis=1)


def perturbed_sphere(subdiv: int, radius: float, amplitude: float, lmax: int = 6, seed: int = 20251203):
    """Icosphere with a smooth radial perturbation sum_{l<=lmax,m} c_lm Y_lm,
    c_lm ~ N(0,1) from a seeded generator, normalised to max |dr| = amplitude."""
    nodes, faces = icosphere(subdiv, 1.0)
    basis = _real_sph_harm_basis(lmax, nodes)
    rng = np.random.default_rng(seed)
    c = rng.standard_normal(basis.shape[1])
    dr = basis @ c
    dr *= amplitude / np.max(np.abs(dr))
    return nodes * (radius + dr)[:, None], faces


def dimer(subdiv: int, radius: float, gap: float):
    """Two icospheres of `radius` centred at -/+(radius + gap/2) along x.
    Returns nodes, faces, ids (15 for the sphere at -x, 16 for +x)."""
    n1, f1 = icosphere(subdiv, radius)
    off = np.array([radius + 0.5 * gap, 0.0, 0.0])
    nodes = np.concatenate([n1 - off, n1 + off])
    faces = np.concatenate([f1, f1 + n1.shape[0]])
    ids = np.concatenate([np.full(f1.shape[0], 15), np.full(f1.shape[0], 16)])
    return nodes, faces, ids


# line-element ids of the mirror-plane boundaries written by octant_sphere
OCTANT_LINE_IDS = {"xplane": 101, "yplane": 102, "zplane": 103}


def _octant_grid(n):
    """Nodes P(i, j) = (i e_x + j e_y + (n-i-j) e_z)/n of the octahedron face,
    i + j <= n, projected to the unit sphere (coordinates on the planes are
    exactly 0), and their index."""
    index, pts = {}, []
    for i in range(n + 1):
        for j in range(n + 1 - i):
            v = np.array([i, j, n - i - j], dtype=np.float64) / n
            index[(i, j)] = len(pts)
            pts.append(v / np.linalg.norm(v))
    return np.array(pts), index


def octant_sphere(n: int, radius: float = 1.0):
    """The octant x, y, z >= 0 of a sphere: the octahedron face (e_x, e_y,
    e_z) split into n^2 triangles, projected to `radius`.  Its edges on the
    planes x = 0, y = 0, z = 0 are open boundaries, returned as line
    elements with the ids of OCTANT_LINE_IDS.  Mirrored in the three planes
    (octasphere) it closes into a sphere of 8 n^2 faces.

    Returns nodes[V,3], faces[F,3] (0-based, outward normals), lines[L,2],
    line_ids[L]."""
    pts, index = _octant_grid(n)
    faces = []
    for i in range(n):
        for j in range(n - i):
            faces.append((index[(i, j)], index[(i + 1, j)], index[(i, j + 1)]))
            if i + j < n - 1:
                faces.append((index[(i + 1, j)], index[(i + 1, j + 1)], index[(i, j + 1)]))
    faces = np.array(faces, dtype=np.int64)
    nodes = pts * radius
    p1, p2, p3 = nodes[faces[:, 0]], nodes[faces[:, 1]], nodes[faces[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(p2 - p1, p3 - p1), (p1 + p2 + p3) / 3.0) < 0
    faces[flip] = faces[flip][:, ::-1]
    lines, lids = [], []
    for j in range(n):  # x = 0: i = 0
        lines.append((index[(0, j)], index[(0, j + 1)]))
        lids.append(OCTANT_LINE_IDS["xplane"])
    for i in range(n):  # y = 0: j = 0
        lines.append((index[(i, 0)], index[(i + 1, 0)]))
        lids.append(OCTANT_LINE_IDS["yplane"])
    for i in range(n):  # z = 0: i + j = n
        lines.append((index[(i, n - i)], index[(i + 1, n - i - 1)]))
        lids.append(OCTANT_LINE_IDS["zplane"])
    return nodes, faces, np.array(lines, dtype=np.int64), np.array(lids, dtype=np.int64)


def sphere_cell(n: int, radius: float = 1.0, reflect=()):
    """The octant of octant_sphere(n) together with its images in the planes
    listed in `reflect` (0 = x, 1 = y, 2 = z): a symmetry cell of the sphere
    whose open boundaries lie on the remaining planes, with their line
    elements (ids of OCTANT_LINE_IDS).  reflect=() is the octant itself,
    (0, 1, 2) the closed sphere (octasphere).

    Returns nodes, faces (outward), lines[L,2], line_ids[L]."""
    on, of, ol, oid = octant_sphere(n, radius)
    key_of, nodes, faces, lines, lids = {}, [], [], [], []
    signs = [[1.0, -1.0] if a in reflect else [1.0] for a in range(3)]
    for sx in signs[0]:
        for sy in signs[1]:
            for sz in signs[2]:
                js = np.array([sx, sy, sz])
                idx = []
                for p in on * js:
                    k = tuple(float(c) + 0.0 for c in p)  # -0.0 and 0.0 are one node
                    if k not in key_of:
                        key_of[k] = len(nodes)
                        nodes.append(k)
                    idx.append(key_of[k])
                for a, b, c in of:
                    f = (idx[a], idx[b], idx[c])
                    faces.append(f if sx * sy * sz > 0 else f[::-1])
                for (a, b), lid in zip(ol, oid):
                    axis = {OCTANT_LINE_IDS["xplane"]: 0, OCTANT_LINE_IDS["yplane"]: 1,
                            OCTANT_LINE_IDS["zplane"]: 2}[int(lid)]
                    if axis not in reflect:
                        lines.append((idx[a], idx[b]))
                        lids.append(int(lid))
    return (np.array(nodes, dtype=np.float64), np.array(faces, dtype=np.int64),
            np.array(lines, dtype=np.int64).reshape(-1, 2), np.array(lids, dtype=np.int64))


def octasphere(n: int, radius: float = 1.0):
    """The closed sphere made of the 8 mirror images of octant_sphere(n)
    (shared nodes on the planes merged), outward normals.  The same
    triangles as the octant problem with the group mxp myp mzp."""
    on, of, _, _ = octant_sphere(n, radius)
    key_of, nodes, faces = {}, [], []
    for sx in (1.0, -1.0):
        for sy in (1.0, -1.0):
            for sz in (1.0, -1.0):
                js = np.array([sx, sy, sz])
                idx = []
                for p in on * js:
                    k = tuple(float(c) + 0.0 for c in p)  # -0.0 and 0.0 are one node
                    if k not in key_of:
                        key_of[k] = len(nodes)
                        nodes.append(k)
                    idx.append(key_of[k])
                for a, b, c in of:
                    f = (idx[a], idx[b], idx[c])
                    faces.append(f if sx * sy * sz > 0 else f[::-1])
    nodes = np.array(nodes, dtype=np.float64)
    faces = np.array(faces, dtype=np.int64)
    return nodes, faces


def write_gmsh(path: str, nodes: np.ndarray, faces: np.ndarray, ids=15, lines=None, line_ids=None) -> None:
    """Write a Gmsh 2.2 ASCII triangle mesh (and optional line elements,
    written after the triangles: the reference's reader fails on a file
    whose line elements come first); node coordinates use %.17g so that
    every reader recovers the exact doubles."""
    ids = np.broadcast_to(np.asarray(ids), (faces.shape[0],))
    nl = 0 if lines is None else len(lines)
    nf = faces.shape[0]
    out = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(nodes.shape[0])]
    out += ["%d %.17g %.17g %.17g" % (i + 1, x, y, z) for i, (x, y, z) in enumerate(nodes)]
    out += ["$EndNodes", "$Elements", str(nf + nl)]
    out += ["%d 2 2 %d %d %d %d %d" % (i + 1, ids[i], ids[i], a + 1, b + 1, c + 1)
            for i, (a, b, c) in enumerate(faces)]
    if nl:
        out += ["%d 1 2 %d %d %d %d" % (nf + i + 1, line_ids[i], line_ids[i], a + 1, b + 1)
                for i, (a, b) in enumerate(lines)]
    lines = out
    lines += ["$EndElements", ""]
    with open(path, "w") as fh:
        fh.write("\n".join(lines))


def read_gmsh(path: str):
    """Read a Gmsh 2.x ASCII mesh the way load_mesh_gmsh does
    (mesh.f90:610-736): node coordinates indexed by node number, triangles
    (element type 2) with their first tag as the surface id.

    Returns nodes[V,3], faces[F,3] (0-based), ids[F].
    """
    with open(path) as fh:
        lines = fh.read().split("\n")
    i = 0
    nodes = faces = ids = None
    while i < len(lines):
        tok = lines[i].strip()
        if tok == "$MeshFormat":
            ver = lines[i + 1].strip()[:3]
            if ver not in ("2.1", "2.2"):
                raise ValueError("Mesh version is unsupported!")
            i += 2
        elif tok == "$Nodes":
            nn = int(lines[i + 1])
            nodes = np.zeros((nn, 3))
            for k in range(nn):
                parts = lines[i + 2 + k].split()
                nodes[int(parts[0]) - 1] = [float(parts[1]), float(parts[2]), float(parts[3])]
            i += 2 + nn
        elif tok == "$Elements":
            ne = int(lines[i + 1])
            fl, il = [], []
            for k in range(ne):
                v = [int(x) for x in lines[i + 2 + k].split()]
                if v[1] == 2:
                    nt = v[2]
                    il.append(v[3])
                    fl.append([v[3 + nt] - 1, v[4 + nt] - 1, v[5 + nt] - 1])
            faces = np.array(fl, dtype=np.int64).reshape(-1, 3)
            ids = np.array(il, dtype=np.int64)
            i += 2 + ne
        else:
            i += 1
    if nodes is None or faces is None:
        raise ValueError("incomplete mesh file")
    return nodes, faces, ids
