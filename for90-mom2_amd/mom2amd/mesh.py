"""Host-side problem set-up: master mesh, domain meshes, RWG edge numbering.

This mirrors what the reference's `mesh`, `sdom` and `fdom` commands build
(interface.f90:357-366, 44-105, 107-123) and produces the flattened arrays
that the C-ABI (`mom_mesh_create`) takes.  The numbering rules are the
reference's, because every matrix entry's index and sign depends on them:

- edges are numbered by first appearance in face order, the pair being
  (node m, node m+1) of the face (mesh.f90:1066-1108); the first face that
  holds an edge is its "+" face and the face's third vertex its "+" free
  vertex; a later face is the "-" face;
- geometry (normals, tangents s, in-plane normals m, area, centroid, edge
  length, average edge length) follows mesh.f90:1265-1322 and 138-153 with
  the same IEEE operation order, so the arrays equal the reference's bits;
- a domain mesh holds the master faces whose id is listed, in master order,
  with the nodes of negative ids reversed (mesh.f90:278-376);
- every domain edge gets the master edge with the same node pair as parent
  (mesh.f90:378-452), and the children of each master edge are oriented
  against the first child (mesh.f90:454-558).

Indices are 0-based; -1 marks "none".  The reference uses 1-based indices.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .meshgen import read_gmsh

__all__ = ["DomainMesh", "Problem", "build_problem", "load_problem"]


@dataclass
class DomainMesh:
    """Flattened mesh_container (mesh.f90:75-83) of one domain."""

    p: np.ndarray            # (V,3) float64 node coordinates [m]
    fnodes: np.ndarray       # (F,3) int32
    fedges: np.ndarray       # (F,3) int32 local edge of each face side
    fid: np.ndarray          # (F,) int32 physical id
    fparent: np.ndarray      # (F,) int32 master face index
    node_parent: np.ndarray  # (V,) int32 master node index
    fn: np.ndarray = None    # (F,3) unit normal
    fcp: np.ndarray = None   # (F,3) centroid
    fs: np.ndarray = None    # (F,3,3) fs[f,e,:] = unit tangent of side e
    fm: np.ndarray = None    # (F,3,3) fm[f,e,:] = s x n (unit)
    farea: np.ndarray = None  # (F,)
    enodes: np.ndarray = None  # (E,2)
    ebn: np.ndarray = None   # (E,2) free vertex on the + / - face
    efaces: np.ndarray = None  # (E,2) + / - face
    elen: np.ndarray = None  # (E,)
    eparent: np.ndarray = None  # (E,) master edge index
    avelen: float = 0.0

    @property
    def nnodes(self):
        return self.p.shape[0]

    @property
    def nfaces(self):
        return self.fnodes.shape[0]

    @property
    def nedges(self):
        return self.enodes.shape[0]


@dataclass
class Problem:
    master: DomainMesh
    domains: list = field(default_factory=list)

    @property
    def nedges(self):
        """Number of RWG functions per current; the system has 2*nedges unknowns."""
        return self.master.nedges


def _norm3(v):
    # SQRT(SUM(v*v)) in the order v1*v1 + v2*v2 + v3*v3 (linalg.f90:54-59)
    return np.sqrt(v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1] + v[..., 2] * v[..., 2])


def _cross(a, b):
    r = np.empty(np.broadcast(a, b).shape)
    r[..., 0] = a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1]
    r[..., 1] = a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2]
    r[..., 2] = a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]
    return r


def _build_edges(m: DomainMesh) -> None:
    """First-appearance edge numbering (mesh.f90:1044-1137)."""
    F = m.nfaces
    fn = m.fnodes
    lookup = {}
    enodes, ebn, efaces = [], [], []
    fedges = np.empty((F, 3), dtype=np.int32)
    for f in range(F):
        a0, a1, a2 = int(fn[f, 0]), int(fn[f, 1]), int(fn[f, 2])
        tri = (a0, a1, a2)
        for s in range(3):
            a, b, c = tri[s], tri[(s + 1) % 3], tri[(s + 2) % 3]
            key = (a, b) if a < b else (b, a)
            l = lookup.get(key)
            if l is None:
                l = len(enodes)
                lookup[key] = l
                enodes.append((a, b))
                ebn.append([c, -1])
                efaces.append([f, -1])
            else:
                # more than two faces on one edge: the last one is kept
                # as the "-" face (mesh.f90:1082-1086)
                efaces[l][1] = f
                ebn[l][1] = c
            fedges[f, s] = l
    m.fedges = fedges
    m.enodes = np.array(enodes, dtype=np.int32).reshape(-1, 2)
    m.ebn = np.array(ebn, dtype=np.int32).reshape(-1, 2)
    m.efaces = np.array(efaces, dtype=np.int32).reshape(-1, 2)
    m.eparent = np.full(m.enodes.shape[0], -1, dtype=np.int32)


def _basis_data(m: DomainMesh) -> None:
    """Face and edge geometry (mesh.f90:1265-1322) and avelen (138-153)."""
    p = m.p
    p1, p2, p3 = p[m.fnodes[:, 0]], p[m.fnodes[:, 1]], p[m.fnodes[:, 2]]
    b = _cross(p2 - p1, p3 - p1)
    m.fn = b / _norm3(b)[:, None]
    p21, p32, p13 = p2 - p1, p3 - p2, p1 - p3
    s = np.stack([p21 / _norm3(p21)[:, None], p32 / _norm3(p32)[:, None],
                  p13 / _norm3(p13)[:, None]], axis=1)
    mm = _cross(s, m.fn[:, None, :])
    mm = mm / _norm3(mm)[..., None]
    m.fs, m.fm = s, mm
    m.farea = 0.5 * _norm3(_cross(-p13, p32))
    m.fcp = (p1 + p2 + p3) / 3.0
    d = p[m.enodes[:, 0]] - p[m.enodes[:, 1]]
    m.elen = _norm3(d)
    # sequential sum, as the reference loop does
    acc = 0.0
    for v in m.elen.tolist():
        acc = acc + v
    m.avelen = acc / m.nedges


def _build_mesh(m: DomainMesh, scale: float) -> None:
    m.p = m.p * scale
    _build_edges(m)
    _basis_data(m)


def _extract(master: DomainMesh, ids) -> DomainMesh:
    """extract_submesh (mesh.f90:290-376), surface faces only."""
    sel = np.isin(master.fid, np.asarray(ids))
    faces = master.fnodes[sel]
    mask = np.zeros(master.nnodes, dtype=bool)
    mask[faces.ravel()] = True
    node_parent = np.nonzero(mask)[0].astype(np.int32)
    loc = np.full(master.nnodes, -1, dtype=np.int64)
    loc[node_parent] = np.arange(node_parent.size)
    return DomainMesh(
        p=master.p[node_parent].copy(),
        fnodes=loc[faces].astype(np.int32),
        fedges=None,
        fid=master.fid[sel].astype(np.int32),
        fparent=np.nonzero(sel)[0].astype(np.int32),
        node_parent=node_parent,
    )


def _connect_and_orient(master: DomainMesh, doms) -> None:
    """submesh_edge_connectivity (mesh.f90:378-452) then orient_basis
    (mesh.f90:454-558), identity group and no periodic boundaries."""
    lookup = {}
    for n, (a, b) in enumerate(master.enodes.tolist()):
        lookup[(a, b) if a < b else (b, a)] = n
    children = [[] for _ in range(master.nedges)]
    for d, dm in enumerate(doms):
        par = dm.node_parent[dm.enodes]
        for l, (a, b) in enumerate(par.tolist()):
            n = lookup.get((a, b) if a < b else (b, a))
            if n is not None:
                dm.eparent[l] = n
                children[n].append((d, l))
    # children are visited by (domain, local edge) order, which is the order
    # the reference's n-m-l loop records them in.
    for n, ch in enumerate(children):
        if len(ch) < 2:
            raise RuntimeError("Submesh decomposition is inconsistent (invalid child edges)!")
    for n, ch in enumerate(children):
        nchild = len(ch)
        oriented = [False] * nchild
        oriented[0] = True
        exited_at = None
        for it in range(1, nchild + 1):
            for l in range(1, nchild):
                if oriented[l]:
                    continue
                dl, el = ch[l]
                for k in range(nchild):
                    if not oriented[k]:
                        continue
                    dk, ek = ch[k]
                    Ml, Mk = doms[dl], doms[dk]
                    pf1, nf1 = (int(x) for x in Ml.efaces[el])
                    pf2, nf2 = (int(x) for x in Mk.efaces[ek])
                    pf1 = int(Ml.fparent[pf1]) if pf1 != -1 else -1
                    nf1 = int(Ml.fparent[nf1]) if nf1 != -1 else -1
                    pf2 = int(Mk.fparent[pf2]) if pf2 != -1 else -1
                    nf2 = int(Mk.fparent[nf2]) if nf2 != -1 else -1
                    if pf1 == pf2 or nf1 == nf2:
                        Ml.efaces[el] = Ml.efaces[el][::-1].copy()
                        Ml.ebn[el] = Ml.ebn[el][::-1].copy()
                    if pf1 == pf2 or nf1 == nf2 or pf1 == nf2 or nf1 == pf2:
                        oriented[l] = True
                        break
            if all(oriented):
                exited_at = it
                break
        # The reference checks its loop counter after the loop; success in the
        # last pass is reported as failure and ends the orientation of the
        # remaining edges (mesh.f90:552-555).
        if exited_at is not None and exited_at == nchild:
            import sys
            print("Basis orientation was unsuccessful!", file=sys.stderr)
            break


def build_problem(nodes, faces, ids, scale: float, domains) -> Problem:
    """Build the master mesh and domain meshes.

    nodes: (V,3) coordinates in file units; faces: (F,3) 0-based; ids: (F,)
    physical ids; scale: the `mesh` command's scale factor; domains: list of
    lists of signed surface ids, one list per `sdom` (negative = reversed).
    """
    master = DomainMesh(p=np.asarray(nodes, dtype=np.float64).copy(),
                        fnodes=np.asarray(faces, dtype=np.int32).copy(), fedges=None,
                        fid=np.asarray(ids, dtype=np.int32).copy(),
                        fparent=np.arange(len(faces), dtype=np.int32),
                        node_parent=np.arange(len(nodes), dtype=np.int32))
    _build_mesh(master, scale)
    doms = []
    for sids in domains:
        dm = _extract(master, [abs(s) for s in sids])
        for s in sids:
            if s == 0:
                raise ValueError("Zero is not a valid surface id!")
            if s < 0:
                inv = dm.fid == -s
                dm.fnodes[inv] = dm.fnodes[inv][:, ::-1]
        _build_mesh(dm, 1.0)
        doms.append(dm)
    _connect_and_orient(master, doms)
    return Problem(master=master, domains=doms)


def load_problem(path: str, scale: float, domains) -> Problem:
    nodes, faces, ids = read_gmsh(path)
    return build_problem(nodes, faces, ids, scale, domains)
