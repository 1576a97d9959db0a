"""Host-side problem set-up: master mesh, domain meshes, RWG edge numbering.

This mirrors what the reference's `mesh` and `sdom` commands build
(interface.f90:357-366, 44-105) and produces the flattened arrays that the
C-ABI (`mom_mesh_create`) takes.  The work is done by the library's native
set-up (mom_problem_build, csrc/setup.hip); this module holds the arrays.
The numbering rules are the reference's, because every matrix entry's index
and sign depends on them:

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

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import capi

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
    bnd: np.ndarray = None   # (E,) boundary class (mesh_bnd_*: 1 none, 2/3/4 the x/y/z mirror planes)

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


def _fetch(L, h, d) -> DomainMesh:
    nt, V, F, E = (C.c_int64() for _ in range(4))
    capi.check(L.mom_problem_sizes(h, d, C.byref(nt), C.byref(V), C.byref(F), C.byref(E)))
    V, F, E = V.value, F.value, E.value
    m = DomainMesh(p=np.zeros((V, 3)), fnodes=np.zeros((F, 3), np.int32), fedges=np.zeros((F, 3), np.int32),
                   fid=np.zeros(F, np.int32), fparent=np.zeros(F, np.int32), node_parent=np.zeros(V, np.int32),
                   fn=np.zeros((F, 3)), fcp=np.zeros((F, 3)), fs=np.zeros((F, 3, 3)), fm=np.zeros((F, 3, 3)),
                   farea=np.zeros(F), enodes=np.zeros((E, 2), np.int32), ebn=np.zeros((E, 2), np.int32),
                   efaces=np.zeros((E, 2), np.int32), elen=np.zeros(E), eparent=np.zeros(E, np.int32))
    flat = lambda a: a.reshape(-1)
    av = C.c_double()
    capi.check(L.mom_problem_domain(h, d, flat(m.p), flat(m.fnodes), flat(m.fedges), flat(m.fn), flat(m.fcp),
                                    flat(m.fs), flat(m.fm), m.farea, flat(m.efaces), flat(m.ebn), m.elen,
                                    m.eparent, C.byref(av)))
    capi.check(L.mom_problem_topology(h, d, flat(m.enodes), m.fid, m.fparent, m.node_parent))
    m.bnd = np.zeros(E, np.int32)
    capi.check(L.mom_problem_edge_bnd(h, d, m.bnd))
    m.avelen = av.value
    return m


def _domain_lists(domains):
    dom_ptr = np.cumsum([0] + [len(s) for s in domains]).astype(np.int32)
    dom_ids = np.array([int(s) for sids in domains for s in sids], dtype=np.int32)
    return dom_ptr, dom_ids


# boundary classes of the `sbnd` command (interface.f90:120-157, mesh.f90:16-25)
BND_CLASSES = {"none": 1, "xplane": 2, "yplane": 3, "zplane": 4}


def _classify(L, h, lines, sbnd):
    if lines is not None:
        ln, lid = lines
        ln = np.ascontiguousarray(ln, dtype=np.int32).reshape(-1)
        lid = np.ascontiguousarray(lid, dtype=np.int32).reshape(-1)
        capi.check(L.mom_problem_set_lines(h, lid.size, ln, lid))
    for lid, cls in sbnd:
        code = BND_CLASSES[cls] if isinstance(cls, str) else int(cls)
        capi.check(L.mom_problem_classify_edges(h, int(lid), code, None))


def _collect(L, h, ndom, lines=None, sbnd=()) -> Problem:
    try:
        _classify(L, h, lines, sbnd)
        master = _fetch(L, h, -1)
        doms = [_fetch(L, h, d) for d in range(ndom)]
    finally:
        L.mom_problem_destroy(h)
    return Problem(master=master, domains=doms)


def build_problem(nodes, faces, ids, scale: float, domains, lines=None, sbnd=()) -> Problem:
    """Build the master mesh and domain meshes with the library's set-up
    (mom_problem_build: build_mesh, extract_submesh,
    submesh_edge_connectivity and orient_basis of mesh.f90).

    nodes: (V,3) coordinates in file units; faces: (F,3) 0-based; ids: (F,)
    physical ids; scale: the `mesh` command's scale factor; domains: list of
    lists of signed surface ids, one list per `sdom` (negative = reversed);
    lines: (node pairs (L,2) 0-based, ids (L,)) of the mesh's line elements;
    sbnd: (line id, class) pairs of the `sbnd` command, class one of
    BND_CLASSES (mom_problem_classify_edges).
    """
    L = capi.lib()
    nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 3)
    faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
    ids = np.ascontiguousarray(ids, dtype=np.int32).reshape(-1)
    if ids.shape[0] != faces.shape[0]:
        raise ValueError("one physical id per face is needed")
    dom_ptr, dom_ids = _domain_lists(domains)
    h = C.c_void_p()
    capi.check(L.mom_problem_build(nodes.shape[0], nodes.reshape(-1), faces.shape[0], faces.reshape(-1), ids,
                                   float(scale), len(domains), dom_ptr, dom_ids, C.byref(h)))
    return _collect(L, h, len(domains), lines, sbnd)


def load_problem(path: str, scale: float, domains, sbnd=()) -> Problem:
    """`mesh` + `sbnd` + `sdom` + `fdom` from a Gmsh 2.x file, read and built
    by the library (mom_problem_load_gmsh; load_mesh_gmsh, mesh.f90:610-736);
    sbnd classifies edges by the file's line elements as build_problem's."""
    L = capi.lib()
    dom_ptr, dom_ids = _domain_lists(domains)
    h = C.c_void_p()
    capi.check(L.mom_problem_load_gmsh(str(path).encode(), float(scale), len(domains), dom_ptr, dom_ids,
                                       C.byref(h)))
    return _collect(L, h, len(domains), None, sbnd)
