"""CPU checks of the library's host set-up (mom_problem_build and
mom_problem_load_gmsh: the `mesh`, `sdom` and `fdom` commands, mesh.f90:610-736,
1044-1137, 1265-1322, 290-558).  No GPU work: the set-up runs on the host
inside libmomgpu.so.

Pins: the reference's own mesh dumps (tests/golden, bit for bit, in
test_oracle.py::test_host_mesh_builder_matches_reference), the oracle's
independent builder on larger and non-manifold meshes, and the orientation
invariant the reference's orient_basis establishes.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from cases import dimer, sphere
from mom2amd import build_problem, capi, meshgen

KEYS = ("p", "fnodes", "fedges", "fn", "fcp", "fs", "fm", "farea", "enodes", "ebn", "efaces", "elen")


def _same_as_oracle(path, doms, P):
    op = oracle.OracleProblem(path, 1e-9, doms)
    for d, m in enumerate([P.master] + P.domains):
        ref = op.mesh(d - 1)
        for k in KEYS:
            assert np.array_equal(getattr(m, k), ref[k]), (path, d, k)
        if d > 0:
            assert np.array_equal(m.eparent, ref["eparent"]), (path, d)
        assert m.avelen == ref["avelen"], (path, d)


@pytest.mark.parametrize("kind,s", [("sphere", 3), ("sphere", 4), ("dimer", 2)])
def test_setup_matches_oracle(msh_dir, kind, s):
    case = sphere(msh_dir, s) if kind == "sphere" else dimer(msh_dir, s)
    _same_as_oracle(case["path"], case["domains"], case["problem"])


def _fan(msh_dir):
    """Three triangles on the edge (0,1) (a non-manifold fin) closed into a
    double cover by the second domain."""
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0.5, 1, 0], [0.5, -1, 0], [0.5, 0, 1]], dtype=np.float64)
    faces = np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], dtype=np.int32)
    path = str(msh_dir / "fan.msh")
    meshgen.write_gmsh(path, nodes, faces, 7)
    return path, nodes, faces


def test_edge_on_three_faces_keeps_last_face(msh_dir):
    """An edge on more than two faces takes the first as "+" and the last as
    "-" face (mesh.f90:1082-1086)."""
    path, nodes, faces = _fan(msh_dir)
    doms = [[7], [-7]]
    P = build_problem(*meshgen.read_gmsh(path), 1e-9, doms)
    m = P.master
    assert m.enodes[0].tolist() == [0, 1]
    assert m.efaces[0].tolist() == [0, 2] and m.ebn[0].tolist() == [2, 4]
    assert m.nedges == 7 and (m.fedges[:, 0] == 0).all()
    _same_as_oracle(path, doms, P)


def test_orientation_invariant(msh_dir):
    """After orient_basis, two children of one master edge that share a
    master face hold it in opposite roles (one's "+" face is the other's "-"
    face), so the global RWG function has one orientation."""
    P = dimer(msh_dir, 2)["problem"]
    roles = {}
    for dm in P.domains:
        for l in range(dm.nedges):
            pf, nf = (int(dm.fparent[f]) if f >= 0 else -1 for f in dm.efaces[l])
            roles.setdefault(int(dm.eparent[l]), []).append((pf, nf))
    assert len(roles) == P.master.nedges
    for n, ch in roles.items():
        assert len(ch) >= 2
        for i in range(len(ch)):
            for j in range(i + 1, len(ch)):
                (p1, n1), (p2, n2) = ch[i], ch[j]
                assert not (p1 == p2 and p1 >= 0) and not (n1 == n2 and n1 >= 0), n
    # every child's nodes map onto its parent's node pair
    for dm in P.domains:
        a = np.sort(dm.node_parent[dm.enodes], axis=1)
        b = np.sort(P.master.enodes[dm.eparent], axis=1)
        assert np.array_equal(a, b)


def test_domain_faces_and_nodes(msh_dir):
    """A domain takes the master faces of its ids in master order and the
    used master nodes in ascending order; negative ids reverse the nodes."""
    case = dimer(msh_dir, 1)
    P = case["problem"]
    master = P.master
    for dm, sids in zip(P.domains, case["domains"]):
        sel = np.isin(master.fid, np.abs(sids))
        assert np.array_equal(dm.fparent, np.nonzero(sel)[0])
        assert np.array_equal(dm.node_parent, np.unique(master.fnodes[sel]))
        glob = dm.node_parent[dm.fnodes]
        for s in sids:
            rows = master.fid[dm.fparent] == abs(s)
            want = master.fnodes[dm.fparent[rows]]
            assert np.array_equal(glob[rows], want[:, ::-1] if s < 0 else want)


def _raw_build(nodes, faces, ids, scale, doms):
    L = capi.lib()
    nodes = np.ascontiguousarray(nodes, np.float64)
    faces = np.ascontiguousarray(faces, np.int32)
    ptr = np.cumsum([0] + [len(s) for s in doms]).astype(np.int32)
    di = np.array([s for d in doms for s in d], np.int32)
    h = C.c_void_p()
    st = L.mom_problem_build(len(nodes), nodes.ravel(), len(faces), faces.ravel(),
                             np.ascontiguousarray(ids, np.int32), scale, len(doms), ptr, di, C.byref(h))
    if st == 0:
        L.mom_problem_destroy(h)
    return st, L.mom_last_error().decode()


def test_setup_errors():
    nodes, faces = meshgen.icosphere(1, 50.0)
    ids = np.full(len(faces), 15)
    assert _raw_build(nodes, faces, ids, 1e-9, [[15], [-15]])[0] == 0
    st, msg = _raw_build(nodes, faces, ids, 1e-9, [[15], [0]])
    assert st == -1 and "Zero is not a valid surface id" in msg
    st, msg = _raw_build(nodes, faces, ids, 1e-9, [[15]])
    assert st == -1 and "inconsistent" in msg
    bad = faces.copy()
    bad[3, 1] = len(nodes)
    st, msg = _raw_build(nodes, bad, ids, 1e-9, [[15], [-15]])
    assert st == -1 and "out of range" in msg
    st, msg = _raw_build(nodes, faces, ids, 1e-9, [[15], []])
    assert st == -1 and "at least one surface id" in msg
    st, msg = _raw_build(nodes, faces, ids, 0.0, [[15], [-15]])
    assert st == -1 and "scale" in msg
    with pytest.raises(capi.MomError):
        build_problem(nodes, faces, ids, 1e-9, [[15]])


@pytest.mark.parametrize("kind,s", [("sphere", 2), ("dimer", 1)])
def test_gmsh_loader_matches_file_arrays(msh_dir, kind, s):
    """mom_problem_load_gmsh reads the file as load_mesh_gmsh does and gives
    the same problem as mom_problem_build on the file's arrays."""
    from mom2amd import load_problem

    case = sphere(msh_dir, s) if kind == "sphere" else dimer(msh_dir, s)
    P = load_problem(case["path"], 1e-9, case["domains"])
    Q = case["problem"]
    for a, b in zip([P.master] + P.domains, [Q.master] + Q.domains):
        for k in KEYS + ("eparent", "fid", "fparent", "node_parent"):
            assert np.array_equal(getattr(a, k), getattr(b, k)), k
        assert a.avelen == b.avelen


def test_gmsh_loader_node_order_and_other_elements(msh_dir):
    """Nodes are stored at their node number, whatever the order of the
    lines; points, lines and tetrahedra are skipped; the first tag is the
    surface id."""
    from mom2amd import load_problem

    nodes, faces = meshgen.icosphere(1, 50.0)
    ids = np.full(len(faces), 15)
    perm = np.random.default_rng(3).permutation(len(nodes))
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(len(nodes))]
    lines += ["%d %.17g %.17g %.17g" % (k + 1, *nodes[k]) for k in perm]
    lines += ["$EndNodes", "$Elements", str(len(faces) + 3)]
    lines += ["1 15 2 99 1 1", "2 1 2 98 3 1 2"]
    lines += ["%d 2 2 15 7 %d %d %d" % (k + 3, *(faces[k] + 1)) for k in range(len(faces))]
    lines += ["%d 4 2 97 1 1 2 3 4" % (len(faces) + 3), "$EndElements", ""]
    path = msh_dir / "perm.msh"
    path.write_text("\n".join(lines))
    P = load_problem(str(path), 1e-9, [[15], [-15]])
    Q = build_problem(nodes, faces, ids, 1e-9, [[15], [-15]])
    for a, b in zip([P.master] + P.domains, [Q.master] + Q.domains):
        for k in KEYS:
            assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_gmsh_loader_errors(msh_dir):
    from mom2amd import load_problem

    with pytest.raises(capi.MomError, match="Could not open mesh file"):
        load_problem(str(msh_dir / "missing.msh"), 1e-9, [[1], [-1]])
    bad = msh_dir / "v4.msh"
    bad.write_text("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")
    with pytest.raises(capi.MomError, match="Mesh version is unsupported"):
        load_problem(str(bad), 1e-9, [[1], [-1]])
    bad = msh_dir / "noend.msh"
    bad.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n1\n1 0 0 0\n$Elements\n")
    with pytest.raises(capi.MomError, match="EndNodes"):
        load_problem(str(bad), 1e-9, [[1], [-1]])
