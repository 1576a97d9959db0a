"""Host side of the mirror-symmetry decomposition, without a GPU: the group
table of mom_symm_group against the characters of Z2^k, the edge classes of
the set-up and the removed unknowns of mom_symm_edge_bc against the
reference's own symmetric run (tests/golden/octant_n4.npz, quarter_n4.npz:
the reference's expanded solutions are exactly 0 at the unknowns that
edge_bc removes, bc.f90:199-232)."""
import itertools

import numpy as np
import pytest

from cases import sphere_cell
from conftest import golden
from mom2amd import symm_edge_bc, symm_group


@pytest.mark.parametrize("gens", [["mxp"], ["myp", "mzp"], ["mxp", "myp", "mzp"], ["mzp", "mxp"]])
def test_group_is_the_product_of_mirrors(gens):
    """Element g and representation r in product_group's order (the first
    generator in the highest bit): J_g reflects the planes of g's bits,
    det J_g = (-1)^|g|, ef[g][r] = prod_i (-1)^(g_i r_i), and the
    characters are orthogonal (ef^T ef = n I)."""
    G = symm_group(gens)
    k = len(gens)
    n = 1 << k
    assert G["n"] == n
    axis = {"mxp": 0, "myp": 1, "mzp": 2}
    for g in range(n):
        bits = [(g >> (k - 1 - i)) & 1 for i in range(k)]
        js = np.ones(3)
        for i, b in enumerate(bits):
            if b:
                js[axis[gens[i]]] = -1
        assert np.array_equal(G["jdiag"][g], js)
        assert G["detj"][g] == (-1) ** sum(bits)
        # identifiers: identity, a single mirror (its symmetry.f90 id), else 0
        want = 1 if sum(bits) == 0 else (2 + axis[gens[bits.index(1)]] if sum(bits) == 1 else 0)
        assert G["id"][g] == want
        for r in range(n):
            rb = [(r >> (k - 1 - i)) & 1 for i in range(k)]
            assert G["ef"][g, r] == (-1) ** sum(a * b for a, b in zip(bits, rb))
    assert np.array_equal(G["ef"].T @ G["ef"], n * np.eye(n))


def test_group_rejects_bad_generators():
    from mom2amd.capi import MomError
    with pytest.raises(MomError):
        symm_group(["mxp", "mxp"])
    with pytest.raises(MomError):
        symm_group([4])


@pytest.mark.parametrize("name,reflect", [("octant_n4", ()), ("quarter_n4", (2,))])
def test_setup_edge_classes_match_reference(msh_dir, name, reflect):
    g = golden(name)
    case = sphere_cell(msh_dir, int(g["n"]), reflect)
    P = case["problem"]
    for d, m in enumerate([P.master] + P.domains):
        assert np.array_equal(m.bnd, g["mesh%d_bnd" % d]), d
        for k in ("efaces", "ebn"):
            ref = g["mesh%d_%s" % (d, k)].copy()
            ref[ref == -2] = -1  # the reference's -1 ("none") read as 1-based
            assert np.array_equal(getattr(m, k), ref), (d, k)
        for k in ("enodes", "eparent", "fnodes", "fedges"):
            assert np.array_equal(getattr(m, k), g["mesh%d_%s" % (d, k)]), (d, k)


@pytest.mark.parametrize("name,reflect", [("octant_n4", ()), ("quarter_n4", (2,))])
def test_removed_unknowns_match_reference(msh_dir, name, reflect):
    g = golden(name)
    case = sphere_cell(msh_dir, int(g["n"]), reflect)
    bnd = case["problem"].master.bnd
    gens = [str(s) for s in g["gens"]]
    x = g["x"]
    total = 0
    for r in range(x.shape[1]):
        ids, dim = symm_edge_bc(bnd, gens, r)
        removed = ids == -1
        assert dim == (~removed).sum()
        assert np.all(x[removed, r] == 0)
        assert np.all(x[~removed, r] != 0)
        # each boundary edge loses exactly one of its J and M coefficients
        N = bnd.size
        onplane = bnd > 1
        assert np.array_equal(removed[:N] ^ removed[N:], onplane)
        total += dim
    assert total == x.shape[1] * 2 * bnd.size - x.shape[1] * int((bnd > 1).sum())


def test_sphere_cells_close_into_the_octasphere():
    """The images of the octant in all three planes give the octasphere;
    faces keep outward normals; every mirror-plane node lies exactly on its
    plane."""
    from mom2amd import meshgen
    n = 3
    on, of, ol, oid = meshgen.octant_sphere(n, 1.0)
    assert of.shape[0] == n * n and ol.shape[0] == 3 * n
    for (a, b), lid in zip(ol, oid):
        axis = {101: 0, 102: 1, 103: 2}[int(lid)]
        assert on[a, axis] == 0.0 and on[b, axis] == 0.0
    fn, ff = meshgen.octasphere(n, 1.0)
    V, F = fn.shape[0], ff.shape[0]
    E = 3 * F // 2
    assert F == 8 * n * n and V - E + F == 2  # closed genus-0 surface
    p1, p2, p3 = fn[ff[:, 0]], fn[ff[:, 1]], fn[ff[:, 2]]
    assert np.all(np.einsum("ij,ij->i", np.cross(p2 - p1, p3 - p1), p1 + p2 + p3) > 0)
    cn, cf, cl, _ = meshgen.sphere_cell(n, 1.0, (0, 1, 2))
    assert cf.shape == ff.shape and cl.shape[0] == 0
    for sub in itertools.chain.from_iterable(itertools.combinations(range(3), k) for k in range(3)):
        _, sf, sl, _ = meshgen.sphere_cell(n, 1.0, sub)
        assert sf.shape[0] == n * n * 2 ** len(sub) and sl.shape[0] == n * (3 - len(sub)) * 2 ** len(sub)


def test_edge_bc_rejects_periodic_and_rotation_classes():
    """Only the mirror classes are built: a periodic (5..8) or rotation (9,
    10) boundary class is an error, not a silently unbounded edge."""
    from mom2amd.capi import MomError
    bnd = np.ones(10, np.int32)
    bnd[3] = 5
    with pytest.raises(MomError):
        symm_edge_bc(bnd, ["mxp"], 0)
    bnd[3] = 2
    ids, dim = symm_edge_bc(bnd, ["mxp"], 0)
    assert dim == 19 and ids[3] == -1
