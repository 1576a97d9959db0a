"""Host side of the symmetry decomposition, without a GPU: the group table of
mom_symm_group against the characters of Z2^k, the rotation groups of
mom_symm_group_c against group_rz's definition (rotation matrices, the
characters of Z_n, product_group's order), the edge classes of the set-up
and the removed unknowns of mom_symm_edge_bc against the reference's own
symmetric runs (tests/golden/octant_n4.npz, quarter_n4.npz, trimer_r3.npz,
ring_r4_mzp.npz: the reference's expanded solutions are exactly 0 at the
unknowns that edge_bc removes, bc.f90:199-232)."""
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
    for gens in (["r1"], ["r9"], ["r2", "r3"], ["r8", "mzp"], ["r3", "mxp", "myp"], ["r4", "mxp"], ["myp", "r3"]):
        with pytest.raises(MomError):
            symm_group(gens)
    # r2 (rotation by pi, diagonal) commutes with every mirror
    assert symm_group(["r2", "mxp"])["n"] == 4


def _rz(a):
    return np.array([[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]])


@pytest.mark.parametrize("n", [2, 3, 4, 6, 8])
def test_rotation_group(n):
    """group_rz(n) (symmetry.f90:105-156): element k rotates by 2 pi k / n
    counter-clockwise about z, det 1, the identity first and the rz
    identifier (5) on the last element; ef[k][l] = conj(exp(i 2 pi k l / n)),
    the characters of Z_n, orthogonal: ef^H ef = n I."""
    # the reference's pi is the single-precision literal 3.141592653589
    # (constants.f90:10), so the rotations and characters are exact only to
    # about 1e-7 and orthogonal to about 1e-6
    pi = float(np.float32(3.141592653589))
    G = symm_group(["r%d" % n])
    assert G["n"] == n
    for k in range(n):
        assert np.abs(G["J"][k] - _rz(2 * pi * k / n)).max() < 1e-15
        assert abs(np.linalg.det(G["J"][k]) - 1) < 1e-14 and G["detj"][k] == 1.0
        for l in range(n):
            assert abs(G["ef"][k, l] - np.exp(-2j * pi * k * l / n)) < 1e-14
    assert np.abs(G["J"][1] - _rz(2 * np.pi / n)).max() < 1e-6
    assert list(G["id"]) == [1] + [0] * (n - 2) + [5]
    assert np.abs(G["ef"].conj().T @ G["ef"] - n * np.eye(n)).max() < 1e-5
    # the real table of mom_symm_group refuses a group with full Jacobians
    # (with the reference's pi even r2 is not diagonal)
    import ctypes as C
    from mom2amd import capi
    from mom2amd.capi import MomError
    p = np.array([100 + n], np.int32)
    with pytest.raises(MomError):
        capi.check(capi.lib().mom_symm_group(1, p, C.byref(C.c_int32()), np.zeros(64), np.zeros(8), np.zeros(24),
                                             np.zeros(8, np.int32)))


def test_rotation_times_mirror():
    """product_group(r4, mzp): element k = 2 n + m has J = J_n J_m, det the
    product, ef[k][2 n2 + m2] = ef_r4[n][n2] ef_mzp[m][m2], and the pure
    mirror (n = 0, m = 1) keeps the mzp identifier (4) that edge_bc uses."""
    R, M, G = symm_group(["r4"]), symm_group(["mzp"]), symm_group(["r4", "mzp"])
    assert G["n"] == 8
    for n in range(4):
        for m in range(2):
            k = 2 * n + m
            assert np.array_equal(G["J"][k], R["J"][n] @ M["J"][m])
            assert G["detj"][k] == R["detj"][n] * M["detj"][m]
            for n2 in range(4):
                for m2 in range(2):
                    assert G["ef"][k, 2 * n2 + m2] == R["ef"][n, n2] * M["ef"][m, m2]
    assert G["id"][0] == 1 and G["id"][1] == 4 and G["id"][6] == 5
    assert np.abs(G["ef"].conj().T @ G["ef"] - 8 * np.eye(8)).max() < 1e-5


CELLS = [("octant_n4", ()), ("quarter_n4", (2,)), ("trimer_r3", (0, 1, 2)), ("ring_r4_mzp", (0, 1))]


def _cell(msh_dir, g, reflect):
    radius = float(g["radius"]) if "radius" in g else 50.0
    shift = float(g["shift"]) if "shift" in g else 0.0
    return sphere_cell(msh_dir, int(g["n"]), reflect, radius=radius, shift=shift)


@pytest.mark.parametrize("name,reflect", CELLS)
def test_setup_edge_classes_match_reference(msh_dir, name, reflect):
    g = golden(name)
    case = _cell(msh_dir, g, reflect)
    P = case["problem"]
    for d, m in enumerate([P.master] + P.domains):
        assert np.array_equal(m.bnd, g["mesh%d_bnd" % d]), d
        for k in ("efaces", "ebn"):
            ref = g["mesh%d_%s" % (d, k)].copy()
            ref[ref == -2] = -1  # the reference's -1 ("none") read as 1-based
            assert np.array_equal(getattr(m, k), ref), (d, k)
        for k in ("enodes", "eparent", "fnodes", "fedges"):
            assert np.array_equal(getattr(m, k), g["mesh%d_%s" % (d, k)]), (d, k)


@pytest.mark.parametrize("name,reflect", CELLS)
def test_removed_unknowns_match_reference(msh_dir, name, reflect):
    g = golden(name)
    case = _cell(msh_dir, g, reflect)
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
