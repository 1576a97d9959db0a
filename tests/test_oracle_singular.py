"""The int.f90 closed forms of the oracle (Hanninen et al. 2006) against
brute-force numerical integration over a triangle, for observation points
off the plane, in the plane outside the triangle and inside it.

The integrals are int_T R^p dS' and the RWG-weighted vector forms; the
brute force uses a Duffy-type split of T into three sub-triangles meeting at
the projection of r and tensor Gauss-Legendre rules on each (the 1/R
singularity is removed by the Jacobian)."""
import numpy as np
import pytest

import oracle
from cases import sphere


def tri_rule(P, order=60):
    """Points and weights on triangle P (3x3) -- collapsed Gauss-Legendre."""
    x, w = np.polynomial.legendre.leggauss(order)
    u = 0.5 * (x + 1)
    wu = 0.5 * w
    U, V = np.meshgrid(u, u, indexing="ij")
    W = np.outer(wu, wu) * U
    s, t = U * (1 - V), U * V  # barycentric of p2, p3; Jacobian U
    pts = P[0] + s[..., None] * (P[1] - P[0]) + t[..., None] * (P[2] - P[0])
    area = 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
    return pts.reshape(-1, 3), (W * 2 * area).ravel()


def brute(P, r, order=60):
    """Integrals of 1/R, R and of (r' - r) R^p with the triangle split at the
    projection rho of r (singularity at a sub-triangle vertex)."""
    n = np.cross(P[1] - P[0], P[2] - P[0])
    n /= np.linalg.norm(n)
    rho = r - np.dot(r - P[0], n) * n
    subs = [np.array([rho, P[i], P[(i + 1) % 3]]) for i in range(3)]
    # signed areas: rho outside -> some sub-triangles count negatively
    total_area = np.dot(np.cross(P[1] - P[0], P[2] - P[0]), n)
    res = {"m1": 0.0, "p1": 0.0, "vm1": np.zeros(3), "vp1": np.zeros(3)}
    for S in subs:
        sgn = np.sign(np.dot(np.cross(S[1] - S[0], S[2] - S[0]), n)) * np.sign(total_area)
        if sgn == 0:
            continue
        pts, w = tri_rule(S, order)
        R = np.linalg.norm(r - pts, axis=1)
        res["m1"] += sgn * np.sum(w / R)
        res["p1"] += sgn * np.sum(w * R)
        res["vm1"] += sgn * np.sum(w[:, None] * (pts - r) / R[:, None], axis=0)
        res["vp1"] += sgn * np.sum(w[:, None] * (pts - r) * R[:, None], axis=0)
    return res


@pytest.fixture(scope="module")
def op(msh_dir):
    case = sphere(msh_dir, 1)
    return oracle.OracleProblem(case["path"], 1e-9, case["domains"]), oracle.OracleProblem(
        case["path"], 1e-9, case["domains"]).mesh(0)


@pytest.mark.parametrize("where", ["above", "in_plane_outside", "inside"])
def test_intSm1_intS1_against_brute_force(op, where):
    O, m = op
    f = 7
    P = m["p"][m["fnodes"][f]]
    c = P.mean(axis=0)
    n = m["fn"][f]
    L = np.linalg.norm(P[1] - P[0])
    r = {"above": c + 0.3 * L * n + 0.1 * L * (P[1] - P[0]) / L,
         "in_plane_outside": c + 1.3 * (P[2] - c),
         "inside": 0.2 * P[0] + 0.5 * P[1] + 0.3 * P[2]}[where]
    b = brute(P, r)
    sm1 = O.int_scalar(0, 4, f, 0, r)
    s1 = O.int_scalar(0, 5, f, 0, r)
    assert abs(sm1 / b["m1"] - 1) < 1e-9, (sm1, b["m1"])
    assert abs(s1 / b["p1"] - 1) < 1e-9, (s1, b["p1"])


@pytest.mark.parametrize("e", [0, 1, 2])
def test_rwg_weighted_forms_against_brute_force(op, e):
    """intK2m1 = sign L/(2A) int (r' - p) / R, intK21 = ... * R, and
    intK1m1/intK11 = sign L/A int R^-1 / R, at an off-plane point."""
    O, m = op
    f = 11
    P = m["p"][m["fnodes"][f]]
    c = P.mean(axis=0)
    n = m["fn"][f]
    L0 = np.linalg.norm(P[1] - P[0])
    r = c + 0.25 * L0 * n + 0.2 * (P[0] - c)
    ge = m["fedges"][f, e]
    sign = 1.0 if m["efaces"][ge, 0] == f else -1.0
    bn = m["ebn"][ge, 0] if sign > 0 else m["ebn"][ge, 1]
    p = m["p"][bn]
    L, A = m["elen"][ge], m["farea"][f]
    b = brute(P, r)
    k2m1 = O.int_vector(0, 1, f, e, r)
    k21 = O.int_vector(0, 0, f, e, r)
    # int (r' - p) R^k = int (r' - r) R^k + (r - p) int R^k
    ref_m1 = sign * L / (2 * A) * (b["vm1"] + (r - p) * b["m1"])
    ref_p1 = sign * L / (2 * A) * (b["vp1"] + (r - p) * b["p1"])
    assert np.abs(k2m1 - ref_m1).max() / np.abs(ref_m1).max() < 1e-9
    assert np.abs(k21 - ref_p1).max() / np.abs(ref_p1).max() < 1e-9
    assert abs(O.int_scalar(0, 7, f, e, r) / (sign * L / A * b["m1"]) - 1) < 1e-9
    assert abs(O.int_scalar(0, 6, f, e, r) / (sign * L / A * b["p1"]) - 1) < 1e-9


def test_solid_angle_closed_form(op):
    """intSm3h is the signed solid angle: -> +/- 2 pi just above/below the
    face centre, and the three faces' angles seen from... sum consistently."""
    O, m = op
    f = 3
    P = m["p"][m["fnodes"][f]]
    c = P.mean(axis=0)
    n = m["fn"][f]
    L = np.linalg.norm(P[1] - P[0])
    assert abs(O.int_scalar(0, 3, f, 0, c + 1e-9 * L * n) - 2 * np.pi) < 1e-6
    assert abs(O.int_scalar(0, 3, f, 0, c - 1e-9 * L * n) + 2 * np.pi) < 1e-6
    # far away: solid angle ~ A cos / d^2
    d = 1e3 * L
    r = c + d * n
    A = m["farea"][f]
    assert abs(O.int_scalar(0, 3, f, 0, r) / (A / d**2) - 1) < 1e-5
