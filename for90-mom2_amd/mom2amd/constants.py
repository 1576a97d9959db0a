"""Physical constants and triangle rules exactly as the reference defines them.

constants.f90:8-16 writes pi as a default-REAL literal, so the value every
derived constant (mu0, degtorad, omega = 2 pi c0 / lambda) uses is the
float32 rounding of 3.141592653589 widened to double.  quad.f90:115-171
writes the 13-point rule the same way (float32 literals).
"""
import numpy as np

PI = float(np.float32(3.141592653589))
EPS0 = 8.8541878176e-12
MU0 = 4 * PI * 1e-7
C0 = 2.997924580003e8
ETA0 = 3.767303134622e2
DEGTORAD = PI / 180.0

_F = np.float32
_W13 = [-0.149570044467670, 0.175615257433204, 0.175615257433204, 0.175615257433204,
        0.053347235608839, 0.053347235608839, 0.053347235608839, 0.077113760890257,
        0.077113760890257, 0.077113760890257, 0.077113760890257, 0.077113760890257,
        0.077113760890257]
_N13 = [[0.333333333333333, 0.479308067841923, 0.260345966079038, 0.260345966079038,
         0.869739794195568, 0.065130102902216, 0.065130102902216, 0.638444188569809,
         0.638444188569809, 0.312865496004875, 0.312865496004875, 0.048690315425316,
         0.048690315425316],
        [0.333333333333333, 0.260345966079038, 0.260345966079038, 0.479308067841923,
         0.065130102902216, 0.065130102902216, 0.869739794195568, 0.048690315425316,
         0.312865496004875, 0.048690315425316, 0.638444188569809, 0.312865496004875,
         0.638444188569809],
        [0.333333333333333, 0.260345966079038, 0.479308067841923, 0.260345966079038,
         0.065130102902216, 0.869739794195568, 0.065130102902216, 0.312865496004875,
         0.048690315425316, 0.638444188569809, 0.048690315425316, 0.638444188569809,
         0.312865496004875]]


def tri_quad_data(name="tri_gl13"):
    """Triangle rule (quad.f90:277-315): weights[Q], nodes[Q,3] (barycentric;
    the point is node[:,2] p1 + node[:,0] p2 + node[:,1] p3, quad.f90:365)."""
    if name == "tri_gl13":
        w = np.array([float(_F(v)) for v in _W13])
        nodes = np.array([[float(_F(v)) for v in col] for col in _N13]).T.copy()
        return w, nodes
    if name == "tri_gl1":
        return np.array([1.0]), np.full((1, 3), 0.333333333333333)
    raise ValueError("Unrecognized quadrature type %s!" % name)


def omega_of(wl):
    """Angular frequency of vacuum wavelength wl (solver.f90:78)."""
    return 2.0 * PI * C0 / wl
