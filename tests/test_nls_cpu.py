"""Host logic of the second-harmonic path (no GPU): determine_epsp
(solver.f90:393-433) against the reference's values on the dimer."""
import numpy as np

from cases import dimer
from conftest import golden
from mom2amd import determine_epsp


def test_determine_epsp_matches_reference(msh_dir):
    g = golden("dimer_s1")
    case = dimer(msh_dir, 1)
    for d in (1, 2):
        ep = determine_epsp(case["problem"], d, list(g["ri"]))
        assert np.array_equal(ep, g["sh_epsp_%d" % d])
    # the exterior domain's faces lie on both spheres: their outside is the
    # first other domain holding each face
    ep0 = determine_epsp(case["problem"], 0, list(g["ri"]))
    assert np.allclose(ep0, (np.asarray(g["ri"])[1] ** 2) * 8.8541878176e-12)
