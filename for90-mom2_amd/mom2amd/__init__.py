"""Host-side mirror of the for90-mom2 hot-path interface over libmomgpu.

Routine names and argument meaning follow the reference (sysmat.f90,
source.f90, linalg.f90, cs.f90); all arithmetic runs in the HIP library.
"""
from .constants import C0, DEGTORAD, EPS0, ETA0, MU0, PI, omega_of, tri_quad_data  # noqa: F401
from .mesh import DomainMesh, Problem, build_problem, load_problem  # noqa: F401
from .gpu import (Context, DeviceMesh, System, computeD, computeH, computeK,  # noqa: F401
                  cs_prtsrf, far_fields, rcs, rcs_ext, solve_multi_linsys, srcvec, sysmat_pmchwt)
from .nlsurf import (determine_epsp, nlsurf_coef, sh_power, sh_power_translation_gradient, solve_sh,  # noqa: F401
                     sysmat_pmchwt_nls)
from .adjoint import cs_abs_gradient, pmchwt_deps_apply, sigma_abs_eps_gradient, sigma_abs_vertex_gradient  # noqa: F401,E501
from .symmetry import (SymmetricSystems, cs_prtsrf_symm, rcs_symm, srcvec_symm,  # noqa: F401
                       edge_bc as symm_edge_bc, group as symm_group)
