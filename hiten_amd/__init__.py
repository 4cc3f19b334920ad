"""hiten_amd -- MI355X (gfx950) implementation of hiten's CR3BP propagation path.

The vector field with its 6x6 variational equations, fixed-step RK8, adaptive
DOP853 with dense output, and the Poincare section-crossing search run as HIP
kernels (libhiten_hip.so, C-ABI in include/hiten_hip.h).  The Python modules
mirror the reference's call signatures:

* hiten_amd.dynamics    rtbp_dynsys, variational_dynsys, _propagate_dynsys,
                        _compute_stm, _compute_monodromy
* hiten_amd.integrators RungeKutta, AdaptiveRK, FixedRK, _Solution
* hiten_amd.singlehit   _x/_y/_z_plane_crossing
* hiten_amd.batch       batched entry points over many trajectories
"""

__version__ = "0.1.0"
