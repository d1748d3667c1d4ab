"""MI355X (gfx950) hot path of COVOF 2D: Weymouth-Yue VOF advection with
PLIC/ELVIRA reconstruction, the SP velocity predictor (momentum block) and the
variable-density pressure-Poisson solve of the standard projection, as HIP
kernels behind the C-ABI of ``include/covof_hip.h``.

This Python layer mirrors the reference's operator interface for the path
(``advect_vof``, ``reconstruct_interface``, ``update_physical_properties``,
``cut_box_area``/``compute_alpha``, the momentum block's
``assemble_preconditioner_blocks`` / ``assemble_rhs`` / ``apply_system_matrix``
/ ``solve_linear_system``, the SP Poisson block, the velocity correction) over torch CUDA tensors in the reference's field layout: a field
``F(0:nx+1, 0:ny+1)`` is a float64 tensor of shape ``(ny+2, nx+2)`` with
``F[j, i]``; Poisson work vectors are ``(ny, nx)``.  PyTorch only supplies
device memory and streams; every computation runs in libcovof_hip.
"""
from ._lib import (BC_DIRICHLET, BC_FREESLIP, BC_NEUMANN, BC_PERIODIC, LINEAR_BICGSTAB2,  # noqa: F401
                   LINEAR_CG, PRECISION_FP64, PRECISION_MIXED, PRECOND_JACOBI, PRECOND_MG, PRECOND_NONE,
                   PRECOND_SGS_RB, PRECOND_SSOR, TIME_SCHEME_EULER, TIME_SCHEME_GEAR, CONV_UPWIND,
                   CONV_CENTERED, CONV_TVD_SUPERBEE, CONV_LIN_PICARD, CONV_LIN_ADAMS_BASHFORTH, CONV_LIN_NEWTON,
                   PRECOND_BLOCK_TRIANGULAR, SOLVER_MONOLITHIC, SOLVER_STANDARD_PROJECTION, CURVATURE_NORMALS,
                   CURVATURE_HEIGHT_FUNCTION, REDUCTION_TREE, REDUCTION_REFERENCE, OPT_MIXED_STALL_WINDOW,
                   OPT_MIXED_CYCLE, OPT_MIXED_LOOP, OPT_MG_COARSE, OPT_MG_SWEEPS, OPT_MG_OMEGA, OPT_MG_FUSED,
                   OPT_MIXED_MG_FORM, OPT_MG_PROJECT, CovofError)
from .context import Context, SolveResult  # noqa: F401
from .autograd import PressureSolve, Projection, pressure_solve, projection  # noqa: F401
from .sp_run import StandardProjectionRun  # noqa: F401
from .coupled_adjoint import KineticEnergyAdjoint, kinetic_energy_after  # noqa: F401
