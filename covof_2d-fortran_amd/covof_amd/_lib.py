"""ctypes loader of libcovof_hip.so (the HIP kernels + C-ABI of include/covof_hip.h).

The library is built in-tree (``covof_2d-fortran_amd/libcovof_hip.so``) by
``__graft_entry__.build()`` / ``make -C covof_2d-fortran_amd/csrc``.  There is
no fallback: if the library or a GPU is missing, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as ct
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libcovof_hip.so")

# status codes (include/covof_hip.h)
OK, NOT_CONVERGED, BREAKDOWN, NONFINITE, BAD_ARGUMENT, HIP_ERROR = range(6)
STATUS_NAMES = {OK: "ok", NOT_CONVERGED: "not converged", BREAKDOWN: "breakdown",
                NONFINITE: "non-finite", BAD_ARGUMENT: "bad argument", HIP_ERROR: "HIP error"}

# reference selector values (src/module_constants.f90)
BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC, BC_FREESLIP = 1, 2, 3, 4
LINEAR_BICGSTAB2, LINEAR_CG = 1, 3
PRECOND_NONE, PRECOND_JACOBI, PRECOND_SSOR, PRECOND_BLOCK_TRIANGULAR = 0, 1, 2, 4
SOLVER_MONOLITHIC, SOLVER_STANDARD_PROJECTION = 2, 3
CURVATURE_NORMALS, CURVATURE_HEIGHT_FUNCTION = 0, 1
# parallel preconditioners of this library (no reference counterpart)
PRECOND_SGS_RB, PRECOND_MG = 16, 17
PRECISION_FP64, PRECISION_MIXED = 0, 1
OPT_MIXED_STALL_WINDOW, OPT_MIXED_CYCLE, OPT_MIXED_LOOP, OPT_MG_COARSE, OPT_MG_SWEEPS, OPT_MG_OMEGA = 1, 2, 3, 4, 5, 6
OPT_MG_FUSED, OPT_MIXED_MG_FORM, OPT_MG_PROJECT = 7, 8, 9
(RESERVE_MOMENTUM, RESERVE_MULTIGRID, RESERVE_SGS, RESERVE_MIXED_LOOP, RESERVE_SURFACE, RESERVE_ADJOINT,
 RESERVE_HOST_MIRRORS, RESERVE_ALL) = 1, 2, 4, 8, 16, 32, 64, 127
REDUCTION_TREE, REDUCTION_REFERENCE = 0, 1
TIME_SCHEME_EULER, TIME_SCHEME_GEAR = 1, 2
CONV_UPWIND, CONV_CENTERED, CONV_TVD_SUPERBEE = 1, 2, 3
CONV_LIN_PICARD, CONV_LIN_ADAMS_BASHFORTH, CONV_LIN_NEWTON = 1, 2, 3


class MomentumCfg(ct.Structure):
    """covof_momentum_cfg (include/covof_hip.h)."""
    _fields_ = [("time_scheme", ct.c_int), ("has_nm1", ct.c_int), ("enable_convection", ct.c_int),
                ("convective_scheme", ct.c_int), ("convective_linearization", ct.c_int),
                ("gx", ct.c_double), ("gy", ct.c_double), ("bc_u", ct.c_double * 4), ("bc_v", ct.c_double * 4),
                ("solver_type", ct.c_int)]

class Obstacle(ct.Structure):
    """covof_obstacle (include/covof_hip.h): config_t's obs_* fields."""
    _fields_ = [("enabled", ct.c_int), ("type", ct.c_int), ("x0", ct.c_double), ("y0", ct.c_double),
                ("x1", ct.c_double), ("y1", ct.c_double), ("xc", ct.c_double), ("yc", ct.c_double),
                ("r", ct.c_double), ("perm", ct.c_double)]


OBS_BOX, OBS_CYLINDER = 1, 2

# every symbol the header declares: name -> (restype, argtypes)
_P = ct.c_void_p
_I = ct.c_int
_D = ct.c_double
_S = ct.c_char_p
_IP = ct.POINTER(ct.c_int)
_DP = ct.POINTER(ct.c_double)
SIGNATURES = {
    "covof_ctx_create": (_I, [_I, _I, _D, _D, _IP, ct.POINTER(_P)]),
    "covof_ctx_destroy": (_I, [_P]),
    "covof_ctx_set_stream": (_I, [_P, _P]),
    "covof_ctx_synchronize": (_I, [_P]),
    "covof_ctx_set_reduction_order": (_I, [_P, _I]),
    "covof_ctx_set_obstacle": (_I, [_P, ct.c_void_p]),
    "covof_ctx_set_vof_top_profile": (_I, [_P, _S, _D, _D]),
    "covof_update_properties_cs_d": (_I, [_P, _P, _P, _D, _D, _D, _D] + [_P] * 5),
    "covof_last_error": (_S, []),
    "covof_geom_batch_d": (_I, [_P, _I] + [_P] * 10),
    "covof_reconstruct_interface_d": (_I, [_P, _S, _P, _P, _P, _P]),
    "covof_reconstruct_interface_noalloc_d": (_I, [_I, _I, _D, _D, _IP, _S, _P, _P, _P, _P, _P]),
    "covof_vof_advect_d": (_I, [_P, _P, _P, _P, _D, _S, _S, _IP, _I, _I, _D, _D, _D, _D] + [_P] * 5),
    "covof_vof_advect": (_I, [_P, _DP, _DP, _DP, _D, _S, _S, _IP, _I, _I, _D, _D, _D, _D] + [_DP] * 5),
    "covof_update_properties_d": (_I, [_P, _P, _D, _D, _D, _D] + [_P] * 5),
    "covof_curvature_csf_d": (_I, [_P, _P, _P, _I, _D, _P, _P, _P]),
    "covof_curvature_csf_method_d": (_I, [_P, _P, _P, _I, _D, _I, _P, _P, _P]),
    "covof_curvature_csf": (_I, [_P, _P, _P, _I, _D, _P, _P, _P]),
    "covof_smooth_diffusion_d": (_I, [_P, _P, _I, _P]),
    "covof_smooth_diffusion": (_I, [_P, _P, _I, _P]),
    "covof_svof_smooth_d": (_I, [_P, _P, _D, _P]),
    "covof_svof_smooth": (_I, [_P, _P, _D, _P]),
    "covof_poisson_assemble_d": (_I, [_P, _P, _P, _P, _P, _D, _P]),
    "covof_poisson_matvec_d": (_I, [_P, _P, _P]),
    "covof_poisson_precond_d": (_I, [_P, _I, _P, _P]),
    "covof_poisson_solve_d": (_I, [_P, _P, _P, _I, _I, _D, _I, _I, _IP, _DP]),
    "covof_sp_poisson_d": (_I, [_P, _P, _P, _P, _P, _P, _D, _I, _I, _D, _I, _I, _IP, _DP]),
    "covof_sp_poisson": (_I, [_P, _DP, _DP, _DP, _DP, _DP, _D, _I, _I, _D, _I, _I, _IP, _DP]),
    "covof_sp_correct_d": (_I, [_P, _P, _P, _P, _D, _P, _P]),
    "covof_sp_correct": (_I, [_P, _DP, _DP, _DP, _D, _DP, _DP]),
    "covof_poisson_adjoint_d": (_I, [_P, _P, _D, _P, _P, _P, _P, _I, _I, _D, _I, _I, _IP, _DP]),
    "covof_sp_pressure_d": (_I, [_P, _P, _P]),
    "covof_projection_adjoint_d": (_I, [_P, _P, _P, _P, _D] + [_P] * 7 + [_I, _D, _I, _I, _IP, _DP]),
    "covof_vof_advect_adjoint_d": (_I, [_P, _P, _P, _P, _D, ct.c_char_p, _I, _I, _I, _P, _P, _P, _P]),
    "covof_vof_fill_halo_d": (_I, [_P, _P]),
    "covof_ctx_set_solver_option": (_I, [_P, _I, _D]),
    "covof_ctx_mixed_solve_info": (_I, [_P, _IP, _IP]),
    "covof_ctx_reserve": (_I, [_P, _I]),
    "covof_kinetic_energy_d": (_I, [_P, _P, _P, _P, _D, _D, _DP, _P, _P, _P]),
    "covof_properties_adjoint_d": (_I, [_P, _P, _D, _D, _P, _P]),
    "covof_inertial_predictor_d": (_I, [_P, _I, _D, _D, _D, _P, _P, _P, _P, _P, _P]),
    "covof_inertial_predictor_adjoint_d": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "covof_momentum_size": (_I, [_P, _I, _IP, _IP, _IP]),
    "covof_momentum_set_fields_d": (_I, [_P, ct.c_void_p] + [_P] * 12 + [_D]),
    "covof_momentum_diag_d": (_I, [_P, _P]),
    "covof_momentum_rhs_d": (_I, [_P, _P]),
    "covof_momentum_matvec_d": (_I, [_P, _P, _P]),
    "covof_momentum_solve_d": (_I, [_P, _P, _P, _I, _I, _D, _I, _IP, _DP]),
    "covof_momentum_precond_d": (_I, [_P, _I, _P, _P]),
    "covof_momentum_set_fields": (_I, [_P, ct.c_void_p] + [_DP] * 12 + [_D]),
    "covof_momentum_precond": (_I, [_P, _I, _I, _DP, _DP]),
    "covof_velocity_bc_d": (_I, [_P, ct.c_void_p, _P, _P, _I]),
    "covof_sp_predictor_d": (_I, [_P, ct.c_void_p] + [_P] * 12 + [_D, _I, _I, _D, _I, _IP, _DP]),
    "covof_sp_predictor": (_I, [_P, ct.c_void_p] + [_DP] * 12 + [_D, _I, _I, _D, _I, _IP, _DP]),
    "covof_monolithic_step_d": (_I, [_P, ct.c_void_p] + [_P] * 13 + [_D, _I, _I, _D, _I, _IP, _DP]),
    "covof_monolithic_step": (_I, [_P, ct.c_void_p] + [_DP] * 13 + [_D, _I, _I, _D, _I, _IP, _DP]),
    "covof_history_d": (_I, [_P] + [_P] * 6 + [_DP]),
    "covof_cell_output_d": (_I, [_P] + [_P] * 5),
    "covof_ctx_profile": (_I, [_P, _I]),
    "covof_ctx_profile_read": (_I, [_P, _I, _DP, ct.POINTER(ct.c_long)]),
}

# kernel tags of covof_ctx_profile_read
KERNEL_PCG_A_FP64, KERNEL_PCG_B_FP64, KERNEL_PCG_A_FP32, KERNEL_PCG_B_FP32, KERNEL_VOF_RECON, \
    KERNEL_VOF_UPDATE, KERNEL_MOM_MATVEC = range(7)

_lib = None


class CovofError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str = ""):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}" + (f" ({msg})" if msg else ""))


def load(path: str = LIB_PATH) -> ct.CDLL:
    """Load the shared library and bind every exported symbol (no GPU needed)."""
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with __graft_entry__.build() "
                          "or `make -C covof_2d-fortran_amd/csrc`")
    L = ct.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def lib() -> ct.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def check(status: int, where: str, allow=(OK,)) -> int:
    if status in allow:
        return status
    msg = lib().covof_last_error()
    raise CovofError(status, where, msg.decode() if msg else "")
