"""ctypes mirror of include/gales_mc.h (structs and constants).

The layouts here must match the C header field for field; tests/test_abi.py
checks sizes against a C compile of the header.
"""
import ctypes as C

GMC_OK = 0
GMC_ERR_ARG = -1
GMC_ERR_HIP = -2
GMC_ERR_MESH = -3
GMC_ERR_NUMERIC = -4
GMC_ERR_UNSUPPORTED = -5
GMC_ERR_STATE = -6
GMC_ERR_RANGE = -7

GMC_MAX_COMP = 4
GMC_MAX_DOF = 8
GMC_MAX_FLAG = 16

GMC_COMP_CUSTOM = 0
GMC_COMP_IDEAL_GAS = 1
GMC_COMP_MAGMA = 2
GMC_MU_NONE = 0
GMC_MU_ISHII_ZUBER = 1
GMC_MU_PAL_2003 = 2
GMC_MU_EINSTEIN_ROSCOE = 1
GMC_MU_CARICCHI = 2
GMC_TAU_NON_DIAG_COMP_2001 = 2001
GMC_TAU_NON_DIAG_COMP_2019 = 2019
GMC_TAU_DIAG_INCOMP_2007 = 2007
GMC_TAU_DIAG_2014 = 2014
GMC_PRECOND_ILU0 = 3
GMC_PREC_FP64 = 0
GMC_PREC_FP32 = 1

GMC_N_TAU11, GMC_N_TAU22, GMC_N_TAU33 = 1, 2, 4
GMC_N_TAU12, GMC_N_TAU13, GMC_N_TAU23 = 8, 16, 32

GMC_ASM_MATRIX = 1
GMC_ASM_RHS = 2
GMC_ASM_FP32 = 4

GMC_VJP_EXACT = 0
GMC_VJP_FROZEN = 1


class Magma(C.Structure):
    _fields_ = [("oxide_wf", C.c_double * 10), ("melt_kappa", C.c_double),
                ("h2o_wf", C.c_double), ("co2_wf", C.c_double),
                ("h2o_l_cp", C.c_double), ("h2o_l_kappa", C.c_double),
                ("co2_l_cp", C.c_double), ("co2_l_kappa", C.c_double),
                ("crystal_vf", C.c_double), ("crystal_rho", C.c_double),
                ("gas_on_mu", C.c_int), ("crystal_on_mu", C.c_int),
                ("n_p", C.c_int), ("n_T", C.c_int),
                ("p_table", C.POINTER(C.c_double)), ("T_table", C.POINTER(C.c_double)),
                ("wf_h2o_l", C.POINTER(C.c_double)), ("wf_co2_l", C.POINTER(C.c_double))]


class Component(C.Structure):
    _fields_ = [("kind", C.c_int),
                ("rho", C.c_double), ("mu", C.c_double), ("cp", C.c_double),
                ("kappa", C.c_double), ("alpha", C.c_double), ("beta", C.c_double),
                ("molar_mass", C.c_double), ("gamma", C.c_double),
                ("diffusivity", C.c_double),
                ("magma", Magma)]


class Neumann(C.Structure):
    _fields_ = [("tau_mask", C.c_uint), ("tau", C.c_double * 6),
                ("q_mask", C.c_uint), ("q", C.c_double * 3),
                ("J_mask", C.c_uint * 3), ("J", (C.c_double * GMC_MAX_COMP) * 3),
                ("mflux_mask", C.c_uint), ("mflux", C.c_double * 3)]


class Params(C.Structure):
    _fields_ = [("dim", C.c_int), ("nc", C.c_int),
                ("comp", Component * GMC_MAX_COMP),
                ("gravity", C.c_double * 3),
                ("p_ref", C.c_double), ("v_ref", C.c_double * 3), ("T_ref", C.c_double),
                ("dt", C.c_double),
                ("steady_state", C.c_int),
                ("tau_variant", C.c_int),
                ("dc_2006", C.c_int),
                ("dc_sharp", C.c_double),
                ("dc_scale_fact", C.c_double),
                ("dirichlet_mask", C.c_uint * GMC_MAX_FLAG),
                ("neumann", Neumann * GMC_MAX_FLAG),
                ("precision", C.c_int)]


class Mesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("n_nodes", C.c_int), ("n_elems", C.c_int),
                ("n_sides", C.c_int),
                ("coords", C.POINTER(C.c_double)),
                ("conn", C.POINTER(C.c_int)),
                ("node_flag", C.POINTER(C.c_int)),
                ("side_elem", C.POINTER(C.c_int)),
                ("side_local", C.POINTER(C.c_int)),
                ("side_flag", C.POINTER(C.c_int))]


class Info(C.Structure):
    _fields_ = [("dim", C.c_int), ("nc", C.c_int), ("ndof", C.c_int), ("nen", C.c_int),
                ("n_nodes", C.c_int), ("n_elems", C.c_int), ("n_sides", C.c_int),
                ("nnzb", C.c_longlong), ("n_colors", C.c_int), ("n_bd_elems", C.c_int),
                ("n_bd_colors", C.c_int), ("device_bytes", C.c_size_t), ("scatter", C.c_int),
                ("ebuf_chunks", C.c_int),
                ("fp32_mfma", C.c_int)]


class Solver(C.Structure):
    _fields_ = [("restart", C.c_int), ("max_iters", C.c_int), ("rel_tol", C.c_double),
                ("precond", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("rel_residual", C.c_double),
                ("abs_residual", C.c_double), ("seconds", C.c_double)]


class Timing(C.Structure):
    _fields_ = [("assemble_s", C.c_double), ("clear_s", C.c_double),
                ("interior_s", C.c_double), ("boundary_s", C.c_double), ("gather_s", C.c_double),
                ("interior_launches", C.c_int), ("boundary_launches", C.c_int)]


# every symbol declared in include/gales_mc.h
EXPORTS = [
    "gmc_create", "gmc_destroy", "gmc_last_error", "gmc_get_info", "gmc_set_stream",
    "gmc_set_dt", "gmc_set_material", "gmc_set_state", "gmc_get_state",
    "gmc_set_state_device", "gmc_assemble", "gmc_residual", "gmc_get_timing",
    "gmc_element", "gmc_get_pattern", "gmc_get_system", "gmc_spmv", "gmc_jvp", "gmc_solve",
    "gmc_get_solution", "gmc_newton_update", "gmc_predictor", "gmc_refine", "gmc_vjp",
    "gmc_vjp_device", "gmc_residual_device", "gmc_corrector", "gmc_dof_trimmer", "gmc_rhs_norm",
    "gmc_rhs_norm_noalloc", "gmc_adaptive_dt", "gmc_hydrostatic_profile",
]
