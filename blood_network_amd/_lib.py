"""ctypes binding of libbnamg.so (the MI355X AMG-PCG library).

The library is built in-tree (``make -C blood_network_amd/csrc``, or
``__graft_entry__.build()``).  Importing this module loads it and raises
if it is missing: there is no Python or CPU fallback for the solve.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbnamg.so")

IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)
VP = C.c_void_p


class CSR(C.Structure):
    """dcsr_mat (C_amg1/include/agg_lib.h:37-44)."""
    _fields_ = [("row", C.c_int), ("col", C.c_int), ("nnz", C.c_int),
                ("ia", IP), ("ja", IP), ("val", DP)]


class IcholFactor(C.Structure):
    """ichol_factor (C_amg1/include/amg.h:29-35)."""
    _fields_ = [("n", C.c_int), ("ia", IP), ("ja", IP), ("val", DP), ("nnz", C.c_int)]


class AMGLevel(C.Structure):
    """amg_level (C_amg1/include/amg.h:42-53)."""
    _fields_ = [("A", CSR), ("P", CSR), ("A_filtered", CSR), ("mis", IP), ("isolated", IP),
                ("cf_order", IP), ("n", C.c_int), ("n_coarse", C.c_int), ("L_ichol", IcholFactor),
                ("l1_diag", DP)]


AMG_MAX_LEVELS = 20


class AMGData(C.Structure):
    """amg_data (C_amg1/include/amg.h:56-69): hierarchy plus options, the
    argument of the library-level entry points."""
    _fields_ = [("levels", AMGLevel * AMG_MAX_LEVELS), ("num_levels", C.c_int),
                ("threshold", C.c_double), ("min_size", C.c_int), ("max_levels", C.c_int),
                ("maxit_prol", C.c_int), ("lump", C.c_int), ("nu", C.c_int), ("smoother", C.c_int),
                ("omega_jacobi", C.c_double), ("coarsening", C.c_int)]


class Options(C.Structure):
    """bn_options (include/bn_amg_ext.h)."""
    _fields_ = [("coarsening", C.c_int), ("threshold", C.c_double),
                ("min_size", C.c_int), ("max_levels", C.c_int), ("nu", C.c_int),
                ("smoother", C.c_int), ("omega", C.c_double),
                ("precision", C.c_int), ("coarse_solver", C.c_int),
                ("fine_ic", C.c_int), ("ic_sweeps", C.c_int),
                ("max_colours", C.c_int), ("flexible", C.c_int),
                ("stop", C.c_int), ("symmetrize", C.c_int), ("device", C.c_int),
                ("verbose", C.c_int), ("graph", C.c_int),
                ("block_smoother", C.c_int), ("chunk_rows", C.c_int), ("strength", C.c_int),
                ("gs_exact", C.c_int)]


class Nwk(C.Structure):
    """bn_nwk (include/bn_netgen.h): a vessel-network case."""
    _fields_ = [("np", C.c_int), ("nf", C.c_int), ("nbc", C.c_int), ("pts", C.POINTER(C.c_double)),
                ("face", C.POINTER(C.c_int)), ("dia", C.POINTER(C.c_double)), ("bc_node", C.POINTER(C.c_int)),
                ("bc_val", C.POINTER(C.c_double))]


class NwkSys(C.Structure):
    """bn_nwk_sys (include/bn_netgen.h): its interior pressure system."""
    _fields_ = [("n", C.c_int), ("nnz", C.c_int), ("ia", C.POINTER(C.c_int)), ("ja", C.POINTER(C.c_int)),
                ("val", C.POINTER(C.c_double)), ("b", C.POINTER(C.c_double)), ("interior", C.POINTER(C.c_int))]


class SolveInfo(C.Structure):
    """bn_solve_info (include/bn_amg_ext.h)."""
    _fields_ = [("iterations", C.c_int), ("rel_bnorm", C.c_double),
                ("relres", C.c_double), ("solve_ms", C.c_double),
                ("wall_ms", C.c_double), ("setup_ms", C.c_double),
                ("upload_ms", C.c_double), ("num_levels", C.c_int)]


class KernelStat(C.Structure):
    """bn_kernel_stat (include/bn_amg_ext.h): one kernel symbol's census."""
    _fields_ = [("name", C.c_char * 80), ("launches", C.c_int), ("shapes", C.c_int), ("bytes", C.c_double),
                ("us_solve", C.c_double), ("us_cold", C.c_double)]


class NetParams(C.Structure):
    """bn_netgen_params (include/bn_netgen.h)."""
    _fields_ = [("L", C.c_int), ("depth", C.c_int), ("n_trees", C.c_int),
                ("seed", C.c_uint64), ("p_art", C.c_double), ("p_ven", C.c_double),
                ("mu", C.c_double), ("spacing", C.c_double), ("cap_r", C.c_double),
                ("cap_sigma", C.c_double), ("root_r", C.c_double), ("boundary_roots", C.c_int)]


class Network(C.Structure):
    """bn_network (include/bn_netgen.h)."""
    _fields_ = [("n_nodes", C.c_int), ("n", C.c_int), ("n_edges", C.c_int),
                ("nnz", C.c_int), ("n_dirichlet", C.c_int),
                ("ia", IP), ("ja", IP), ("val", DP), ("b", DP),
                ("eu", IP), ("ev", IP), ("radius", DP), ("length", DP), ("g", DP),
                ("pos_uv", IP), ("pos_vu", IP), ("diag_pos", IP),
                ("p_dirichlet", DP), ("dirichlet_is_art", IP), ("node_unknown", IP)]


BN_ERR_NO_DEVICE = -1000001
BN_ERR_HIP = -1000002
BN_ERR_ARG = -1000003

SMOOTHER_L1, SMOOTHER_GS, SMOOTHER_CF_GS, SMOOTHER_JACOBI = 0, 1, 2, 3
COARSEN_RS, COARSEN_PMIS, COARSEN_HMIS = 0, 1, 2
PREC_FP64, PREC_FP32 = 0, 1
COARSE_AUTO, COARSE_DENSE, COARSE_ICPCG = 0, 1, 2
IC_EXACT, IC_JACOBI, IC_NONE = 0, 1, 2
# lanes-argument bits of the SpMV launchers: 8 or 16 load rounds per lane
LANES_DEEP8, LANES_DEEP16 = 0x100, 0x200
STOP_BNORM, STOP_RELRES = 0, 1
STRENGTH_C, STRENGTH_MEX = 0, 1


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C blood_network_amd/csrc` "
            "(the solve has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "amg_solver_create": (VP, [C.c_int, C.c_int, IP, IP, DP, C.c_double, C.c_int, C.c_int, C.c_int]),
        "amg_solver_update": (C.c_int, [VP, C.c_int, C.c_int, IP, IP, DP]),
        "amg_solver_set_matrix": (C.c_int, [VP, C.c_int, C.c_int, IP, IP, DP]),
        "amg_solver_solve": (C.c_int, [VP, DP, DP, C.c_double, C.c_int, C.c_int]),
        "amg_solver_set_smoother": (None, [VP, C.c_int, C.c_double]),
        "amg_solver_size": (C.c_int, [VP]),
        "amg_solver_num_levels": (C.c_int, [VP]),
        "amg_solver_free": (None, [VP]),
        "bn_options_default": (None, [C.POINTER(Options)]),
        "bn_solver_create": (VP, [C.c_int, C.c_int, IP, IP, DP, C.POINTER(Options)]),
        "bn_solver_solve_device": (C.c_int, [VP, VP, VP, C.c_double, C.c_int, C.c_int, C.POINTER(SolveInfo)]),
        "bn_solver_solve_info": (C.c_int, [VP, DP, DP, C.c_double, C.c_int, C.c_int, C.POINTER(SolveInfo)]),
        "bn_solver_precond": (C.c_int, [VP, DP, DP, C.c_int]),
        "bn_solver_last_info": (C.c_int, [VP, C.POINTER(SolveInfo)]),
        "bn_solver_set_mode": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "bn_solver_level_dims": (C.c_int, [VP, C.c_int, IP]),
        "bn_solver_level_csr": (C.c_int, [VP, C.c_int, C.c_int, IP, IP, DP]),
        "bn_solver_level_split": (C.c_int, [VP, C.c_int, IP, IP, DP]),
        "bn_solver_ichol": (C.c_int, [VP, C.c_int, IP, IP, DP]),
        "bn_solver_schedule": (C.c_int, [VP, C.c_int, IP, IP, IP]),
        "bn_solver_ic_levels": (C.c_int, [VP, IP, IP]),
        "bn_solver_time_kernel": (C.c_int, [VP, C.c_int, C.c_int, DP, DP]),
        "bn_solver_kernel_census": (C.c_int, [VP, VP, VP, C.c_double, C.c_int, C.c_int, C.POINTER(KernelStat),
                                              C.c_int, IP]),
        "bn_dev_alloc": (C.c_int, [C.POINTER(VP), C.c_size_t]),
        "bn_dev_free": (C.c_int, [VP]),
        "bn_dev_copy": (C.c_int, [VP, VP, C.c_size_t, C.c_int]),
        "bn_dev_sync": (C.c_int, []),
        "bn_device_available": (C.c_int, []),
        "bn_last_error": (C.c_char_p, []),
        "bn_solver_adjoint_grad": (C.c_int, [VP, DP, DP, C.c_int, IP, IP, DP, DP, C.c_double, C.c_int, DP, DP, DP]),
        "bn_solver_adjoint_grad_device": (C.c_int, [VP, VP, VP, C.c_int, VP, VP, VP, VP, C.c_int, C.c_double,
                                                    C.c_int, VP, VP, VP]),
        "bn_solver_adjoint_grad_noalloc": (C.c_int, [VP, DP, DP, C.c_int, IP, IP, DP, DP, C.c_double, C.c_int,
                                                     DP, DP, DP]),
        "bn_solver_set_values_device": (C.c_int, [VP, C.c_int, VP]),
        "bn_solver_rebuild_values_device": (C.c_int, [VP, C.c_int, VP]),
        "bn_solver_coarse_inverse": (C.c_int, [VP, IP, DP]),
        "bn_dense_inverse": (C.c_int, [C.POINTER(CSR), DP]),
        "bn_solver_solve_batch": (C.c_int, [VP, C.c_int, DP, DP, C.c_double, C.c_int, IP, C.POINTER(SolveInfo)]),
        "bn_solver_solve_segments": (C.c_int, [VP, C.c_int, IP, DP, DP, C.c_double, C.c_int, IP,
                                               C.POINTER(SolveInfo)]),
        "bn_solver_solve_segments_device": (C.c_int, [VP, C.c_int, IP, VP, VP, C.c_double, C.c_int, IP,
                                                      C.POINTER(SolveInfo)]),
        "bn_solver_solve_block": (C.c_int, [VP, C.c_int, DP, DP, C.c_double, C.c_int, IP, C.POINTER(SolveInfo)]),
        "bn_solver_solve_block_device": (C.c_int, [VP, C.c_int, VP, VP, C.c_double, C.c_int, IP,
                                                   C.POINTER(SolveInfo)]),
        "bn_netgen_default_params": (None, [C.POINTER(NetParams)]),
        "bn_netgen_build": (C.c_int, [C.POINTER(NetParams), C.POINTER(Network)]),
        "bn_netgen_assemble": (None, [C.POINTER(Network), DP, DP, DP, DP]),
        "bn_netgen_conductance": (None, [C.POINTER(Network), DP, C.c_double, DP]),
        "bn_netgen_free": (None, [C.POINTER(Network)]),
        "bn_nwk_read": (C.c_int, [C.c_char_p, C.POINTER(Nwk)]),
        "bn_nwk_free": (None, [C.POINTER(Nwk)]),
        "bn_nwk_system": (C.c_int, [C.POINTER(Nwk), C.POINTER(NwkSys)]),
        "bn_nwk_sys_free": (None, [C.POINTER(NwkSys)]),
        "bn_write_pressures": (C.c_int, [C.c_char_p, DP, C.c_int]),
        "bn_write_labelled": (C.c_int, [C.c_char_p, IP, DP, C.c_int]),
        "amg_build_hierarchy": (None, [VP, C.POINTER(CSR)]),
        "amg_free": (None, [VP]),
        "ichol_compute": (None, [C.POINTER(CSR), VP]),
        "ichol_free": (None, [VP]),
        "ichol_solve": (None, [VP, DP, DP]),
        "amg_backslash": (None, [VP, C.c_int, DP, DP, C.c_int]),
        "amg_fwdslash": (None, [VP, C.c_int, DP, DP, C.c_int]),
        "amg_vcycle_precond": (None, [VP, DP, DP]),
        "amg_ichol_precond": (None, [VP, C.POINTER(CSR), VP, DP, DP]),
        # kernel launchers (device pointers, stream)
        "bn_csr_apply": (None, [C.c_char, C.c_int, C.c_int, VP, VP, VP, VP, VP, VP, C.c_int, VP, VP, VP]),
        "bn_csr_apply_binned": (None, [C.c_char, C.c_int, C.c_int, VP, VP, VP, VP, VP, VP, C.c_int, C.c_int,
                                        C.c_int, VP, VP]),
        "bn_gs_color_blocks": (None, [C.c_char, C.c_int, C.c_int, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
        "bn_gs_color_map": (None, [C.c_char, C.c_int, C.c_int, VP, VP, VP, VP, VP, VP, VP, C.c_int, VP, VP, VP]),
        "bn_gs_color_zero_map": (None, [C.c_char, C.c_int, C.c_int, C.c_int, VP, VP, VP, VP, VP]),
        "bn_gs_color_zero_colour": (None, [C.c_char, C.c_int, C.c_int, VP, VP, VP, VP, VP]),
        "bn_block_gs": (None, [C.c_char, C.c_int, C.c_int, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP,
                                C.c_int, C.c_int, VP, VP, VP]),
        "bn_chunk_order": (C.c_int, [C.POINTER(CSR), IP, C.c_int, C.c_int, IP, C.POINTER(IP), C.POINTER(IP),
                                     C.POINTER(IP)]),
        "bn_csr_apply_seg": (None, [C.c_char, C.c_int, C.c_int, VP, VP, VP, VP, VP, VP, C.c_int, C.c_int, C.c_int,
                                     VP, VP, VP, C.c_int, VP, VP, VP]),
        "bn_seg_reset": (None, [C.c_int, VP, C.c_int, VP, VP]),
        "bn_csr_apply_blocks": (None, [C.c_char, C.c_int, C.c_int, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
        "bn_csr_resid_blocks_scaled": (None, [C.c_char, C.c_int, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
        "bn_row_blocks": (C.c_int, [IP, C.c_int, C.c_int, C.c_int, IP]),
        "bn_colour_order": (C.c_int, [C.POINTER(CSR), IP, C.c_int, IP, C.POINTER(IP)]),
        "bn_level_sets": (C.c_int, [IP, IP, C.c_int, C.c_int, IP, C.POINTER(IP)]),
        "bn_dev_scan_i32": (C.c_int, [VP, VP, C.c_longlong, VP]),
        "bn_dev_scan_i64": (C.c_int, [VP, VP, C.c_longlong, VP]),
        "bn_dev_sort_pairs": (C.c_int, [VP, VP, VP, VP, C.c_int, C.c_int, VP]),
        "bn_sweep_levels": (C.c_int, [C.POINTER(CSR), IP, C.c_int, IP, C.POINTER(IP)]),
        "cf_gauss_seidel": (C.c_int, [C.c_int, IP, IP, DP, DP, DP, DP, C.c_int, C.c_int, DP]),
        "bn_cf_gauss_seidel_device": (C.c_int, [C.c_int, C.c_int, VP, VP, VP, VP, VP, VP, C.c_int, C.c_int]),
        "bn_set_library_options": (None, [C.POINTER(Options)]),
        "bn_trsv_level": (None, [C.c_char, C.c_int, VP, VP, VP, VP, VP, VP, VP, C.c_int, VP, VP, VP]),
        "bn_gs_color": (None, [C.c_char, C.c_int, C.c_int, VP, VP, VP, VP, VP, VP, C.c_int, VP, VP, VP]),
        "bn_jacobi_rows": (None, [C.c_char, C.c_int, VP, VP, VP, VP, VP, VP, VP, C.c_int, VP, VP, VP]),
        "bn_jacobi_rows_scaled": (None, [C.c_char, C.c_int, VP, VP, VP, VP, VP, VP, VP, VP, C.c_int, VP]),
        "bn_dense_gemv": (None, [C.c_char, C.c_int, VP, VP, VP, VP]),
        "bn_dot_partial": (None, [C.c_int, VP, VP, VP, VP, VP, VP, VP]),
        "bn_scalar_finish": (None, [C.c_int, VP, VP, VP, VP, C.c_int, VP]),
        "bn_scalar_finish_pap": (None, [VP, C.c_int, VP, VP]),
        "bn_csr_apply_blocks_dot": (None, [C.c_int, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
        "bn_edge_grad": (None, [C.c_int, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def ip(a):
    return a.ctypes.data_as(IP)


def dp(a):
    return a.ctypes.data_as(DP)


def last_error():
    e = lib.bn_last_error()
    return e.decode() if e else ""


class DeviceArray:
    """A device buffer owned through libbnamg's HIP runtime (FP64 or int32
    numpy mirror).  Lets ctypes callers (tests, benches) pass device
    pointers without loading a second HIP runtime."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.dtype, self.shape, self.nbytes = host.dtype, host.shape, host.nbytes
        p = C.c_void_p()
        if lib.bn_dev_alloc(C.byref(p), max(self.nbytes, 8)):
            raise RuntimeError(f"device allocation failed: {last_error()}")
        self.ptr = p.value
        self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if lib.bn_dev_copy(C.c_void_p(self.ptr), host.ctypes.data_as(VP), self.nbytes, 1):
            raise RuntimeError("host-to-device copy failed")

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if lib.bn_dev_copy(out.ctypes.data_as(VP), C.c_void_p(self.ptr), self.nbytes, 2):
            raise RuntimeError("device-to-host copy failed")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib.bn_dev_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_available():
    return bool(lib.bn_device_available())
