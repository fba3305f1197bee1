"""ctypes binding of the C ABI declared in include/meshfemsparse_amd.h (libmfs_amd.so, in-tree)."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmfs_amd.so")

c_i32, c_i64, c_dbl, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
c_double = c_dbl
P_i32, P_i64, P_dbl, P_u8 = (ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint8))

MFS_OK, MFS_ERR_ARG, MFS_ERR_STATE, MFS_ERR_NOT_SPD, MFS_ERR_DEVICE, MFS_ERR_ALLOC, MFS_ERR_UNSUPPORTED = \
    0, -1, -2, -3, -4, -5, -6
MFS_ERR_NOT_CONVERGED = -7
MFS_PREC_FP64, MFS_PREC_FP32, MFS_PREC_BF16, MFS_PREC_BF16X3 = 0, 1, 2, 3
MFS_HE_FULL12, MFS_HE_UPPER78 = 0, 1
MFS_SOLVE_NOALLOC_MAX_BYTES = 96 << 20

# name -> argtypes (all return int status unless listed in _RESTYPE)
_SIGS = {
    "mfs_last_error": [],
    "mfs_version": [],
    "mfs_pattern_build": [P_i32, c_i64, c_i32, c_i64, ctypes.POINTER(c_vp)],
    "mfs_pattern_get": [c_vp, P_i64, P_i64, ctypes.POINTER(P_i64), ctypes.POINTER(P_i64)],
    "mfs_pattern_build_device": [c_vp, c_i64, c_i32, c_i64, c_vp, ctypes.POINTER(c_vp)],
    "mfs_pattern_get_device": [c_vp, ctypes.POINTER(P_i64), ctypes.POINTER(P_i64)],
    "mfs_pattern_destroy": [c_vp],
    "mfs_assembler_create": [P_i64, P_i64, c_i64, P_i32, c_i64, c_i32, ctypes.POINTER(c_vp)],
    "mfs_assemble_hessian": [c_vp, c_vp, ctypes.c_int, c_vp, c_vp],
    "mfs_assemble_hessian_noncontiguous": [c_vp, c_vp, ctypes.c_int, c_vp, c_vp],
    "mfs_assembler_zero": [c_vp, c_vp, c_vp],
    "mfs_assembler_destroy": [c_vp],
    "mfs_neohookean_elements": [c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_dbl, c_vp, c_i32, c_vp, c_vp],
    "mfs_assemble_gradient": [c_vp, c_vp, c_vp, c_vp],
    "mfs_factorizer_create": [ctypes.c_int, ctypes.POINTER(c_vp)],
    "mfs_factorizer_destroy": [c_vp],
    "mfs_factorize_symbolic": [c_vp, P_i64, P_i64, c_i64, c_i32, P_i64, c_i64],
    "mfs_factorize_numeric": [c_vp, c_vp, c_dbl, c_vp, c_vp],
    "mfs_factorize_symbolic_scalar": [c_vp, P_i64, P_i64, c_i64, P_i64, c_i64],
    "mfs_factorize_numeric_scalar": [c_vp, c_vp, c_dbl, c_vp, c_vp],
    "mfs_solve": [c_vp, c_vp, c_vp, c_i32, c_vp],
    "mfs_solve_noalloc": [c_vp, c_vp, c_vp, c_i32, c_vp],
    "mfs_solve_raw_reduced": [c_vp, c_vp, c_vp, c_i32, c_i32, c_vp],
    "mfs_remove_fixed_entries": [c_vp, c_vp, c_vp, c_i32, c_i32, c_vp],
    "mfs_extract_full_solution": [c_vp, c_vp, c_vp, c_i32, c_i32, c_vp],
    "mfs_spmv": [c_vp, c_vp, c_vp, c_vp, c_vp],
    "mfs_solve_refined": [c_vp, c_vp, c_vp, c_vp, c_dbl, c_i32, P_i32, P_dbl, c_vp],
    "mfs_solve_refined_segments": [c_vp, c_vp, c_vp, c_vp, c_i64, P_i64, c_dbl, c_i32, P_i32, P_dbl, P_dbl,
                                   c_vp],
    "mfs_factorizer_set_graph": [c_vp, c_i32],
    "mfs_autotune": [c_vp, P_i64, P_i64, c_i64, c_i32, P_i64, c_i64, c_vp, P_i32, c_i32, P_i32, c_i32, c_i32,
                     P_dbl, P_i32, c_vp],
    "mfs_schedule_stats": [c_vp, P_i64],
    "mfs_learned_create": [P_i64, P_i64, c_i64, P_i64, c_i64, ctypes.c_uint64, ctypes.POINTER(c_vp)],
    "mfs_learned_destroy": [c_vp],
    "mfs_learned_info": [c_vp, P_i64],
    "mfs_learned_param_offsets": [c_vp, P_i64],
    "mfs_learned_graph": [c_vp, P_i32, P_i32, P_i32, P_i32, P_i32, P_i32],
    "mfs_learned_analyze_host": [P_i64, P_i64, c_i64, P_i64, c_i64, P_i64, P_i32, P_i32, P_i32, P_i32, P_i32, P_i32],
    "mfs_learned_set_matrix": [c_vp, c_vp, c_vp],
    "mfs_learned_features": [c_vp, c_vp, c_vp, c_vp, c_vp],
    "mfs_learned_loss": [c_vp, c_vp, c_i32, c_i32, P_dbl, c_vp],
    "mfs_learned_adam_step": [c_vp, ctypes.c_float, c_vp],
    "mfs_learned_params": [c_vp, c_vp, c_vp],
    "mfs_learned_grads": [c_vp, c_vp],
    "mfs_learned_factor": [c_vp, c_vp, c_vp, c_vp],
    "mfs_learned_solve": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_dbl, c_i32, P_i32, P_i32, P_dbl, P_i32, c_vp],
    "mfs_factorization_type": [c_vp, P_i32],
    "mfs_num_successful_pivots": [c_vp, P_i64],
    "mfs_sizes": [c_vp, P_i64, P_i64],
    "mfs_factor_nnz": [c_vp, P_i64],
    "mfs_flop_estimate": [c_vp, P_dbl],
    "mfs_flop_breakdown": [c_vp, P_dbl],
    "mfs_get_inverse_permutation": [c_vp, P_i64],
    "mfs_clear_factors": [c_vp],
    "mfs_stash_factorization": [c_vp],
    "mfs_has_stashed_factorization": [c_vp, P_i32],
    "mfs_swap_stashed_factorization": [c_vp],
    "mfs_clear_stashed_factorization": [c_vp],
    "mfs_supernode_stats": [c_vp, P_i64, P_i64, P_i64, P_i64],
    "mfs_factor_export_csc": [c_vp, P_i64, P_i64, P_dbl],
    "mfs_analyze_host": [P_i64, P_i64, c_i64, P_i64, c_i64, P_i64, P_i64, P_dbl],
    "mfs_nd_ordering": [c_vp, P_i64, P_i64, P_i64, P_i64, P_i64],
    "mfs_set_temporal_reuse_period": [c_vp, c_i64],
    "mfs_get_temporal_reuse_period": [c_vp, P_i64],
    "mfs_reset_temporal_reuse": [c_vp],
    "mfs_temporal_reuse_statistics": [c_vp, P_i64, P_dbl],
    "mfs_nd_create": [ctypes.POINTER(c_vp)],
    "mfs_nd_destroy": [c_vp],
    "mfs_nd_set_temporal_reuse_period": [c_vp, c_i64],
    "mfs_nd_reset": [c_vp],
    "mfs_nd_order": [c_vp, c_i64, P_i64, P_i64, P_i64, P_i64],
    "mfs_nd_forest": [c_vp, P_i64, P_i64],
    "mfs_nd_statistics": [c_vp, P_i64, P_dbl],
    "mfs_debug_emulate_factor": [P_i64, P_i64, c_i64, P_i64, c_i64, P_dbl, c_dbl, P_i64, P_i64, P_dbl, P_i64,
                                 P_dbl],
    "mfs_last_timings": [c_vp, P_dbl],
    "mfs_kernel_profiling": [c_vp, c_i32, c_i32],
    "mfs_kernel_profile": [c_vp, P_dbl, P_i64],
    "mfs_kernel_profile_work": [c_vp, P_dbl],
    "mfs_kernel_profile_bytes": [c_vp, P_dbl],
    "mfs_blockcsc_write": [ctypes.c_char_p, P_i64, P_i64, c_i64, P_dbl],
    "mfs_blockcsc_read": [ctypes.c_char_p, P_i64, P_i64, P_i64, P_i64, P_i64, P_dbl],
    "mfs_cube_mesh": [c_i32, P_i32, P_dbl],
    "mfs_perturb_nodes": [c_i32, ctypes.c_uint64, c_dbl, P_dbl],
    "mfs_neohookean_host": [P_dbl, P_dbl, P_i32, c_i64, P_dbl, P_dbl, c_dbl, c_dbl, P_dbl, P_dbl,
                            P_dbl, c_i32],
    "mfs_test_gemm_nt": [c_i32, c_i32, c_i32, c_dbl, c_vp, c_i32, c_vp, c_i32, c_dbl, c_vp, c_i32,
                         c_vp],
    "mfs_test_gemm_nt_f32": [c_i32, c_i32, c_i32, ctypes.c_float, c_vp, c_i32, c_vp, c_i32, ctypes.c_float,
                             c_vp, c_i32, c_vp],
    "mfs_test_gemm_nt_bf16": [c_i32, c_i32, c_i32, ctypes.c_float, c_vp, c_i32, c_vp, c_i32, ctypes.c_float,
                              c_vp, c_i32, c_i32, c_vp],
    "mfs_test_multiply_lower_block": [c_i32, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64, c_i32, c_vp,
                                      c_i64, c_vp],
}
_RESTYPE = {"mfs_last_error": ctypes.c_char_p}

EXPORTED = tuple(_SIGS)

_lib = None


class MFSError(RuntimeError):
    """Failure reported by the C ABI (status code + message)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class NotPositiveDefiniteError(MFSError):
    pass


class UnsupportedError(MFSError):
    pass


class NotConvergedError(MFSError):
    pass


class InvalidArgumentError(MFSError, ValueError):
    pass


def load(path=None):
    """Load (once) the in-tree shared library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"{p} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(p)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, ctypes.c_int)
    _lib = lib
    return lib


def check(status):
    if status == MFS_OK:
        return
    msg = load().mfs_last_error().decode(errors="replace")
    if status == MFS_ERR_NOT_SPD:
        raise NotPositiveDefiniteError(status, msg)
    if status == MFS_ERR_UNSUPPORTED:
        raise UnsupportedError(status, msg)
    if status == MFS_ERR_ARG:
        raise InvalidArgumentError(status, msg)
    if status == MFS_ERR_NOT_CONVERGED:
        raise NotConvergedError(status, msg)
    raise MFSError(status, msg)


def call(name, *args):
    check(getattr(load(), name)(*args))


def ptr(a, ctype):
    """Pointer to a numpy array's data."""
    return a.ctypes.data_as(ctypes.POINTER(ctype))
