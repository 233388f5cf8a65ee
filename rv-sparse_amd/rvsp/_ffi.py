"""ctypes binding of librvsparse_hip.so (the rv-sparse C API on gfx950).

The structs mirror include/rv_sparse_types.h and include/rv_sparse_hip.h
field for field. Loading fails loudly when the library has not been built:
there is no other implementation behind this module.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "librvsparse_hip.so")

# rvsp_status_t (reference include/rv_sparse_types.h:21-30, plus DEVICE)
SUCCESS = 0
ERROR_NULL_POINTER = -1
ERROR_INVALID_ARGUMENT = -2
ERROR_INVALID_CSR = -3
ERROR_UNSUPPORTED_DTYPE = -4
ERROR_UNSUPPORTED_BACKEND = -5
ERROR_ALLOCATION_FAILED = -6
ERROR_DEVICE = -7

# rvsp_dtype_t
DTYPE_INT8, DTYPE_INT32, DTYPE_BF16, DTYPE_FP32, DTYPE_FP64 = range(5)
FORMAT_CSR = 0
# rvsp_backend_t
BACKEND_SCALAR, BACKEND_SCALAR_UNROLL4, BACKEND_RVV_INTRINSICS, BACKEND_HIP = range(4)
# rvsp_spgemm_algo_t
(ALGO_DEFAULT, ALGO_RVV, ALGO_CONTIG, ALGO_ADAPTIVE,
 ALGO_HIP_FIXED, ALGO_HIP_TUNED, ALGO_HIP_LEARNED) = range(7)
# rvsp_memory_t
MEMORY_HOST, MEMORY_DEVICE = 0, 1
# rvsp_phase_t / rvsp_kind_t
PHASE_COUNT, PHASE_FILL, PHASE_NUMERIC = 0, 1, 2
(KIND_NONE, KIND_DIRECT, KIND_HASH, KIND_SEARCH, KIND_BITMAP, KIND_BLOCK,
 KIND_GLOBAL, KIND_MAP) = range(8)
KIND_NAMES = {KIND_NONE: "none", KIND_DIRECT: "direct", KIND_HASH: "lds_hash",
              KIND_SEARCH: "lds_search", KIND_BITMAP: "lds_bitmap",
              KIND_BLOCK: "lds_block", KIND_GLOBAL: "global_dense", KIND_MAP: "lds_map"}
# rvsp_csr_status_t
CSR_OK, CSR_BAD_ROW_PTR, CSR_COL_OUT_OF_RANGE, CSR_UNSORTED, CSR_DUPLICATE = range(5)


class CsrMatrix(C.Structure):
    """rvsp_csr_matrix_t (reference include/rv_sparse_types.h:53-67)."""
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("nnz", C.c_int32),
                ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p),
                ("values", C.c_void_p), ("dtype", C.c_int), ("format", C.c_int),
                ("owns_data", C.c_int)]


class SpgemmOptions(C.Structure):
    """rvsp_spgemm_options_t (reference include/rv_sparse_types.h:69-74)."""
    _fields_ = [("backend", C.c_int), ("input_dtype", C.c_int),
                ("output_dtype", C.c_int)]


class BinInfo(C.Structure):
    _fields_ = [("phase", C.c_int32), ("bin", C.c_int32), ("kind", C.c_int32),
                ("group", C.c_int32), ("cap", C.c_int32), ("rows", C.c_int32),
                ("lds_bytes", C.c_int32), ("bound", C.c_int32)]


class BinChoice(C.Structure):
    _fields_ = [("phase", C.c_int32), ("bin", C.c_int32), ("kind", C.c_int32),
                ("group", C.c_int32), ("cap", C.c_int32), ("variant", C.c_int32)]


# include/rv_sparse_learned.h
MLP_FEATURES, MLP_HIDDEN, MLP_CLASSES = 17, 64, 56
MLP_PARAMS = (MLP_HIDDEN * MLP_FEATURES + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN
              + MLP_CLASSES * MLP_HIDDEN + MLP_CLASSES)


class LearnedClass(C.Structure):
    _fields_ = [("kind", C.c_int32), ("group", C.c_int32), ("cap_mult", C.c_int32),
                ("variant", C.c_int32), ("cap", C.c_int32)]


class LearnedSample(C.Structure):
    _fields_ = [("phase", C.c_int32), ("bin", C.c_int32), ("fixed_class", C.c_int32),
                ("n_candidates", C.c_int32), ("features", C.c_float * MLP_FEATURES),
                ("us", C.c_float * MLP_CLASSES)]


P = C.c_void_p
PM = C.POINTER(CsrMatrix)
I32, I64, SZ = C.c_int32, C.c_int64, C.c_size_t
FP = C.POINTER(C.c_float)
U64P = C.POINTER(C.c_uint64)
I32P = C.POINTER(C.c_int32)

# name -> (restype, argtypes). Exactly the functions the headers declare.
PROTOTYPES = {
    # include/rv_sparse.h
    "rvsp_get_version": (C.c_char_p, []),
    "rvsp_status_to_string": (C.c_char_p, [C.c_int]),
    "rvsp_csr_create": (C.c_int, [PM, I32, I32, I32, P, P, P, C.c_int]),
    "rvsp_csr_validate": (C.c_int, [PM]),
    "rvsp_csr_destroy": (None, [PM]),
    "rvsp_spgemm_csr": (C.c_int, [PM, PM, PM, C.POINTER(SpgemmOptions)]),
    "rvsp_spgemm_descr_create": (C.c_int, [C.POINTER(P)]),
    "rvsp_spgemm_descr_destroy": (None, [P]),
    "rvsp_spgemm_set_algo": (C.c_int, [P, C.c_int]),
    "rvsp_spgemm_work_estimation": (C.c_int, [P, PM, PM, C.POINTER(SZ), C.POINTER(I32)]),
    "rvsp_spgemm_compute": (C.c_int, [P, PM, PM, PM, P]),
    "rvsp_spgemm_get_op_counts": (C.c_int, [P, P, I32]),
    "rvsp_spgemm_invalidate_structure": (None, [P]),
    # include/rvsp_spgemm.h
    "rvsp_spgemm_buffer_size": (C.c_int, [I32, C.POINTER(SZ)]),
    "rvsp_csr_check": (C.c_int, [I32, I32, P, P, C.POINTER(I32)]),
    "rvsp_csr_status_string": (C.c_char_p, [C.c_int]),
    "rvsp_spgemm_hip_f32": (C.c_int, [I32, I32, I32, P, P, P, P, P, P, P, SZ,
                                      C.POINTER(P), C.POINTER(P), C.POINTER(P),
                                      C.POINTER(I32), P]),
    "rvsp_spgemm_scalar_f32": (C.c_int, [I32, I32, I32, P, P, P, P, P, P, P, SZ,
                                      C.POINTER(P), C.POINTER(P), C.POINTER(P),
                                      C.POINTER(I32), P]),
    # include/rv_sparse_hip.h
    "rvsp_spgemm_set_memory": (C.c_int, [P, C.c_int]),
    "rvsp_spgemm_set_stream": (C.c_int, [P, P]),
    "rvsp_spgemm_get_stream": (C.c_int, [P, C.POINTER(P)]),
    "rvsp_spgemm_set_async": (C.c_int, [P, C.c_int]),
    "rvsp_spgemm_set_position_map": (C.c_int, [P, C.c_int]),
    "rvsp_spgemm_get_plan": (C.c_int, [P, C.POINTER(BinInfo), I32, C.POINTER(I32)]),
    "rvsp_spgemm_set_bin_choices": (C.c_int, [P, C.POINTER(BinChoice), I32]),
    "rvsp_spgemm_autotune": (C.c_int, [P, PM, PM, PM, P, I32]),
    "rvsp_tuning_table_save": (I32, [C.c_char_p]),
    "rvsp_tuning_table_load": (I32, [C.c_char_p]),
    "rvsp_tuning_table_clear": (None, []),
    "rvsp_tuning_table_size": (I32, []),
    "rvsp_spgemm_batch_work_estimation": (C.c_int, [P, I32, PM, PM, I32P]),
    "rvsp_spgemm_batch_compute": (C.c_int, [P, I32, PM, PM, PM]),
    "rvsp_spgemm_batch_groups": (I32, [P]),
    "rvsp_spgemm_batch_graph_launches": (I32, [P]),
    "rvsp_spgemm_batch_set_graph": (C.c_int, [P, C.c_int]),
    "rvsp_spgemm_batch_set_group_limit": (C.c_int, [P, I64]),
    "rvsp_spgemm_batch_set_fuse": (C.c_int, [P, C.c_int]),
    "rvsp_spgemm_batch_set_persistent": (C.c_int, [P, C.c_int]),
    "rvsp_spgemm_batch_stream_segments": (I32, [P]),
    "rvsp_spgemm_batch_numeric_path": (I32, [P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "rvsp_hip_malloc": (C.c_int, [C.POINTER(P), SZ]),
    "rvsp_hip_free": (C.c_int, [P]),
    "rvsp_hip_memcpy": (C.c_int, [P, P, SZ, C.c_int]),
    "rvsp_hip_synchronize": (C.c_int, []),
    "rvsp_hip_pool_release": (C.c_int, []),
    # include/rv_sparse_mtx.h
    "rvsp_csr_read_mtx": (C.c_int, [C.c_char_p, I32, PM]),
    "rvsp_csr_canonicalize": (C.c_int, [PM]),
    "rvsp_hip_pool_cached_bytes": (SZ, []),
    "rvsp_hip_device_count": (I32, []),
    "rvsp_csr_check_device": (C.c_int, [I32, I32, I32, P, P, I32P, I32P]),
    "rvsp_csr_check_device_noalloc": (C.c_int, [I32, I32, I32, P, P, I32P, I32P]),
    # include/rv_sparse_learned.h
    "rvsp_learned_class_info": (C.c_int, [I32, C.POINTER(LearnedClass)]),
    "rvsp_learned_sample_count": (I32, []),
    "rvsp_learned_sample_get": (C.c_int, [I32, C.POINTER(LearnedSample)]),
    "rvsp_learned_sample_add": (C.c_int, [C.POINTER(LearnedSample)]),
    "rvsp_learned_sample_clear": (None, []),
    "rvsp_learned_fit": (C.c_int, [I32, I32, I32, C.c_float, C.c_uint64, FP]),
    "rvsp_learned_model_save": (I32, [C.c_char_p]),
    "rvsp_learned_model_load": (I32, [C.c_char_p]),
    "rvsp_learned_model_clear": (None, []),
    "rvsp_learned_model_ready": (I32, []),
    "rvsp_learned_predict": (C.c_int, [FP, U64P, I32, I32P]),
    "rvsp_mlp_forward": (C.c_int, [FP, FP, U64P, I32, FP, I32P]),
    "rvsp_mlp_gradient": (C.c_int, [FP, FP, FP, U64P, FP, I32, FP, FP]),
    "rvsp_mlp_train": (C.c_int, [FP, FP, FP, U64P, FP, I32, I32, C.c_float, FP]),
}

_lib = None


def lib():
    """The loaded library with prototypes applied. Raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `make -C rv-sparse_amd` or "
                "__graft_entry__.build() first")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class RvspError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = lib().rvsp_status_to_string(status).decode()
        super().__init__(f"{what}: {msg} ({status})" if what else f"{msg} ({status})")


def check(status, what=""):
    if status != SUCCESS:
        raise RvspError(status, what)
    return status
