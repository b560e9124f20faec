"""ctypes bindings of the two native libraries (include/tenet_hip.h, include/tenet_host.h).

The libraries are built in-tree by ``tenet_amd.build``; loading fails loudly
when they are missing -- there is no Python or CPU fallback for the device
path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

MAX_RANK = 8

OK = 0
E_INVALID, E_UNSUPPORTED, E_SHAPE, E_OOB, E_DEVICE_MISMATCH, E_HIP, E_OOM = -1, -2, -3, -4, -5, -6, -7
ERRORS = {-1: "INVALID", -2: "UNSUPPORTED", -3: "SHAPE", -4: "OOB", -5: "DEVICE_MISMATCH",
          -6: "HIP", -7: "OOM"}

PREC_F64, PREC_F32, PREC_SPLIT_BF16, PREC_SPLIT_I8 = 0, 1, 2, 3
OP_IDENTITY, OP_TRANSPOSE, OP_ADJOINT = 0, 1, 2   # MatrixOp order (tenet-dense/src/executor.rs:24-31)
F64, F32, C32, C64 = 0, 1, 2, 3                     # DenseRead / DenseWrite payload types
FAULT_NONE, FAULT_ALLOC, FAULT_LAUNCH = 0, 1, 2
RULES = {"U1": 0, "SU2": 1, "Z2": 2, "U1xZ2": 3}


class TenetError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class GemmJob(C.Structure):
    _fields_ = [("dst_offset", C.c_size_t), ("lhs_offset", C.c_size_t), ("rhs_offset", C.c_size_t),
                ("rows", C.c_size_t), ("contracted", C.c_size_t), ("cols", C.c_size_t)]


class GemmPtrJob(C.Structure):
    _fields_ = [("dst", C.c_void_p), ("lhs", C.c_void_p), ("rhs", C.c_void_p),
                ("rows", C.c_int64), ("contracted", C.c_int64), ("cols", C.c_int64),
                ("ldd", C.c_int64), ("ldl", C.c_int64), ("ldr", C.c_int64),
                ("lhs_op", C.c_int32), ("rhs_op", C.c_int32)]


class Layout(C.Structure):
    _fields_ = [("offset", C.c_int64), ("rank", C.c_int32), ("reserved", C.c_int32),
                ("shape", C.c_int64 * MAX_RANK), ("strides", C.c_int64 * MAX_RANK)]


class TransformBlock(C.Structure):
    _fields_ = [("kind", C.c_int32), ("dst_layout", C.c_uint32), ("dst_count", C.c_uint32),
                ("src_layout", C.c_uint32), ("src_count", C.c_uint32), ("coeff", C.c_uint32)]


class TransformDesc(C.Structure):
    _fields_ = [("dst_len", C.c_int64), ("src_len", C.c_int64),
                ("n_layouts", C.c_size_t), ("layouts", C.POINTER(Layout)),
                ("n_blocks", C.c_size_t), ("blocks", C.POINTER(TransformBlock)),
                ("n_coeffs", C.c_size_t), ("coeffs", C.POINTER(C.c_double)),
                ("n_inactive", C.c_size_t), ("inactive", C.POINTER(C.c_uint32))]


class ScalarOp(C.Structure):
    _fields_ = [("op", C.c_int32), ("out", C.c_int32), ("a", C.c_int32), ("b", C.c_int32), ("imm", C.c_double)]


SCALAR_SET, SCALAR_DIV, SCALAR_COPY, SCALAR_RSQRT, SCALAR_NEG, SCALAR_CG_GUARD = 0, 1, 2, 3, 4, 5


class Region(C.Structure):
    _fields_ = [("offset", C.c_int64), ("len", C.c_int64)]


class ContractDesc(C.Structure):
    _fields_ = [("lhs", C.POINTER(TransformDesc)), ("rhs", C.POINTER(TransformDesc)),
                ("out", C.POINTER(TransformDesc)), ("n_jobs", C.c_size_t),
                ("jobs", C.POINTER(GemmJob)), ("lhs_len", C.c_int64), ("rhs_len", C.c_int64),
                ("core_len", C.c_int64), ("n_uncovered", C.c_size_t),
                ("uncovered", C.POINTER(Region)), ("lhs_op", C.c_int32), ("rhs_op", C.c_int32)]


VP = C.c_void_p
PP = C.POINTER(C.c_void_p)
SZ = C.c_size_t
SZP = C.POINTER(C.c_size_t)

# name -> (restype, argtypes)
HIP_API = {
    "tnt_ctx_create": (C.c_int, [C.c_int, PP]),
    "tnt_ctx_destroy": (C.c_int, [VP]),
    "tnt_last_error": (C.c_char_p, [VP]),
    "tnt_ctx_synchronize": (C.c_int, [VP]),
    "tnt_ctx_stream": (C.c_int, [VP, PP]),
    "tnt_ctx_pool_bytes": (C.c_int, [VP, SZP]),
    "tnt_buf_alloc": (C.c_int, [VP, SZ, C.c_int, PP]),
    "tnt_buf_wrap": (C.c_int, [VP, VP, SZ, C.c_int, PP]),
    "tnt_buf_free": (C.c_int, [VP, VP]),
    "tnt_buf_upload": (C.c_int, [VP, VP, VP, SZ]),
    "tnt_buf_download": (C.c_int, [VP, VP, VP, SZ]),
    "tnt_buf_zero": (C.c_int, [VP, VP]),
    "tnt_buf_fill_random": (C.c_int, [VP, VP, C.c_uint64, C.c_int]),
    "tnt_buf_info": (C.c_int, [VP, PP, SZP, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "tnt_gemm_batch": (C.c_int, [VP, VP, VP, VP, C.POINTER(GemmJob), SZ, SZP, SZ, C.c_int,
                                 C.c_int, C.c_double, C.c_double, C.c_int]),
    "tnt_gemm_batch_complex": (C.c_int, [VP, VP, VP, VP, C.POINTER(GemmJob), SZ, SZP, SZ, C.c_int,
                                         C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
    "tnt_gemm_grouped": (C.c_int, [VP, C.POINTER(GemmPtrJob), SZ, C.c_double, C.c_double, C.c_int]),
    "tnt_plan_upload": (C.c_int, [VP, C.POINTER(TransformDesc), PP]),
    "tnt_plan_destroy": (C.c_int, [VP, VP]),
    "tnt_tree_transform": (C.c_int, [VP, VP, VP, VP, C.c_double, C.c_double, C.c_int]),
    "tnt_contract_upload": (C.c_int, [VP, C.POINTER(ContractDesc), PP]),
    "tnt_contract_destroy": (C.c_int, [VP, VP]),
    "tnt_contract_set_tile_config": (C.c_int, [VP, VP, C.c_int]),
    "tnt_contract": (C.c_int, [VP, VP, VP, VP, VP, C.c_double, C.c_double, C.c_int]),
    "tnt_contract_route": (C.c_int, [VP, C.POINTER(C.c_int)]),
    "tnt_weighted_dot": (C.c_int, [VP, VP, VP, C.POINTER(Region), C.POINTER(C.c_double), SZ,
                                   C.POINTER(C.c_double)]),
    "tnt_axpby": (C.c_int, [VP, VP, VP, C.c_double, C.c_double]),
    "tnt_weighted_dot_async": (C.c_int, [VP, VP, VP, C.POINTER(Region), C.POINTER(C.c_double), SZ, VP, SZ]),
    "tnt_weighted_dot_async_noalloc": (C.c_int, [VP, VP, VP, C.POINTER(Region), C.POINTER(C.c_double), SZ, VP, SZ]),
    "tnt_axpby_dev": (C.c_int, [VP, VP, VP, VP, C.c_int64, C.c_double, C.c_int64, C.c_double, C.c_int64]),
    "tnt_scalars": (C.c_int, [VP, VP, C.POINTER(ScalarOp), SZ]),
    "tnt_event_create": (C.c_int, [VP, PP]),
    "tnt_event_record": (C.c_int, [VP, VP]),
    "tnt_event_query": (C.c_int, [VP, VP]),
    "tnt_event_destroy": (C.c_int, [VP, VP]),
    "tnt_ctx_set_timing": (C.c_int, [VP, C.c_int]),
    "tnt_ctx_set_shadow": (C.c_int, [VP, C.c_int]),
    "tnt_ctx_shadow_stats": (C.c_int, [VP, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64)]),
    "tnt_ctx_kernel_stats": (C.c_int, [VP, C.c_int, SZP, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tnt_ctx_reset_stats": (C.c_int, [VP]),
    "tnt_set_tile_config": (C.c_int, [VP, C.c_int]),
    "tnt_set_split_slices": (C.c_int, [VP, C.c_int]),
    "tnt_set_tile_table": (C.c_int, [VP, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int32), SZ]),
    "tnt_ctx_inject_fault": (C.c_int, [VP, C.c_int, C.c_int]),
    "tnt_ctx_set_table_cache_cap": (C.c_int, [VP, SZ]),
    "tnt_graph_begin": (C.c_int, [VP]),
    "tnt_graph_end": (C.c_int, [VP, PP]),
    "tnt_graph_launch": (C.c_int, [VP, VP]),
    "tnt_graph_destroy": (C.c_int, [VP, VP]),
}

U64P = C.POINTER(C.c_uint64)
I64P = C.POINTER(C.c_int64)
U8P = C.POINTER(C.c_uint8)
HOST_API = {
    "tnth_last_error": (C.c_char_p, []),
    "tnth_u1_id": (C.c_uint64, [C.c_int64]),
    "tnth_u1_charge": (C.c_int64, [C.c_uint64]),
    "tnth_product_id": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "tnth_product_split": (C.c_int, [C.c_uint64, U64P, U64P]),
    "tnth_symbol_f": (C.c_double, [C.c_int] + [C.c_uint64] * 6),
    "tnth_symbol_r": (C.c_double, [C.c_int] + [C.c_uint64] * 3),
    "tnth_symbol_dim": (C.c_double, [C.c_int, C.c_uint64]),
    "tnth_symbol_dual": (C.c_uint64, [C.c_int, C.c_uint64]),
    "tnth_space_create": (C.c_int, [C.c_int, SZ, SZ, SZP, U64P, I64P, U8P, PP]),
    "tnth_space_destroy": (None, [VP]),
    "tnth_space_rule": (C.c_int, [VP]),
    "tnth_space_nout": (SZ, [VP]),
    "tnth_space_nin": (SZ, [VP]),
    "tnth_space_len": (C.c_int64, [VP]),
    "tnth_space_n_blocks": (SZ, [VP]),
    "tnth_space_block": (C.c_int, [VP, SZ, I64P, C.POINTER(C.c_int32), I64P, I64P, U64P]),
    "tnth_space_block_trees": (C.c_int, [VP, SZ, U64P, U64P, U64P]),
    "tnth_space_n_sectors": (SZ, [VP]),
    "tnth_space_sector": (C.c_int, [VP, SZ, U64P, I64P, I64P, I64P, C.POINTER(C.c_double)]),
    "tnth_space_leg": (C.c_int, [VP, SZ, SZP, U64P, I64P, U8P]),
    "tnth_space_permuted": (C.c_int, [VP, SZP, SZ, SZP, SZ, PP]),
    "tnth_space_adjoint": (C.c_int, [VP, PP]),
    "tnth_space_equal": (C.c_int, [VP, VP]),
    "tnth_transform_compile": (C.c_int, [VP, SZP, SZ, SZP, SZ, PP]),
    "tnth_transform_destroy": (None, [VP]),
    "tnth_transform_desc": (C.POINTER(TransformDesc), [VP]),
    "tnth_transform_dst_space": (VP, [VP]),
    "tnth_transform_counts": (C.c_int, [VP, SZP, SZP]),
    "tnth_contract_compile": (C.c_int, [VP, VP, SZP, SZP, SZ, SZP, SZ, C.c_int, C.c_int, PP]),
    "tnth_contract_destroy": (None, [VP]),
    "tnth_contract_desc": (C.POINTER(ContractDesc), [VP]),
    "tnth_contract_dst_space": (VP, [VP]),
    "tnth_contract_flops": (C.c_double, [VP]),
    "tnth_narrow_fusion": (C.c_int, [C.POINTER(ContractDesc), PP]),
    "tnth_contract_batch": (C.c_int, [C.POINTER(ContractDesc), SZ, C.c_int, PP]),
    "tnth_batch_desc": (C.POINTER(ContractDesc), [VP]),
    "tnth_batch_destroy": (None, [VP]),
    "tnth_fusion_destroy": (None, [VP]),
    "tnth_fusion_desc": (C.POINTER(TransformDesc), [VP]),
    "tnth_fusion_gather": (SZ, [VP, C.POINTER(I64P), C.POINTER(I64P), C.POINTER(C.POINTER(C.c_double))]),
    "tnth_bytes_data": (C.POINTER(C.c_uint8), [VP]),
    "tnth_bytes_size": (C.c_size_t, [VP]),
    "tnth_bytes_free": (None, [VP]),
    "tnth_snapshot_tensor": (C.c_int, [VP, C.POINTER(C.c_double), C.c_int64, C.c_int, PP]),
    "tnth_snapshot_tensor_space": (C.c_int, [C.c_char_p, SZ, PP, C.POINTER(C.c_int)]),
    "tnth_snapshot_tensor_payload": (C.c_int, [C.c_char_p, SZ, VP, C.POINTER(C.c_double), C.c_int64]),
    "tnth_snapshot_leg": (C.c_int, [C.c_int, SZ, C.POINTER(C.c_uint64), C.POINTER(C.c_int64), C.c_int, PP]),
    "tnth_snapshot_leg_decode": (C.c_int, [C.c_char_p, SZ, C.POINTER(C.c_int), SZP, C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
}

_host = None
_hip = None


def _load(name, api):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build the native libraries first "
            "(python -c 'import __graft_entry__ as g; g.build()' or python -m tenet_amd.build)")
    lib = C.CDLL(path)
    for fn, (res, args) in api.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    return lib


def host():
    """The host structure library; TENET_AMD_HOST_LIB names another build of
    it in lib/ (the sanitizer build libtenet_host_san.so)."""
    global _host
    if _host is None:
        _host = _load(os.environ.get("TENET_AMD_HOST_LIB", "libtenet_host.so"), HOST_API)
    return _host


def hip():
    global _hip
    if _hip is None:
        _hip = _load("libtenet_hip.so", HIP_API)
    return _hip


def check_host(rc):
    if rc != OK:
        raise TenetError(rc, host().tnth_last_error().decode())


def sz_array(xs):
    return (C.c_size_t * max(1, len(xs)))(*xs)
