"""ctypes binding of libdmm_gapbs.so (the C ABI declared in include/dmm_gapbs.h).

The shared library is built in place by ``__graft_entry__.build()`` (or
``make -C dist_mem_gapbs_amd/csrc``).  Loading fails loudly when it is
missing: there is no CPU fallback for the device entry points.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdmm_gapbs.so")

DMM_OK = 0
DMM_ERR_INVALID = -1
DMM_ERR_NOMEM = -2
DMM_ERR_HIP = -3
DMM_ERR_STATE = -4
DMM_ERR_CAPACITY = -5
DMM_ERR_NODEVICE = -6
DMM_ERR_GRAPH = -7

DMM_MAX_STEPS = 256
DMM_MAX_PR_ITERS = 1024

DMM_STEP_TD = 0
DMM_STEP_BU = 1
DMM_STEP_QB = 2
DMM_STEP_BQ = 3

DMM_LAYOUT_NATURAL = 0
DMM_LAYOUT_DEGREE = 1

DMM_PR_FP32_SERIAL = 0
DMM_PR_FP32_KAHAN = 1
DMM_PR_FP64 = 2
DMM_BATCH_DEPTH_HBM = 1


class BfsStats(C.Structure):
    _fields_ = [
        ("num_steps", C.c_int32),
        ("num_levels", C.c_int32),
        ("step_kind", C.c_int32 * DMM_MAX_STEPS),
        ("step_in", C.c_int64 * DMM_MAX_STEPS),
        ("step_out", C.c_int64 * DMM_MAX_STEPS),
        ("step_edges", C.c_int64 * DMM_MAX_STEPS),
        ("step_scout", C.c_int64 * DMM_MAX_STEPS),
        ("step_ms", C.c_float * DMM_MAX_STEPS),
        ("step_kernel_ms", C.c_float * DMM_MAX_STEPS),
        ("alg_bytes", C.c_int64),
        ("bu_alg_bytes", C.c_int64),
        ("reached", C.c_int64),
        ("reached_edges", C.c_int64),
        ("total_ms", C.c_double),
        ("bu_ms", C.c_double),
        ("host_ms", C.c_double),
    ]


class BatchParams(C.Structure):
    _fields_ = [
        ("td_div", C.c_int32),
        ("grid_per_cu", C.c_int32),
        ("graph_levels", C.c_int32),
        ("flags", C.c_int32),
    ]


class BatchStats(C.Structure):
    _fields_ = [
        ("num_levels", C.c_int32),
        ("level_kind", C.c_int32 * DMM_MAX_STEPS),
        ("level_active", C.c_int64 * DMM_MAX_STEPS),
        ("total_ms", C.c_double),
        ("sweep_ms", C.c_double),
        ("sweep_alg_bytes", C.c_int64),
        ("sweeps", C.c_int32),
        ("reached_edges", C.c_int64),
        ("host_ms", C.c_double),
    ]


class PrStats(C.Structure):
    _fields_ = [
        ("iters", C.c_int32),
        ("err", C.c_double * DMM_MAX_PR_ITERS),
        ("total_ms", C.c_double),
        ("iter_ms", C.c_double),
        ("alg_bytes_per_iter", C.c_int64),
        ("host_ms", C.c_double),
    ]


class BlockLevelStats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("edges_examined", C.c_int64),
        ("alg_bytes", C.c_int64),
        ("kind", C.c_int32),
        ("pad", C.c_int32),
        ("frontier", C.c_int64),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("alpha", C.c_int32),
        ("beta", C.c_int32),
        ("grid_per_cu", C.c_int32),
        ("td_edges_per_block", C.c_int32),
        ("batch", BatchParams),
        ("plain_ms", C.c_double),
        ("batch_ms", C.c_double),
        ("pr_window", C.c_int32),
        ("pad", C.c_int32),
        ("pr_ms", C.c_double),
    ]


_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_PI32 = C.POINTER(C.c_int32)
_PI64 = C.POINTER(C.c_int64)
_PU64 = C.POINTER(C.c_uint64)
_PU8 = C.POINTER(C.c_uint8)
_PF32 = C.POINTER(C.c_float)
_PD = C.POINTER(C.c_double)

# name -> (restype, argtypes); mirrors include/dmm_gapbs.h one to one.
SIGNATURES = {
    "dmm_last_error": (C.c_char_p, [_P]),
    "dmm_version": (C.c_char_p, []),
    "dmm_generate_edgelist": (_I32, [_I32, _I32, _I32, _I32, _PI32]),
    "dmm_build_graph": (_I32, [_I32, _I32, _I32, _I32, C.POINTER(_P)]),
    "dmm_build_graph_from_edgelist": (_I32, [_I64, _PI32, _I64, _I32, _I32, C.POINTER(_P)]),
    "dmm_host_graph_num_nodes": (_I64, [_P]),
    "dmm_host_graph_num_edges_directed": (_I64, [_P]),
    "dmm_host_graph_directed": (_I32, [_P]),
    "dmm_host_graph_out_offsets": (_PI64, [_P]),
    "dmm_host_graph_out_neighs": (_PI32, [_P]),
    "dmm_host_graph_in_offsets": (_PI64, [_P]),
    "dmm_host_graph_in_neighs": (_PI32, [_P]),
    "dmm_host_graph_free": (None, [_P]),
    "dmm_write_serialized_graph": (_I32, [_P, C.c_char_p]),
    "dmm_read_serialized_graph": (_I32, [C.c_char_p, C.POINTER(_P)]),
    "dmm_pick_sources": (_I32, [_I64, _PI64, _I32, _I32, _PI32]),
    "dmm_rng_sequence": (_I32, [_I32, _I64, _I64, _I64, _PI64]),
    "dmm_partition": (_I32, [_I64, _I32, _I32, _PI64, _PI64, _PI64, _PI64]),
    "dmm_device_count": (_I32, [_PI32]),
    "dmm_ctx_create": (_I32, [_I32, C.POINTER(_P)]),
    "dmm_ctx_destroy": (None, [_P]),
    "dmm_ctx_synchronize": (_I32, [_P]),
    "dmm_build_graph_device": (_I32, [_P, _I64, _PI32, _I64, _I32, C.POINTER(_P)]),
    "dmm_bc_forward": (_I32, [_P, _I32, _PI32, _PU64, _PU64, _PI32]),
    "dmm_bc_read": (_I32, [_P, _PI32, _PU64, _PU64]),
    "dmm_graph_upload": (_I32, [_P, _I64, _I64, _I64, _PI64, _PI32, _PI64, _PI32, _I64]),
    "dmm_ctx_device_bytes": (_I32, [_P, _PI64]),
    "dmm_ctx_set_layout": (_I32, [_P, _I32]),
    "dmm_bfs": (_I32, [_P, _I32, _I32, _I32, _PI32, _PI32, _I32, C.POINTER(BfsStats)]),
    "dmm_bfs_read": (_I32, [_P, _PI32, _PI32]),
    "dmm_step_init": (_I32, [_P, _I32]),
    "dmm_step_td": (_I32, [_P, _PI64, _PI64]),
    "dmm_step_queue_to_bitmap": (_I32, [_P, _PI64]),
    "dmm_step_bu": (_I32, [_P, _PI64]),
    "dmm_step_bitmap_to_queue": (_I32, [_P, _PI64]),
    "dmm_step_finish": (_I32, [_P]),
    "dmm_state_get_parent": (_I32, [_P, _PI32]),
    "dmm_state_set_parent": (_I32, [_P, _PI32]),
    "dmm_state_get_queue": (_I32, [_P, _PI32, _I64, _PI64]),
    "dmm_state_set_queue": (_I32, [_P, _PI32, _I64]),
    "dmm_state_get_bitmap": (_I32, [_P, _I32, _PU64]),
    "dmm_state_set_bitmap": (_I32, [_P, _I32, _PU64]),
    "dmm_state_set_level": (_I32, [_P, _I32]),
    "dmm_generate_edgelist_device": (_I32, [_P, _I32, _I32, _I32, _PI32]),
    "dmm_make_graph_device": (_I32, [_P, _I32, _I32, _I32, C.POINTER(_P)]),
    "dmm_batch_read_depth": (_I32, [_P, _PU8]),
    "dmm_bfs_batch64": (_I32, [_P, _PI32, _I32, _PU8, C.POINTER(BatchParams), C.POINTER(BatchStats)]),
    "dmm_batch_read_level": (_I32, [_P, _I32, _PU64]),
    "dmm_batch_block_level": (_I32, [_P, _PU64, _PU64, _PU64, C.POINTER(BlockLevelStats)]),
    "dmm_pr_pull": (_I32, [_P, _I32, C.c_double, _I32, _P, C.POINTER(PrStats)]),
    "dmm_pr_block_iteration": (_I32, [_P, _PF32, _PF32, _I32, _PD, _PD]),
    "dmm_pr_block_iteration_noalloc": (_I32, [_P, _P, _P, _P, _I32, _P]),
    "dmm_pr_tiles_check": (_I32, [_I64, _I64, _I64, _PI64, _PI32, _I64, _PF32, _PF32, _PI64]),
    "dmm_pr_blocks_check": (_I32, [_I64, _I64, _PI64, _PI32, _PF32, _PF32, _PI64]),
    "dmm_autotune": (_I32, [_P, _PI32, _I32, _I32, C.POINTER(Tuning)]),
    "dmm_set_tuning": (_I32, [_P, C.POINTER(Tuning)]),
    "dmm_get_tuning": (_I32, [_P, C.POINTER(Tuning)]),
}

_lib = None


def load():
    """Loads libdmm_gapbs.so once; raises RuntimeError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(make -C dist_mem_gapbs_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class DmmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(rc, ctx=None):
    if rc != DMM_OK:
        msg = load().dmm_last_error(ctx)
        raise DmmError(rc, msg.decode() if msg else "")
    return rc


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)
