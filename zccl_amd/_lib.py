"""ctypes binding of zccl_amd/libzccl_amd.so (the C ABI in include/zccl_amd.h).

The library must have been built (``python -m zccl_amd.build``); there is no
fallback of any kind -- a missing library raises ``LibraryMissing``.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libzccl_amd.so")

ZCCL_OK = 0
STATUS_NAMES = {
    0: "ZCCL_OK",
    -1: "ZCCL_EINVAL",
    -2: "ZCCL_ENOSPC",
    -3: "ZCCL_ECORRUPT",
    -4: "ZCCL_EDEVICE",
    -5: "ZCCL_ENOMEM",
    -6: "ZCCL_EWORKSPACE",
}

# every symbol include/zccl_amd.h declares
EXPORTS = (
    "zccl_last_error", "zccl_strerror", "zccl_set_num_chunks", "zccl_get_num_chunks",
    "ZCCL_float_openmp_threadblock_arg", "ZCCL_float_single_thread_arg",
    "ZCCL_float_openmp_threadblock", "ZCCL_float_decompress_openmp_threadblock_arg",
    "ZCCL_float_decompress_single_thread_arg", "ZCCL_float_decompress_openmp_threadblock",
    "ZCCL_float_homomophic_add_openmp_threadblock", "ZCCL_float_homomophic_add_single_thread",
    "ZCCL_float_homomophic_add_openmp_threadblock_record",
    "zccl_decompress_sized", "zccl_homomorphic_add_sized",
    "zccl_worst_case_bytes", "zccl_index_bytes", "zccl_num_tiles", "zccl_num_blocks",
    "zccl_index_size_offset", "zccl_workspace_bytes",
    "zccl_workspace_init", "zccl_workspace_errors",
    "zccl_compress_dev", "zccl_decompress_dev", "zccl_homsum_dev", "zccl_homsum_dev_noalloc", "zccl_parse_dev",
    "zccl_sum_stage_bytes", "zccl_homsum_dev_staged", "zccl_batch_launch_staged", "zccl_calib_copy",
    "zccl_allreduce_workspace_bytes", "zccl_allreduce_create", "zccl_allreduce_run", "zccl_allreduce_segment",
    "zccl_allreduce_errors", "zccl_allreduce_destroy", "zccl_parse_workspace_bytes",
    "zccl_case_count_dev",
    "zccl_batch_homsum_create", "zccl_batch_compress_create", "zccl_batch_decompress_create",
    "zccl_batch_tiles", "zccl_batch_launch", "zccl_batch_destroy",
    "zccl_batch_homsum_plan", "zccl_batch_compress_plan", "zccl_batch_decompress_plan",
    "zccl_batch_desc_bytes", "zccl_batch_desc_host", "zccl_batch_bind",
    "zccl_tune_threads", "zccl_tune_get_threads", "zccl_tune_layout", "zccl_tune_get_layout",
    "zccl_tune_stage_bytes", "zccl_tune_get_stage_bytes", "zccl_tune_two_pass_min_tiles",
    "zccl_tune_get_two_pass_min_tiles", "zccl_tune_decompress_groups", "zccl_tune_get_decompress_groups",
    "SZp_compress3D_fast_openmp", "SZp_decompress3D_fast_openmp",
    "zccl_sz3d_sizes", "zccl_sz3d_compress_dev", "zccl_sz3d_decompress_dev", "zccl_sz3d_status",
)


class LibraryMissing(ImportError):
    pass


class ZcclError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        name = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{what}: {name}" if what else name)


_LIB = None


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} is not built; run `python -m zccl_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, i32, f32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_float
    ull_p = ctypes.POINTER(ctypes.c_ulonglong)
    sz_p = ctypes.POINTER(ctypes.c_size_t)
    i32_p = ctypes.POINTER(ctypes.c_int)
    sig = {
        "zccl_last_error": ([], i32),
        "zccl_strerror": ([i32], ctypes.c_char_p),
        "zccl_set_num_chunks": ([i32], None),
        "zccl_get_num_chunks": ([], i32),
        "ZCCL_float_openmp_threadblock_arg": ([vp, vp, sz_p, f32, sz, i32], None),
        "ZCCL_float_single_thread_arg": ([vp, vp, sz_p, f32, sz, i32], None),
        "ZCCL_float_openmp_threadblock": ([vp, sz_p, f32, sz, i32], vp),
        "ZCCL_float_decompress_openmp_threadblock_arg": ([vp, sz, f32, i32, vp], None),
        "ZCCL_float_decompress_single_thread_arg": ([vp, sz, f32, i32, vp], None),
        "ZCCL_float_decompress_openmp_threadblock": ([ctypes.POINTER(vp), sz, f32, i32, vp], None),
        "ZCCL_float_homomophic_add_openmp_threadblock": ([vp, sz_p, sz, f32, i32, vp, vp], None),
        "ZCCL_float_homomophic_add_single_thread": ([vp, sz_p, sz, f32, i32, vp, vp], None),
        "ZCCL_float_homomophic_add_openmp_threadblock_record":
            ([vp, sz_p, sz, f32, i32, vp, vp, i32_p, i32_p, i32_p, i32_p], None),
        "zccl_decompress_sized": ([vp, sz, f32, i32, i32, vp, sz], i32),
        "zccl_homomorphic_add_sized": ([vp, sz, sz_p, sz, i32, i32, vp, sz, vp, sz], i32),
        "zccl_worst_case_bytes": ([sz, i32, i32], sz),
        "zccl_index_bytes": ([sz, i32, i32], sz),
        "zccl_num_tiles": ([sz, i32, i32], sz),
        "zccl_num_blocks": ([sz, i32, i32], sz),
        "zccl_index_size_offset": ([sz, i32, i32], sz),
        "zccl_workspace_bytes": ([sz], sz),
        "zccl_workspace_init": ([vp, sz, vp], i32),
        "zccl_workspace_errors": ([vp, ull_p, i32, vp], i32),
        "zccl_compress_dev": ([vp, sz, f32, i32, i32, vp, sz, vp, vp, vp, sz, vp], i32),
        "zccl_decompress_dev": ([vp, vp, sz, f32, i32, i32, vp, vp, sz, vp], i32),
        "zccl_homsum_dev": ([vp, vp, i32, sz, i32, i32, vp, sz, vp, vp, vp, sz, vp], i32),
        "zccl_homsum_dev_noalloc": ([vp, vp, i32, sz, i32, i32, vp, sz, vp, vp, vp], i32),
        "zccl_sum_stage_bytes": ([sz], sz),
        "zccl_calib_copy": ([vp, vp, sz, i32, vp], i32),
        "zccl_allreduce_workspace_bytes": ([i32, sz, i32, i32], sz),
        "zccl_allreduce_create": ([i32, sz, ctypes.c_float, i32, i32, vp, sz, vp], vp),
        "zccl_allreduce_run": ([vp, vp, vp, vp], i32),
        "zccl_allreduce_segment": ([vp, i32, vp, vp, vp, vp], i32),
        "zccl_allreduce_errors": ([vp, vp, i32, vp], i32),
        "zccl_allreduce_destroy": ([vp], None),
        "zccl_parse_workspace_bytes": ([sz, i32, i32], sz),
        "zccl_homsum_dev_staged": ([vp, vp, i32, sz, i32, i32, vp, sz, vp, vp, vp, sz, vp, sz, vp], i32),
        "zccl_batch_launch_staged": ([vp, vp, sz, vp, sz, vp], i32),
        "zccl_parse_dev": ([vp, sz, sz, i32, i32, vp, vp, vp, sz, vp], i32),
        "zccl_case_count_dev": ([vp, vp, sz, i32, i32, vp, vp], i32),
        "zccl_batch_homsum_create": ([i32, i32, vp, i32, i32, vp, vp, vp, vp, vp, i32_p], vp),
        "zccl_batch_compress_create": ([i32, vp, vp, i32, i32, vp, vp, vp, vp, i32_p], vp),
        "zccl_batch_decompress_create": ([i32, vp, vp, i32, i32, vp, vp, vp, i32_p], vp),
        "zccl_batch_tiles": ([vp], sz),
        "zccl_batch_launch": ([vp, vp, sz, vp], i32),
        "zccl_batch_destroy": ([vp], None),
        "zccl_batch_homsum_plan": ([i32, i32_p, vp, i32, i32, vp, vp, vp, vp, vp, i32_p], vp),
        "zccl_batch_compress_plan": ([i32, vp, vp, i32, i32, vp, vp, vp, vp, i32_p], vp),
        "zccl_batch_decompress_plan": ([i32, vp, vp, i32, i32, vp, vp, vp, i32_p], vp),
        "zccl_batch_desc_bytes": ([vp], sz),
        "zccl_batch_desc_host": ([vp], vp),
        "zccl_batch_bind": ([vp, vp], i32),
        "SZp_compress3D_fast_openmp": ([vp, vp, sz, sz, sz, i32, ctypes.c_double, sz_p], None),
        "SZp_decompress3D_fast_openmp": ([vp, vp, sz, sz, sz, i32, ctypes.c_double], None),
        "zccl_sz3d_sizes": ([sz, sz, sz, i32, i32, sz_p, sz_p], i32),
        "zccl_sz3d_compress_dev": ([vp, sz, sz, sz, i32, ctypes.c_double, i32, vp, sz, vp, sz, vp, vp], i32),
        "zccl_sz3d_decompress_dev": ([vp, sz, sz, sz, sz, i32, ctypes.c_double, i32, vp, vp, sz, vp], i32),
        "zccl_sz3d_status": ([vp, sz_p], i32),
        "zccl_tune_threads": ([i32, i32], i32),
        "zccl_tune_get_threads": ([i32], i32),
        "zccl_tune_layout": ([i32, i32], i32),
        "zccl_tune_get_layout": ([i32], i32),
        "zccl_tune_stage_bytes": ([ctypes.c_uint], i32),
        "zccl_tune_get_stage_bytes": ([], ctypes.c_uint),
        "zccl_tune_two_pass_min_tiles": ([ctypes.c_ulonglong], i32),
        "zccl_tune_get_two_pass_min_tiles": ([], ctypes.c_ulonglong),
        "zccl_tune_decompress_groups": ([i32], i32),
        "zccl_tune_get_decompress_groups": ([], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = L
    return L


def check(status: int, what: str = "") -> None:
    if status != ZCCL_OK:
        raise ZcclError(status, what)
