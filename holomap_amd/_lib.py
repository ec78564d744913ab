"""ctypes binding of libholomap_gpu.so (the C ABI in include/holomap_gpu.h).

There is no CPU fallback: if the shared library is missing or no GPU is
present, the calls raise.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libholomap_gpu.so")

HM_OK = 0
HM_EBADSHAPE = -1
HM_ETOOFEW = -2
HM_EINVAL = -3
HM_ENONFINITE = -4
HM_EDEVICE = -5
HM_ENOMEM = -6
HM_EBUDGET = -7
HM_EUNSUPPORTED = -8
HM_MAX_NEIGHBORS = 1024
HM_NOALLOC_SCRATCH_BYTES = 256 << 20

HM_EUCLIDEAN = 0
HM_COSINE = 1
HM_PREC_EXACT = 0
HM_PREC_FP32 = 1
HM_PREC_BF16 = 2
HM_PREC_BF16X3 = 3
HM_CALIB_REF_F32 = 0
HM_CALIB_F64 = 1
HM_STAGE_KNN = 1
HM_STAGE_CALIB = 2
HM_STAGE_FUZZY = 4
HM_STAGE_ALL = 7
HM_NTIMES = 8
HM_NCOUNTERS = 8
TIME_NAMES = ["prep", "sweep", "rerank", "fallback", "calib", "fuzzy", "total", "threshold"]

# every symbol include/holomap_gpu.h declares
EXPORTS = [
    "hm_default_options", "hm_version", "hm_ctx_create", "hm_ctx_destroy", "hm_last_error",
    "hm_device_count", "hm_validate", "hm_exact_knn", "hm_smooth_knn",
    "hm_fuzzy_simplicial_set", "hm_csr_from_fuzzy", "hm_knn_graph", "hm_job_create",
    "hm_job_destroy", "hm_job_bytes", "hm_job_estimate_bytes", "hm_job_upload",
    "hm_job_upload_device", "hm_job_set_knn", "hm_job_set_calib", "hm_job_run", "hm_job_sync",
    "hm_job_done", "hm_job_download", "hm_job_times", "hm_job_counters",
    "hm_job_device_outputs", "hm_expf_array", "hm_ctx_budget", "hm_ctx_in_use",
    "hm_server_default_config", "hm_server_create", "hm_server_destroy", "hm_server_submit",
    "hm_server_progress", "hm_server_state", "hm_server_pending", "hm_server_result_info",
    "hm_server_fetch", "hm_knn_graph_noalloc", "hm_connected_components", "hm_laplacian_new",
    "hm_laplacian_matvec", "hm_job_connected_components", "hm_job_laplacian_matvec",
    "hm_schedule_edges", "hm_job_schedule_edges", "hm_l2_normalise",
]


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("calib", ctypes.c_int32),
                ("margin", ctypes.c_int32), ("nsplit", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("autotune", ctypes.c_int32),
                ("timing", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class GraphOut(ctypes.Structure):
    _fields_ = [("indices", ctypes.c_void_p), ("dists", ctypes.c_void_p),
                ("sigmas", ctypes.c_void_p), ("rhos", ctypes.c_void_p),
                ("rows", ctypes.c_void_p), ("cols", ctypes.c_void_p),
                ("vals", ctypes.c_void_p), ("indptr", ctypes.c_void_p),
                ("sigmas64", ctypes.c_void_p), ("cap", ctypes.c_uint64),
                ("nnz", ctypes.c_uint64), ("dists64", ctypes.c_void_p),
                ("err64", ctypes.c_double * 3)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_jobs", ctypes.c_uint32), ("max_batch", ctypes.c_uint32),
                ("max_batch_rows", ctypes.c_uint64), ("opt", Options)]


_lib = None


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with python -m holomap_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32
    P = ctypes.POINTER
    sig = {
        "hm_default_options": (None, [P(Options)]),
        "hm_version": (ctypes.c_char_p, []),
        "hm_ctx_create": (i32, [P(vp), ctypes.c_int, u64]),
        "hm_ctx_destroy": (None, [vp]),
        "hm_last_error": (ctypes.c_char_p, [vp]),
        "hm_device_count": (ctypes.c_int, []),
        "hm_validate": (i32, [vp, u64, u32, u32, P(u64)]),
        "hm_exact_knn": (i32, [vp, vp, u64, u32, u32, i32, P(Options), vp, vp]),
        "hm_smooth_knn": (i32, [vp, vp, u64, u32, i32, vp, vp]),
        "hm_fuzzy_simplicial_set": (i32, [vp, vp, vp, vp, vp, u64, u32, vp, vp, vp, u64, P(u64)]),
        "hm_csr_from_fuzzy": (i32, [vp, vp, u64, u64, vp]),
        "hm_knn_graph": (i32, [vp, vp, u64, u32, u32, i32, P(Options), P(GraphOut)]),
        "hm_knn_graph_noalloc": (i32, [vp, vp, u64, u32, u32, i32, P(Options), P(GraphOut)]),
        "hm_connected_components": (i32, [vp, vp, vp, u64, vp, P(u64)]),
        "hm_laplacian_new": (i32, [vp, vp, vp, vp, u64, vp, vp]),
        "hm_laplacian_matvec": (i32, [vp, vp, vp, vp, u64, vp, vp]),
        "hm_job_connected_components": (i32, [vp, u32, vp, P(u64)]),
        "hm_job_laplacian_matvec": (i32, [vp, u32, vp, vp]),
        "hm_schedule_edges": (i32, [vp, vp, vp, vp, u64, u64, u32, vp, vp, vp, P(u64)]),
        "hm_job_schedule_edges": (i32, [vp, u32, u32, vp, vp, vp, P(u64)]),
        "hm_l2_normalise": (i32, [vp, vp, u64, u32, vp]),
        "hm_job_create": (i32, [vp, u32, vp, vp, vp, vp, P(Options), u32, P(vp)]),
        "hm_job_destroy": (None, [vp]),
        "hm_job_bytes": (u64, [vp]),
        "hm_job_estimate_bytes": (u64, [u32, vp, vp, vp, vp, P(Options)]),
        "hm_job_upload": (i32, [vp, u32, vp]),
        "hm_job_upload_device": (i32, [vp, u32, vp]),
        "hm_job_set_knn": (i32, [vp, u32, vp, vp]),
        "hm_job_set_calib": (i32, [vp, u32, vp, vp]),
        "hm_job_run": (i32, [vp]),
        "hm_job_sync": (i32, [vp]),
        "hm_job_done": (ctypes.c_int, [vp]),
        "hm_job_download": (i32, [vp, u32, P(GraphOut)]),
        "hm_job_times": (i32, [vp, vp]),
        "hm_job_counters": (i32, [vp, vp]),
        "hm_job_device_outputs": (i32, [vp, u32, P(vp), P(vp), P(vp), P(vp), P(vp), P(vp)]),
        "hm_expf_array": (i32, [vp, vp, vp, u64]),
        "hm_ctx_budget": (u64, [vp]),
        "hm_ctx_in_use": (u64, [vp]),
        "hm_server_default_config": (None, [P(ServerConfig)]),
        "hm_server_create": (i32, [vp, P(ServerConfig), P(vp)]),
        "hm_server_destroy": (None, [vp]),
        "hm_server_submit": (i32, [vp, vp, u64, u32, u32, i32, P(u64)]),
        "hm_server_progress": (i32, [vp, P(u32)]),
        "hm_server_state": (ctypes.c_int, [vp, u64]),
        "hm_server_pending": (u32, [vp]),
        "hm_server_result_info": (i32, [vp, u64, P(u64), P(u32), P(u64)]),
        "hm_server_fetch": (i32, [vp, u64, P(GraphOut)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def default_options(**kw) -> Options:
    o = Options()
    load().hm_default_options(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, int(v))
    return o
