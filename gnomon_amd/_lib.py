"""ctypes binding of libgnomon_hip.so (the C ABI in include/gnomon_hip.h).

The shared library is built in-tree by ``gnomon_amd.build.build_library()``.
Loading fails loudly when it is missing: there is no CPU fallback in this
package.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GNOMON_HIP_LIB") or os.path.join(HERE, "libgnomon_hip.so")

GNM_OK = 0
GNM_DECLINED = 1
GNM_ERR_NO_DEVICE = 2
GNM_ERR_INVALID_ARG = 3
GNM_ERR_OUT_OF_MEMORY = 4
GNM_ERR_HIP = 5
GNM_ERR_INTERNAL = 6
GNM_ERR_UNSUPPORTED = 7

STATUS_NAMES = {
    GNM_OK: "OK", GNM_DECLINED: "DECLINED", GNM_ERR_NO_DEVICE: "NO_DEVICE",
    GNM_ERR_INVALID_ARG: "INVALID_ARG", GNM_ERR_OUT_OF_MEMORY: "OUT_OF_MEMORY",
    GNM_ERR_HIP: "HIP_ERROR", GNM_ERR_INTERNAL: "INTERNAL", GNM_ERR_UNSUPPORTED: "UNSUPPORTED",
}

PRECISION = {"fp32": 0, "bf16x2": 1, "bf16x3": 2, "auto": 3, "i8x3": 4, "i8x4": 5, "bf16x4": 6}

# Every entry point declared in include/gnomon_hip.h.
EXPORTS = (
    "gnm_last_error", "gnm_abi_version", "gnm_size_gate", "gnm_device_count", "gnm_create",
    "gnm_create_forced", "gnm_score_rows", "gnm_score_rows_device", "gnm_read_accumulators",
    "gnm_seed_accumulators", "gnm_finish", "gnm_destroy", "gnm_run", "gnm_unpack_device",
    "gnm_dense_forward_device", "gnm_dense_adjoint_device", "gnm_dense_adjoint_device_noalloc",
    "gnm_synth_rows_device", "gnm_hbm_probe", "gnm_context_shape",
    "gnm_last_timing", "gnm_set_launch_config", "gnm_set_sparse_ingest", "gnm_ingest_stats",
    "gnm_sparse_encode_row",
)


class GnmError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Problem(ctypes.Structure):
    _fields_ = [
        ("num_people", ctypes.c_uint64),
        ("total_people_in_fam", ctypes.c_uint64),
        ("bytes_per_variant", ctypes.c_uint64),
        ("num_scores", ctypes.c_uint64),
        ("num_variants", ctypes.c_uint64),
        ("nnz", ctypes.c_uint64),
        ("sparse_weights", ctypes.c_void_p),
        ("sparse_missing_corrections", ctypes.c_void_p),
        ("sparse_score_columns", ctypes.c_void_p),
        ("sparse_row_offsets", ctypes.c_void_p),
        ("output_idx_to_fam_idx", ctypes.c_void_p),
        ("precision", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("reserved", ctypes.c_uint64 * 4),
    ]


_lib = None


def lib():
    """Load libgnomon_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with gnomon_amd.build.build_library() "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32
    st = ctypes.c_int
    L.gnm_last_error.restype = ctypes.c_char_p
    L.gnm_last_error.argtypes = []
    L.gnm_abi_version.restype = ctypes.c_int
    L.gnm_size_gate.restype = ctypes.c_int
    L.gnm_size_gate.argtypes = [u64, u64, u64]
    L.gnm_device_count.restype = ctypes.c_int
    L.gnm_device_count.argtypes = []
    for name in ("gnm_create", "gnm_create_forced"):
        f = getattr(L, name)
        f.restype = st
        f.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(vp)]
    for name in ("gnm_score_rows", "gnm_score_rows_device"):
        f = getattr(L, name)
        f.restype = st
        f.argtypes = [vp, vp, u64, u64, u64]
    L.gnm_read_accumulators.restype = st
    L.gnm_read_accumulators.argtypes = [vp, vp, vp, ctypes.POINTER(u64)]
    L.gnm_seed_accumulators.restype = st
    L.gnm_seed_accumulators.argtypes = [vp, vp, vp, u64]
    L.gnm_finish.restype = st
    L.gnm_finish.argtypes = [vp, vp, vp]
    L.gnm_destroy.restype = None
    L.gnm_destroy.argtypes = [vp]
    L.gnm_run.restype = st
    L.gnm_run.argtypes = [ctypes.POINTER(Problem), vp, u64, vp, vp]
    L.gnm_unpack_device.restype = st
    L.gnm_unpack_device.argtypes = [vp, u64, u64, u64, vp, vp, vp, vp]
    L.gnm_dense_forward_device.restype = st
    L.gnm_dense_forward_device.argtypes = [vp, u64, u64, u64, vp, vp, u64, i32, vp, vp, vp]
    for name in ("gnm_dense_adjoint_device", "gnm_dense_adjoint_device_noalloc"):
        f = getattr(L, name)
        f.restype = st
        f.argtypes = [vp, u64, u64, u64, vp, vp, u64, i32, vp, vp]
    L.gnm_synth_rows_device.restype = st
    L.gnm_synth_rows_device.argtypes = [u64, u32, u32, u32, u64, u64, u64, vp, u64, vp]
    L.gnm_context_shape.restype = st
    L.gnm_context_shape.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u32)]
    L.gnm_hbm_probe.restype = st
    L.gnm_hbm_probe.argtypes = [i32, vp, u64, vp, u64, vp]
    L.gnm_set_launch_config.restype = st
    L.gnm_set_launch_config.argtypes = [vp, u32, u32]
    L.gnm_last_timing.restype = st
    L.gnm_last_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    L.gnm_set_sparse_ingest.restype = st
    L.gnm_set_sparse_ingest.argtypes = [vp, ctypes.c_double]
    L.gnm_ingest_stats.restype = st
    L.gnm_ingest_stats.argtypes = [vp] + [ctypes.POINTER(u64)] * 4
    L.gnm_sparse_encode_row.restype = ctypes.c_int64
    L.gnm_sparse_encode_row.argtypes = [vp, u64, vp, u64, vp, vp, vp]
    _lib = L
    return L


def last_error() -> str:
    return lib().gnm_last_error().decode("utf-8", "replace")


def check(status: int) -> None:
    if status != GNM_OK:
        raise GnmError(status, last_error())
