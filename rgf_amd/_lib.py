"""ctypes binding of librgf_hip.so (the C ABI declared in include/rgf.h).

The shared library is built in-tree by ``rgf_amd.build.build()``; loading fails
loudly when it is missing -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librgf_hip.so")

RGF_OK, RGF_EINVAL, RGF_ENOTSPD, RGF_ENOMEM, RGF_EDEVICE, RGF_ESTATE = range(6)
RGF_FP64, RGF_MIXED_BF16X3 = 0, 2
RGF_SELINV_EXACT, RGF_SELINV_REFERENCE = range(2)
RGF_CHOLESKY_AUTO, RGF_CHOLESKY_TASK_GRAPH, RGF_CHOLESKY_TILE_LAUNCHES = range(3)

STATUS_NAMES = {0: "RGF_OK", 1: "RGF_EINVAL", 2: "RGF_ENOTSPD", 3: "RGF_ENOMEM", 4: "RGF_EDEVICE",
                5: "RGF_ESTATE"}

# Every symbol include/rgf.h declares (checked by tests/test_abi.py).
EXPORTS = [
    "rgf_default_options", "rgf_default_gen_params", "rgf_create_csc", "rgf_create_generated",
    "rgf_destroy", "rgf_factorize", "rgf_solve", "rgf_logdet", "rgf_selinv_diag",
    "rgf_residual_norm", "rgf_residual_norm_normalized", "rgf_last_error", "rgf_solve_device",
    "rgf_selinv_diag_device", "rgf_selinv_blocks_device", "rgf_synchronize", "rgf_get_info", "rgf_export_matrix_mf",
    "rgf_export_factor_mf", "rgf_kernel_timing_enable", "rgf_kernel_timing_read",
    "rgf_create_generated_batch", "rgf_batch_count", "rgf_select", "rgf_regenerate", "rgf_footprint_bytes",
    "rgf_logdet_batch", "rgf_selinv_diag_batch", "rgf_refine_stats", "rgf_set_device_cache_limit",
    "rgf_trim_device_cache", "rgf_solve_device_noalloc", "rgf_set_autotune", "rgf_autotune_entries",
    "rgf_autotune_clear", "rgf_autotune_save", "rgf_autotune_load", "rgf_group_create", "rgf_group_destroy",
    "rgf_group_factorize", "rgf_group_selinv_diag_device", "rgf_group_synchronize",
    "rgf_z_create_csc", "rgf_z_destroy", "rgf_z_factorize", "rgf_z_solve", "rgf_z_logdet", "rgf_z_selinv_diag",
    "rgf_z_residual_norm", "rgf_z_residual_norm_normalized", "rgf_z_last_error",
]
# include/rgf_testing.h
TEST_EXPORTS = ["rgf_debug_gemm", "rgf_debug_potrf", "rgf_debug_bench_gemm", "rgf_debug_bench_potrf",
                "rgf_debug_trace_potrf", "rgf_debug_dag_order",
                "rgf_debug_cache_stats", "rgf_debug_set_tuning"]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int), ("selinv_mode", C.c_int), ("keep_matrix", C.c_int),
                ("device", C.c_int), ("refine_iters", C.c_int), ("refine_tol", C.c_double),
                ("graphs", C.c_int), ("keep_splits", C.c_int), ("cholesky", C.c_int), ("keep_w", C.c_int)]


class GenParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("matrix_id", C.c_uint32), ("sigma_c", C.c_double),
                ("sigma_a", C.c_double), ("theta_d", C.c_double), ("theta_c", C.c_double),
                ("theta_mode", C.c_int)]


class Info(C.Structure):
    _fields_ = [("ns", C.c_size_t), ("nt", C.c_size_t), ("nd", C.c_size_t), ("n", C.c_size_t),
                ("device_bytes", C.c_size_t), ("flops_factor", C.c_double),
                ("flops_selinv", C.c_double), ("bytes_solve_per_rhs", C.c_double)]


_lib = None
_lock = threading.Lock()


def load():
    """Load librgf_hip.so; raises FileNotFoundError if it was not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(there is no non-HIP fallback)")
        lib = C.CDLL(LIB_PATH)
        H = C.c_void_p
        dp = C.POINTER(C.c_double)
        szp = C.POINTER(C.c_size_t)
        lib.rgf_default_options.argtypes = [C.POINTER(Options)]
        lib.rgf_default_options.restype = None
        lib.rgf_default_gen_params.argtypes = [C.POINTER(GenParams)]
        lib.rgf_default_gen_params.restype = None
        lib.rgf_create_csc.argtypes = [szp, szp, dp, C.c_size_t, C.c_size_t, C.c_size_t,
                                       C.POINTER(Options), C.POINTER(H)]
        lib.rgf_create_generated.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, C.POINTER(GenParams),
                                             C.POINTER(Options), C.POINTER(H)]
        lib.rgf_destroy.argtypes = [H]
        lib.rgf_factorize.argtypes = [H]
        lib.rgf_solve.argtypes = [H, dp, dp, C.c_size_t]
        lib.rgf_logdet.argtypes = [H, dp]
        lib.rgf_selinv_diag.argtypes = [H, dp]
        lib.rgf_residual_norm.argtypes = [H, dp, dp, dp]
        lib.rgf_residual_norm_normalized.argtypes = [H, dp, dp, dp]
        lib.rgf_last_error.argtypes = [H]
        lib.rgf_last_error.restype = C.c_char_p
        lib.rgf_solve_device.argtypes = [H, C.c_void_p, C.c_void_p, C.c_size_t]
        lib.rgf_solve_device_noalloc.argtypes = [H, C.c_void_p, C.c_void_p, C.c_size_t]
        lib.rgf_selinv_diag_device.argtypes = [H, C.c_void_p]
        lib.rgf_selinv_blocks_device.argtypes = [H, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.rgf_synchronize.argtypes = [H]
        lib.rgf_get_info.argtypes = [H, C.POINTER(Info)]
        lib.rgf_export_matrix_mf.argtypes = [H, dp]
        lib.rgf_export_factor_mf.argtypes = [H, dp]
        lib.rgf_kernel_timing_enable.argtypes = [C.c_int]
        lib.rgf_kernel_timing_read.argtypes = [dp, dp, C.POINTER(C.c_longlong)]
        lib.rgf_create_generated_batch.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                                   C.POINTER(GenParams), C.POINTER(Options), C.POINTER(H)]
        lib.rgf_batch_count.argtypes = [H]
        lib.rgf_select.argtypes = [H, C.c_size_t]
        lib.rgf_regenerate.argtypes = [H]
        lib.rgf_footprint_bytes.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.POINTER(Options)]
        lib.rgf_logdet_batch.argtypes = [H, dp]
        lib.rgf_refine_stats.argtypes = [H, C.POINTER(C.c_int), dp]
        lib.rgf_selinv_diag_batch.argtypes = [H, dp]
        lib.rgf_set_device_cache_limit.argtypes = [C.c_size_t, C.POINTER(C.c_size_t)]
        lib.rgf_z_create_csc.argtypes = [szp, szp, dp, C.c_size_t, C.c_size_t, C.c_size_t, C.POINTER(Options),
                                         C.POINTER(H)]
        lib.rgf_z_destroy.argtypes = [H]
        lib.rgf_z_factorize.argtypes = [H]
        lib.rgf_z_solve.argtypes = [H, dp, dp, C.c_size_t]
        lib.rgf_z_logdet.argtypes = [H, dp]
        lib.rgf_z_selinv_diag.argtypes = [H, dp]
        lib.rgf_z_residual_norm.argtypes = [H, dp, dp, dp]
        lib.rgf_z_residual_norm_normalized.argtypes = [H, dp, dp, dp]
        lib.rgf_z_last_error.argtypes = [H]
        lib.rgf_z_last_error.restype = C.c_char_p
        lib.rgf_trim_device_cache.argtypes = [C.c_int]
        lib.rgf_set_autotune.argtypes = [C.c_int, C.POINTER(C.c_int)]
        lib.rgf_autotune_entries.argtypes = []
        lib.rgf_autotune_clear.argtypes = []
        lib.rgf_autotune_save.argtypes = [C.c_char_p]
        lib.rgf_autotune_load.argtypes = [C.c_char_p]
        lib.rgf_group_create.argtypes = [C.POINTER(H), C.c_size_t, C.POINTER(H)]
        lib.rgf_group_destroy.argtypes = [H]
        lib.rgf_group_factorize.argtypes = [H]
        lib.rgf_group_selinv_diag_device.argtypes = [H, C.POINTER(C.c_void_p)]
        lib.rgf_group_synchronize.argtypes = [H]
        lib.rgf_debug_gemm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, dp, C.c_int,
                                       dp, C.c_int, C.c_double, dp, C.c_int, C.c_int]
        lib.rgf_debug_gemm.restype = C.c_int
        lib.rgf_debug_potrf.argtypes = [C.c_int, dp, C.c_int, C.c_int, dp]
        lib.rgf_debug_potrf.restype = C.c_int
        lib.rgf_debug_bench_gemm.argtypes = [C.c_int] * 7 + [dp]
        lib.rgf_debug_bench_gemm.restype = C.c_int
        lib.rgf_debug_bench_potrf.argtypes = [C.c_int, C.c_int, dp, dp]
        lib.rgf_debug_bench_potrf.restype = C.c_int
        lp = C.POINTER(C.c_longlong)
        lib.rgf_debug_trace_potrf.argtypes = [C.c_int, lp, C.c_longlong, lp, dp]
        lib.rgf_debug_trace_potrf.restype = C.c_int
        lib.rgf_debug_dag_order.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int]
        lib.rgf_debug_dag_order.restype = C.c_int
        lib.rgf_debug_cache_stats.argtypes = [lp, lp, lp, dp]
        lib.rgf_debug_cache_stats.restype = C.c_int
        lib.rgf_debug_set_tuning.argtypes = [C.c_int, C.c_int]
        lib.rgf_debug_set_tuning.restype = C.c_int
        for name in EXPORTS:
            getattr(lib, name).restype = getattr(lib, name).restype if name in (
                "rgf_last_error", "rgf_z_last_error", "rgf_default_options", "rgf_default_gen_params") else C.c_int
        lib.rgf_footprint_bytes.restype = C.c_size_t
        _lib = lib
        return lib


# ---------------------------------------------------------------- minimal HIP runtime access
_hip = None


def hip():
    """libamdhip64 for device buffers used by the device-resident entry points."""
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _hip = h
    return _hip


class DeviceBuffer:
    """A hipMalloc'ed FP64 buffer (for the *_device entry points)."""

    H2D, D2H = 1, 2

    def __init__(self, count: int):
        self.count = int(count)
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), max(self.count, 1) * 8)
        if rc != 0:
            raise MemoryError(f"hipMalloc({self.count * 8}) failed with {rc}")

    def upload(self, arr) -> None:
        import numpy as np
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.size <= self.count
        rc = hip().hipMemcpy(self.ptr, a.ctypes.data, a.size * 8, self.H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed with {rc}")

    def download(self, count: int | None = None):
        import numpy as np
        cnt = self.count if count is None else int(count)
        out = np.empty(cnt, dtype=np.float64)
        rc = hip().hipMemcpy(out.ctypes.data, self.ptr, cnt * 8, self.D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed with {rc}")
        return out

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gpu_count() -> int:
    try:
        n = C.c_int(0)
        if hip().hipGetDeviceCount(C.byref(n)) != 0:
            return 0
        return n.value
    except OSError:
        return 0
