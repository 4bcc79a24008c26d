"""ctypes binding of libluxum.so (the C ABI in include/luxum.h).

The library is built in-tree by ``make -C luxum_amd`` (or
``__graft_entry__.build()``).  There is no fallback: if the shared library is
missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

from .params import LxParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libluxum.so")

LX_OK, LX_EINVAL, LX_ENOMEM, LX_EDEVICE, LX_EIO, LX_EBUSY, LX_EUNSTABLE = range(7)
STATUS_NAMES = {0: "LX_OK", 1: "LX_EINVAL", 2: "LX_ENOMEM", 3: "LX_EDEVICE", 4: "LX_EIO", 5: "LX_EBUSY",
                6: "LX_EUNSTABLE"}

# Every symbol the header declares (checked by tests/test_abi.py).
EXPORTS = (
    "lx_params_default", "lx_params_load_yaml", "lx_params_check",
    "lx_ctx_create", "lx_ctx_destroy", "lx_ctx_mem_info",
    "lx_sim_create", "lx_sim_destroy", "lx_sim_set_materials", "lx_sim_set_materials3", "lx_sim_set_probes",
    "lx_sim_step", "lx_sim_sync", "lx_sim_read_field", "lx_sim_read_field_noalloc", "lx_sim_snapshot_async",
    "lx_sim_read_probes", "lx_sim_monitor", "lx_sim_max_e2", "lx_sim_max_e2_start", "lx_sim_max_e2_finish",
    "lx_sim_get_info", "lx_sim_stream",
    "lx_sim_time_kernels", "lx_sim_autotune",
    "lx_batch_create", "lx_batch_destroy", "lx_batch_add", "lx_batch_remove",
    "lx_batch_step", "lx_batch_sync",
    "lx_params_footprint", "lx_server_create", "lx_server_destroy", "lx_server_submit", "lx_server_run",
    "lx_server_info", "lx_server_probes", "lx_server_field", "lx_server_get_stats",
    "lx_writer_open", "lx_writer_write", "lx_writer_close", "lx_writer_count",
    "lx_h5_write", "lx_h5_write_dataset", "lx_h5_read", "lx_h5_dataset_name",
    "lx_last_error", "lx_version",
)


class LuxumError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class SimInfo(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("c0", C.c_double), ("eps0", C.c_double), ("mu0", C.c_double),
        ("eta0", C.c_double),
        ("cells", C.c_int64),
        ("slab_cells", C.c_int64 * 3),
        ("device_bytes", C.c_int64),
        ("bytes_per_step", C.c_int64),
        ("step", C.c_int32),
        ("slab_lo", C.c_int32 * 3),
        ("slab_hi", C.c_int32 * 3),
        ("fused_step", C.c_int32),
        ("kz", C.c_int32),
        ("waves", C.c_int32),
        ("core_bytes", C.c_int64),
        ("cells_per_lane", C.c_int32),
        ("shell_kz", C.c_int32),
        ("shell_concurrent", C.c_int32),
        ("shell_pipelined", C.c_int32),
        ("interior_rows", C.c_int32),
    ]


class Request(C.Structure):
    _fields_ = [
        ("params", C.c_void_p),
        ("ids", C.c_void_p),
        ("n_cells", C.c_size_t),
        ("eps_r", C.c_void_p),
        ("n_ids", C.c_int32),
        ("probes", C.c_void_p),
        ("n_probes", C.c_int32),
        ("steps", C.c_int32),
        ("arrival", C.c_double),
        ("keep_fields", C.c_int32),
    ]


class RequestInfo(C.Structure):
    _fields_ = [
        ("state", C.c_int32),
        ("steps_done", C.c_int32),
        ("arrival", C.c_double), ("admitted", C.c_double), ("finished", C.c_double),
        ("footprint", C.c_int64),
        ("cells", C.c_int64),
        ("n_probes", C.c_int32),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int32), ("completed", C.c_int32), ("failed", C.c_int32),
        ("peak_active", C.c_int32),
        ("peak_bytes", C.c_int64),
        ("ticks", C.c_int32),
        ("makespan", C.c_double),
        ("cell_updates", C.c_double),
        ("alg_bytes", C.c_double),
    ]


_lib = None


def lib():
    """Load libluxum.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C luxum_amd` "
                           "or __graft_entry__.build(); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    P = C.POINTER(LxParams)
    st = C.c_int
    sig = {
        "lx_params_default": ([P], st),
        "lx_params_load_yaml": ([C.c_char_p, P], st),
        "lx_params_check": ([P], st),
        "lx_ctx_create": ([C.c_int32, C.POINTER(vp)], st),
        "lx_ctx_destroy": ([vp], None),
        "lx_ctx_mem_info": ([vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], st),
        "lx_sim_create": ([vp, P, C.POINTER(vp)], st),
        "lx_sim_destroy": ([vp], None),
        "lx_sim_set_materials": ([vp, vp, vp, C.c_int32], st),
        "lx_sim_set_materials3": ([vp, vp, vp, C.c_int32], st),
        "lx_sim_set_probes": ([vp, vp, C.c_int32], st),
        "lx_sim_step": ([vp, C.c_int32], st),
        "lx_sim_sync": ([vp], st),
        "lx_sim_read_field": ([vp, C.c_int32, vp, C.c_size_t], st),
        "lx_sim_read_field_noalloc": ([vp, C.c_int32, vp, C.c_size_t], st),
        "lx_sim_snapshot_async": ([vp, C.c_int32, vp, C.c_size_t], st),
        "lx_sim_read_probes": ([vp, vp, C.c_size_t, C.POINTER(C.c_int32)], st),
        "lx_sim_monitor": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double)], st),
        "lx_sim_max_e2": ([vp, C.POINTER(C.c_double)], st),
        "lx_sim_max_e2_start": ([vp], st),
        "lx_sim_max_e2_finish": ([vp, C.POINTER(C.c_double)], st),
        "lx_sim_get_info": ([vp, C.POINTER(SimInfo)], st),
        "lx_sim_stream": ([vp], vp),
        "lx_sim_time_kernels": ([vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)], st),
        "lx_sim_autotune": ([vp, C.POINTER(C.c_int32)], st),
        "lx_batch_create": ([vp, C.POINTER(vp)], st),
        "lx_batch_destroy": ([vp], None),
        "lx_batch_add": ([vp, vp], st),
        "lx_batch_remove": ([vp, vp], st),
        "lx_batch_step": ([vp, C.c_int32], st),
        "lx_batch_sync": ([vp], st),
        "lx_params_footprint": ([P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int64)], st),
        "lx_server_create": ([vp, C.c_int32, C.c_int64, C.c_int32, C.POINTER(vp)], st),
        "lx_server_destroy": ([vp], None),
        "lx_server_submit": ([vp, C.POINTER(Request), C.POINTER(C.c_int64)], st),
        "lx_server_run": ([vp], st),
        "lx_server_info": ([vp, C.c_int64, C.POINTER(RequestInfo)], st),
        "lx_server_probes": ([vp, C.c_int64, vp, C.c_size_t], st),
        "lx_server_field": ([vp, C.c_int64, C.c_int32, vp, C.c_size_t], st),
        "lx_server_get_stats": ([vp, C.POINTER(ServerStats)], st),
        "lx_writer_open": ([vp, C.c_char_p, C.c_int32, C.c_int32, C.POINTER(vp)], st),
        "lx_writer_write": ([vp, C.c_double], st),
        "lx_writer_close": ([vp], st),
        "lx_writer_count": ([vp], C.c_int64),
        "lx_h5_write": ([C.c_char_p, C.c_double, vp, C.POINTER(C.c_int64), C.c_int32], st),
        "lx_h5_write_dataset": ([C.c_char_p, C.c_char_p, vp, C.POINTER(C.c_int64), C.c_int32], st),
        "lx_h5_read": ([C.c_char_p, C.c_char_p, vp, C.c_size_t, C.POINTER(C.c_int64)], st),
        "lx_h5_dataset_name": ([C.c_double, C.c_char_p, C.c_size_t], st),
        "lx_last_error": ([], C.c_char_p),
        "lx_version": ([], C.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(status: int) -> None:
    if status != LX_OK:
        msg = lib().lx_last_error()
        raise LuxumError(status, msg.decode() if msg else "")
