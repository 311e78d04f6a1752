"""ctypes binding of ``libriga_hip.so`` (the C ABI declared in include/riga_hip.h).

The library is built in-tree by ``riga_flow_amd/csrc/Makefile``.  There is no
CPU fallback: when the library or the device is missing every pricing call
raises ``RigaError("gpu_unavailable", ...)``, as the reference's GPU build
does (crates/riga-engine/src/lib.rs:188-202, gpu.rs:490-497).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import threading

import numpy as np

from .types import RigaError

HERE = os.path.dirname(os.path.abspath(__file__))
# RIGA_HIP_LIB names another build of the library (e.g. one built with other
# compile-time options); by default the in-tree one is loaded.
LIB_PATH = os.environ.get("RIGA_HIP_LIB") or os.path.join(HERE, "libriga_hip.so")
TRACE_CAP = 64

FP64 = 0
FP32 = 1
VR_NONE, VR_HALTON, VR_SOBOL = 0, 1, 2
TRACE = {"off": 0, "aggregated": 1, "sampled": 2, "full": 3}

_D = C.POINTER(C.c_double)


class Error(C.Structure):
    _fields_ = [("code", C.c_char * 48), ("message", C.c_char * 464)]


class ImpliedGrid(C.Structure):
    _fields_ = [("n_maturities", C.c_int32), ("n_strikes", C.c_int32),
                ("maturities", _D), ("strikes", _D), ("vols", _D)]


class LvGrid(C.Structure):
    _fields_ = [("n_t", C.c_int32), ("n_x", C.c_int32), ("t_axis", _D), ("log_s", _D), ("v", _D)]


class Request(C.Structure):
    _fields_ = [
        ("factors", C.c_int32),
        ("spots", _D), ("vols", _D), ("dividend_yields", _D), ("correlation", _D),
        ("risk_free_rate", C.c_double),
        ("initial_spots", _D),
        ("maturity_years", C.c_double),
        ("steps", C.c_int64),
        ("strike", C.c_double), ("knock_in_barrier", C.c_double),
        ("has_knock_out", C.c_int32), ("knock_out_barrier", C.c_double),
        ("has_coupon_barrier", C.c_int32), ("coupon_barrier", C.c_double),
        ("has_coupon_rate", C.c_int32), ("coupon_rate", C.c_double),
        ("memory_coupon", C.c_int32),
        ("coupon_periods", C.c_int64),
        ("fixed_first_coupon", C.c_int32),
        ("call_start_step", C.c_int64),
        ("paths", C.c_uint64), ("path_block_size", C.c_uint64), ("seed", C.c_uint64),
        ("trace_level", C.c_int32), ("ladder", C.c_int32), ("ladder_dimension", C.c_int32),
        ("max_dimension", C.c_int32),
        ("implied", C.POINTER(ImpliedGrid)),
        ("calibrated", C.POINTER(LvGrid)),
        ("grid_x", C.c_int32),
        ("surfaces", C.POINTER(C.c_void_p)),
    ]


class Result(C.Structure):
    _fields_ = [
        ("pv", C.c_double), ("mc_error", C.c_double), ("ki_probability", C.c_double),
        ("leg_coupon", C.c_double), ("leg_memory_coupon", C.c_double), ("leg_put", C.c_double),
        ("leg_funding", C.c_double), ("leg_total", C.c_double),
        ("paths", C.c_uint64), ("ko_count", C.c_uint64), ("ki_count", C.c_uint64),
        ("put_eligible_count", C.c_uint64), ("coupon_path_count", C.c_uint64),
        ("ko_mean", C.c_double), ("ko_variance", C.c_double), ("ki_mean", C.c_double),
        ("ki_variance", C.c_double), ("coupon_mean", C.c_double), ("coupon_variance", C.c_double),
        ("n_trace", C.c_int32),
        ("trace_path", C.c_int64 * TRACE_CAP),
        ("trace_step", C.c_int64 * TRACE_CAP),
        ("device_ms", C.c_double),
    ]


class PathRecord(C.Structure):
    _fields_ = [("ko_step", C.c_int32), ("hit_ki", C.c_int32), ("final_worst", C.c_double),
                ("funding", C.c_double), ("put", C.c_double), ("coupon_pv", C.c_double),
                ("memory_pv", C.c_double), ("pv", C.c_double)]


class CouponRow(C.Structure):
    _fields_ = [("period", C.c_int64), ("years", C.c_double), ("qualified_steps", C.c_int64),
                ("period_steps", C.c_int64), ("accrual", C.c_double), ("memory_paid", C.c_double),
                ("paid", C.c_double), ("memory_accrued", C.c_double), ("discount_factor", C.c_double),
                ("pv", C.c_double), ("knocked_out", C.c_int32), ("pad", C.c_int32)]


COUPON_DTYPE = np.dtype([("period", "<i8"), ("years", "<f8"), ("qualified_steps", "<i8"),
                         ("period_steps", "<i8"), ("accrual", "<f8"), ("memory_paid", "<f8"),
                         ("paid", "<f8"), ("memory_accrued", "<f8"), ("discount_factor", "<f8"),
                         ("pv", "<f8"), ("knocked_out", "<i4"), ("pad", "<i4")])


class WorstQuery(C.Structure):
    _fields_ = [("percentile", C.c_double), ("threshold_worst_final", C.c_double),
                ("selected", C.c_uint64), ("total", C.c_uint64), ("n_sample", C.c_uint64)]


RECORD_DTYPE = np.dtype([("ko_step", "<i4"), ("hit_ki", "<i4"), ("final_worst", "<f8"),
                         ("funding", "<f8"), ("put", "<f8"), ("coupon_pv", "<f8"),
                         ("memory_pv", "<f8"), ("pv", "<f8")])


class LaunchConfig(C.Structure):
    _fields_ = [("paths_per_thread", C.c_int32), ("use_graph", C.c_int32),
                ("owen_table", C.c_int32), ("normal_table", C.c_int32),
                ("stream_ordered_alloc", C.c_int32), ("waves_per_block", C.c_int32),
                ("serial_launches", C.c_int32), ("fp32_full_ladder", C.c_int32),
                ("stack_split", C.c_int32), ("persistent", C.c_int32)]


class PlanInfo(C.Structure):
    _fields_ = [("n_requests", C.c_int32), ("n_items", C.c_int32), ("n_launches", C.c_int32),
                ("blocks", C.c_int32), ("path_steps", C.c_double), ("lds_bytes", C.c_int64),
                ("device_bytes", C.c_int64), ("graph", C.c_int32), ("normal_tables", C.c_int32),
                ("paths_per_thread", C.c_int32), ("waves_per_block", C.c_int32),
                ("owen_tables", C.c_int32), ("reserved", C.c_int32)]


class TuneTrial(C.Structure):
    _fields_ = [("paths_per_thread", C.c_int32), ("waves_per_block", C.c_int32),
                ("device_ms", C.c_double)]


class TuneReport(C.Structure):
    _fields_ = [("n_trials", C.c_int32), ("best", C.c_int32), ("trials", TuneTrial * 16)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_inflight", C.c_int32),
                ("max_batch_requests", C.c_int32), ("max_queue_seconds", C.c_double),
                ("paths_per_thread", C.c_int32), ("normal_table", C.c_int32),
                ("max_batch_path_steps", C.c_double)]


class TicketInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("n_requests", C.c_int32), ("batch_tickets", C.c_int32),
                ("reserved", C.c_int32), ("submit_s", C.c_double), ("start_s", C.c_double),
                ("end_s", C.c_double), ("latency_s", C.c_double), ("device_bytes", C.c_int64),
                ("path_steps", C.c_double), ("device_start_s", C.c_double),
                ("device_end_s", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted_tickets", C.c_int64), ("completed_tickets", C.c_int64),
                ("completed_requests", C.c_int64), ("shed_tickets", C.c_int64),
                ("failed_tickets", C.c_int64), ("batches", C.c_int64), ("peak_bytes", C.c_int64),
                ("bytes_in_use", C.c_int64), ("path_steps", C.c_double),
                ("queued_tickets", C.c_int32), ("running_plans", C.c_int32),
                ("max_concurrent", C.c_int32), ("overlapped_plans", C.c_int32)]


EXPORTS = ("riga_hip_init", "riga_hip_destroy", "riga_hip_device_name", "riga_hip_arch",
           "riga_validate", "riga_price_put_leg", "riga_price_put_leg_noalloc", "riga_price_batch",
           "riga_replay_paths",
           "riga_plan_create", "riga_plan_estimate", "riga_plan_launch", "riga_plan_ready",
           "riga_plan_wait", "riga_plan_run", "riga_plan_query", "riga_plan_device_times",
           "riga_plan_destroy",
           "riga_server_create", "riga_server_submit", "riga_server_wait",
           "riga_server_stats_get", "riga_server_destroy",
           "riga_surface_create", "riga_surface_destroy", "riga_autotune",
           "riga_lv_calibrate", "riga_lv_sigma", "riga_debug_sobol", "riga_debug_sobol_f32",
           "riga_debug_halton", "riga_debug_exp_log", "riga_debug_mc_starts",
           "riga_replay_ledger", "riga_query_worst_percentile", "riga_path_cube",
           "riga_worst_of", "riga_resident_worst_of_create", "riga_resident_worst_of_reduce",
           "riga_resident_worst_of_download", "riga_resident_worst_of_synchronize",
           "riga_resident_worst_of_destroy", "riga_lv_terminal_spots", "riga_resident_lv_create",
           "riga_resident_lv_run", "riga_resident_lv_download", "riga_resident_lv_synchronize",
           "riga_resident_lv_destroy")

_lib = None
_lock = threading.Lock()


def build() -> str:
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "csrc")], check=True)
    return LIB_PATH


def lib():
    """Load the in-tree library (no implicit rebuild, no fallback)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RigaError("gpu_unavailable",
                                f"{LIB_PATH} is not built; run riga_flow_amd/csrc/Makefile")
            L = C.CDLL(LIB_PATH)
            L.riga_hip_init.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.POINTER(Error)]
            L.riga_hip_destroy.argtypes = [C.c_void_p]
            L.riga_hip_destroy.restype = None
            L.riga_hip_device_name.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(Error)]
            L.riga_hip_arch.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(Error)]
            L.riga_validate.argtypes = [C.POINTER(Request), C.POINTER(Error)]
            L.riga_price_put_leg.argtypes = [C.c_void_p, C.POINTER(Request), C.c_int,
                                             C.POINTER(Result), C.POINTER(Error)]
            L.riga_price_put_leg_noalloc.argtypes = [C.c_void_p, C.POINTER(Request), C.c_int,
                                                     C.POINTER(Result), C.POINTER(Error)]
            L.riga_price_batch.argtypes = [C.c_void_p, C.POINTER(Request), C.c_int, C.c_int,
                                           C.POINTER(LaunchConfig), C.POINTER(Result),
                                           C.POINTER(Error)]
            L.riga_replay_paths.argtypes = [C.c_void_p, C.POINTER(Request), C.c_int,
                                            C.POINTER(PathRecord), C.POINTER(Result),
                                            C.POINTER(Error)]
            L.riga_plan_create.argtypes = [C.c_void_p, C.POINTER(Request), C.c_int, C.c_int,
                                           C.POINTER(LaunchConfig), C.POINTER(C.c_void_p),
                                           C.POINTER(Error)]
            L.riga_plan_estimate.argtypes = [C.c_void_p, C.POINTER(Request), C.c_int, C.c_int,
                                             C.POINTER(LaunchConfig), C.POINTER(C.c_int64),
                                             C.POINTER(Error)]
            L.riga_plan_ready.argtypes = [C.c_void_p]
            L.riga_server_create.argtypes = [C.c_void_p, C.POINTER(ServerConfig),
                                             C.POINTER(C.c_void_p), C.POINTER(Error)]
            L.riga_server_submit.argtypes = [C.c_void_p, C.POINTER(Request), C.c_int, C.c_int,
                                             C.POINTER(C.c_uint64), C.POINTER(Error)]
            L.riga_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(Result),
                                           C.POINTER(TicketInfo), C.POINTER(Error)]
            L.riga_server_stats_get.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
            L.riga_server_destroy.argtypes = [C.c_void_p]
            L.riga_server_destroy.restype = None
            L.riga_plan_launch.argtypes = [C.c_void_p, C.POINTER(Error)]
            L.riga_plan_wait.argtypes = [C.c_void_p, C.POINTER(Result), C.POINTER(Error)]
            L.riga_plan_run.argtypes = [C.c_void_p, C.POINTER(Result), C.POINTER(Error)]
            L.riga_plan_query.argtypes = [C.c_void_p, C.POINTER(PlanInfo)]
            L.riga_plan_device_times.argtypes = [C.c_void_p, C.POINTER(C.c_double),
                                                 C.POINTER(C.c_double)]
            L.riga_plan_destroy.argtypes = [C.c_void_p]
            L.riga_plan_destroy.restype = None
            L.riga_lv_calibrate.argtypes = [C.POINTER(ImpliedGrid), C.c_double, C.c_double,
                                            C.c_int32, _D, _D, _D, C.POINTER(Error)]
            L.riga_surface_create.argtypes = [C.c_void_p, C.POINTER(LvGrid), C.POINTER(C.c_void_p),
                                              C.POINTER(Error)]
            L.riga_autotune.argtypes = [C.c_void_p, C.POINTER(Request), C.c_int, C.c_int,
                                        C.POINTER(LaunchConfig), C.c_int, C.POINTER(TuneReport),
                                        C.POINTER(Error)]
            L.riga_surface_destroy.argtypes = [C.c_void_p]
            L.riga_surface_destroy.restype = None
            L.riga_lv_sigma.argtypes = [C.POINTER(LvGrid), C.c_double, C.c_double]
            L.riga_lv_sigma.restype = C.c_double
            L.riga_debug_sobol.argtypes = [C.c_void_p, C.c_uint64, C.c_int32, C.c_uint64,
                                           C.c_int32, C.c_int32, C.c_int32,
                                           C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _D,
                                           C.POINTER(Error)]
            L.riga_debug_sobol_f32.argtypes = [C.c_void_p, C.c_uint64, C.c_int32, C.c_uint64,
                                               C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                               C.POINTER(C.c_uint32), C.POINTER(C.c_float),
                                               C.POINTER(Error)]
            L.riga_debug_halton.argtypes = [C.c_void_p, C.c_uint64, C.c_int32, C.c_uint64,
                                            C.c_int32, C.c_int32, C.c_int32,
                                            C.POINTER(C.c_uint32), _D, _D, C.POINTER(Error)]
            L.riga_debug_exp_log.argtypes = [C.c_void_p, _D, C.c_uint64, _D, _D, C.POINTER(Error)]
            L.riga_debug_mc_starts.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64,
                                               C.POINTER(C.c_uint64), C.POINTER(Error)]
            _U64 = C.POINTER(C.c_uint64)
            L.riga_replay_ledger.argtypes = [C.c_void_p, C.POINTER(Request), C.c_int, _U64, C.c_uint64,
                                             C.POINTER(PathRecord), C.POINTER(CouponRow), C.POINTER(Error)]
            L.riga_query_worst_percentile.argtypes = [C.c_void_p, C.POINTER(Request), C.c_int, C.c_double,
                                                      C.c_uint64, C.POINTER(WorstQuery), _U64,
                                                      C.POINTER(PathRecord), C.POINTER(CouponRow),
                                                      C.POINTER(Error)]
            _F32 = C.POINTER(C.c_float)
            _I32 = C.POINTER(C.c_int32)
            _VP = C.c_void_p
            L.riga_worst_of.argtypes = [_VP, _F32, _F32, C.c_uint64, _F32, C.POINTER(Error)]
            L.riga_resident_worst_of_create.argtypes = [_VP, _F32, _F32, C.c_uint64, C.POINTER(_VP),
                                                        C.POINTER(Error)]
            for name in ("riga_resident_worst_of_reduce", "riga_resident_worst_of_synchronize",
                         "riga_resident_lv_synchronize"):
                getattr(L, name).argtypes = [_VP, C.POINTER(Error)]
            L.riga_resident_worst_of_download.argtypes = [_VP, _F32, C.POINTER(Error)]
            L.riga_resident_worst_of_destroy.argtypes = [_VP]
            L.riga_resident_worst_of_destroy.restype = None
            L.riga_lv_terminal_spots.argtypes = [_VP, _F32, _F32, _F32, C.c_int32, _F32, _F32, C.c_float,
                                                 C.c_int64, C.c_int64, C.c_float, C.c_float, _F32, _F32, _I32,
                                                 C.POINTER(Error)]
            L.riga_resident_lv_create.argtypes = [_VP, _F32, _F32, _F32, C.c_int32, C.c_int64, C.c_int64,
                                                  C.c_float, C.POINTER(_VP), C.POINTER(Error)]
            L.riga_resident_lv_run.argtypes = [_VP, _F32, _F32, C.c_float, C.c_float, _F32, C.POINTER(Error)]
            L.riga_resident_lv_download.argtypes = [_VP, _F32, _I32, C.POINTER(Error)]
            L.riga_resident_lv_destroy.argtypes = [_VP]
            L.riga_resident_lv_destroy.restype = None
            L.riga_path_cube.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int32, C.c_int32, _D, _D,
                                         _D, C.c_double, _D, C.c_double, _F32, _F32, _F32,
                                         C.POINTER(Error)]
            _lib = L
    return _lib


def check(rc: int, err: Error) -> None:
    if rc != 0:
        raise RigaError(err.code.decode() or "gpu_error", err.message.decode())


class Device:
    """The process-global device (``gpu.rs`` ``OnceLock<Gpu>``)."""

    _instance = None
    _inst_lock = threading.Lock()

    def __init__(self, index: int = 0):
        L = lib()
        h = C.c_void_p()
        err = Error()
        check(L.riga_hip_init(index, C.byref(h), C.byref(err)), err)
        self.handle = h

    @classmethod
    def get(cls) -> "Device":
        with cls._inst_lock:
            if cls._instance is None:
                cls._instance = Device(0)
            return cls._instance

    def name(self) -> str:
        buf = C.create_string_buffer(256)
        err = Error()
        check(lib().riga_hip_device_name(self.handle, buf, 256, C.byref(err)), err)
        return buf.value.decode()

    def arch(self) -> str:
        buf = C.create_string_buffer(64)
        err = Error()
        check(lib().riga_hip_arch(self.handle, buf, 64, C.byref(err)), err)
        return buf.value.decode()


def darr(values) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(values, dtype=np.float64))


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)
