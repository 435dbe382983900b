"""ctypes binding of libwgm.so (the C ABI declared in include/wgm.h).

The library is built in-tree by __graft_entry__.build() (make -C
gpu_workgroup_managment_amd/csrc). There is no fallback: if the library is
missing, importing the package raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwgm.so")

WGM_OK = 0
ERRORS = {
    -1: "WGM_ERR_INVALID",
    -2: "WGM_ERR_NOMEM",
    -3: "WGM_ERR_HIP",
    -4: "WGM_ERR_NODEV",
    -5: "WGM_ERR_TOO_MANY_GROUPS",
    -6: "WGM_ERR_MISMATCH",
    -7: "WGM_ERR_TIMEOUT",
    -8: "WGM_ERR_STATE",
    -9: "WGM_ERR_BUDGET",
    -10: "WGM_ERR_IO",
    -11: "WGM_ERR_FORMAT",
}
WGM_ERR_MISMATCH, WGM_ERR_IO, WGM_ERR_FORMAT = -6, -10, -11

# file formats (WGM_FMT_*, WGM_READ_*, WGM_SOL_*)
FMT_GR, FMT_DIMACS, FMT_METIS, FMT_MM, FMT_EDGES = range(5)
READ_UNDIRECTED, READ_TRANSPOSE = 1, 2
SOL_COLON, SOL_COLON_SPACE, SOL_PLAIN = range(3)

# kernel ids (WGM_K_*)
(K_DISCOVERY, K_BARRIER_BENCH, K_BFS, K_SSSP, K_COLOR, K_MIS, K_PAGERANK, K_MEGA, K_BATCHED, K_SERVE, K_MINPLUS,
 K_BC) = range(12)
BIG_NUM = 99999999  # WGM_BIG_NUM: Pannotia SSSP's unreached distance
GRAPH_RMAT, GRAPH_GRID = 0, 1
QUERY_BFS, QUERY_SSSP = 0, 1
BARRIER_FLAGS, BARRIER_SENSE, BARRIER_XCD, BARRIER_COOP = 0, 1, 2, 3
SSSP_AUTO, SSSP_PUSH, SSSP_PULL = 0, 1, 2


class WgmError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")


class Csr(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("m", ctypes.c_int64),
        ("row_ptr", ctypes.POINTER(ctypes.c_int32)),
        ("col", ctypes.POINTER(ctypes.c_int32)),
        ("wt", ctypes.POINTER(ctypes.c_int32)),
    ]


class BatchItem(ctypes.Structure):
    _fields_ = [
        ("index", ctypes.c_int32),
        ("kind", ctypes.c_int32),
        ("query", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("scale", ctypes.c_int32),
        ("rows", ctypes.c_int32),
        ("cols", ctypes.c_int32),
        ("graph_seed", ctypes.c_uint64),
        ("weight_seed", ctypes.c_uint64),
    ]


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("arch", ctypes.c_char * 32),
        ("num_cus", ctypes.c_int32),
        ("max_wg_size", ctypes.c_int32),
        ("wavefront_size", ctypes.c_int32),
        ("wall_clock_hz", ctypes.c_int64),
        ("total_mem", ctypes.c_int64),
        ("lds_per_cu", ctypes.c_int32),
    ]


class BarrierResult(ctypes.Structure):
    _fields_ = [
        ("participating", ctypes.c_int32),
        ("barriers", ctypes.c_int32),
        ("epoch_errors", ctypes.c_int32),
        ("kernel_ms", ctypes.c_double),
        ("device_ms", ctypes.c_double),
        ("ns_per_barrier", ctypes.c_double),
    ]


class Launch(ctypes.Structure):
    _fields_ = [
        ("wg_size", ctypes.c_int32),
        ("num_wgs", ctypes.c_int32),
        ("handshake", ctypes.c_int32),
        ("mode", ctypes.c_int32),
    ]


class RunStats(ctypes.Structure):
    _fields_ = [
        ("participating", ctypes.c_int32),
        ("iterations", ctypes.c_int32),
        ("edges_scanned", ctypes.c_int64),
        ("kernel_ms", ctypes.c_double),
        ("device_ms", ctypes.c_double),
        ("host_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchConfig(ctypes.Structure):
    _fields_ = [
        ("wg_size", ctypes.c_int32),
        ("wgs_per_cu", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
        ("team_edges", ctypes.c_int32),
        ("max_team", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("delta", ctypes.c_int32),
        ("sssp_team_edges", ctypes.c_int32),
        ("delta_high_degree", ctypes.c_int32),
        ("far_flags_lds", ctypes.c_int32),
        ("low_ell", ctypes.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("wall_ms", ctypes.c_double),
        ("device_ms", ctypes.c_double),
        ("gteps", ctypes.c_double),
        ("edges_scanned", ctypes.c_int64),
        ("participating", ctypes.c_int32),
        ("launches", ctypes.c_int32),
        ("grid", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("edges_relaxed", ctypes.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class NpConfig(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("dim", ctypes.c_int32), ("length", ctypes.c_int32), ("seed", ctypes.c_int32)]


class TaskTimes(ctypes.Structure):
    _fields_ = [("response_us", ctypes.c_double), ("execution_us", ctypes.c_double), ("ok", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("dev_assign_us", ctypes.c_double), ("dev_start_us", ctypes.c_double),
                ("dev_body_us", ctypes.c_double), ("dev_retire_us", ctypes.c_double)]

    def as_dict(self):
        return {f: (bool(getattr(self, f)) if f == "ok" else getattr(self, f)) for f, _ in self._fields_
                if f != "reserved"}


class MegaStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int32) for f in ("participating", "tasks", "np_groups", "errors", "redlines", "reserved")] + [
        (f, ctypes.c_double) for f in ("persistent_ms", "response_mean_us", "response_p50_us", "response_p99_us",
                                       "response_max_us", "exec_mean_us", "exec_p50_us", "exec_p99_us", "exec_max_us",
                                       "total_mean_us", "total_p50_us", "total_p99_us", "avg_persistent_groups")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class ServeConfig(ctypes.Structure):
    _fields_ = [("wg_size", ctypes.c_int32), ("num_wgs", ctypes.c_int32), ("max_inflight", ctypes.c_int32),
                ("reserve", ctypes.c_int32), ("budget_bytes", ctypes.c_int64), ("np", NpConfig)]


class ServeDone(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("graph", ctypes.c_int32), ("query", ctypes.c_int32),
                ("src", ctypes.c_int32), ("levels", ctypes.c_int32), ("reached", ctypes.c_int64),
                ("edges_scanned", ctypes.c_int64), ("digest", ctypes.c_uint64), ("queue_us", ctypes.c_double),
                ("service_us", ctypes.c_double), ("device_us", ctypes.c_double), ("preempted", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class ServeLoad(ctypes.Structure):
    _fields_ = [("rate_per_s", ctypes.c_double), ("requests", ctypes.c_int32), ("np_groups", ctypes.c_int32),
                ("np_fraction", ctypes.c_double), ("seed", ctypes.c_uint64)]


class ServeReport(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int32) for f in ("queries", "tasks", "task_errors", "reserved")] + [
        (f, ctypes.c_double) for f in ("wall_s", "throughput_rps", "gteps", "query_p50_us", "query_p99_us",
                                       "query_mean_us", "task_p50_us", "task_p99_us", "task_mean_us",
                                       "task_queue_p99_us", "task_response_p99_us", "task_exec_p99_us",
                                       "max_inflight_bytes")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class PersistentTask(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("src", ctypes.c_int32), ("damping", ctypes.c_float),
                ("iterations", ctypes.c_int32)]


class PersistentApp(ctypes.Structure):
    """wgm_persistent_app: the plugin's state and its nine entry points
    (function pointers, filled by the plugin's own C code)."""
    _fields_ = [("self", ctypes.c_void_p)] + [(f, ctypes.c_void_p) for f in (
        "name", "kernel", "init_for_occupancy", "init_for_real", "set_args_for_occupancy", "set_args_for_real",
        "reset", "check", "clean")]


class NpTask(ctypes.Structure):
    """wgm_np_task: the non-persistent task plugin's state and entry points."""
    _fields_ = [("self", ctypes.c_void_p)] + [(f, ctypes.c_void_p) for f in (
        "name", "kernel", "init", "set_args", "reset", "check", "clean")]


class SkelConfig(ctypes.Structure):
    _fields_ = [("threads_per_wg", ctypes.c_int32), ("num_wgs", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("non_persistent_wgs", ctypes.c_int32), ("non_persistent_frequency_ms", ctypes.c_double),
                ("skip_tasks", ctypes.c_int32), ("use_query_barrier", ctypes.c_int32), ("lds_bytes", ctypes.c_int32),
                ("max_tasks", ctypes.c_int32)]


class SkelReport(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int32) for f in ("occupancy", "num_wgs", "np_wgs", "iterations", "persistent_errors",
                                              "np_errors", "tasks", "redlines")] + [
        (f, ctypes.c_double) for f in ("total_ms", "mean_ms", "min_ms", "max_ms", "std_ms", "response_p50_us",
                                       "response_p99_us", "exec_p50_us", "exec_p99_us")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


EXTRA_STRUCTS = {
    "wgm_persistent_app": PersistentApp,
    "wgm_np_task": NpTask,
    "wgm_skel_config": SkelConfig,
    "wgm_skel_report": SkelReport,
    "wgm_persistent_task": PersistentTask,
    "wgm_serve_config": ServeConfig,
    "wgm_serve_done": ServeDone,
    "wgm_serve_load": ServeLoad,
    "wgm_serve_report": ServeReport,
}

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_P = ctypes.POINTER

# name -> (restype, argtypes); every symbol here is declared in include/wgm.h
SIGNATURES = {
    "wgm_splitmix64_next": (_u64, [_P(_u64)]),
    "wgm_gen_er": (_i32, [_i32, _i64, _u64, _P(Csr)]),
    "wgm_gen_rmat": (_i32, [_i32, _i32, ctypes.c_double, ctypes.c_double, ctypes.c_double, _u64, _i32, _P(Csr)]),
    "wgm_gen_grid": (_i32, [_i32, _i32, _P(Csr)]),
    "wgm_gen_weights": (_i32, [_P(Csr), _i32, _i32, _u64]),
    "wgm_gen_priorities": (_i32, [_P(Csr), _u64, _vp]),
    "wgm_csr_free": (None, [_P(Csr)]),
    "wgm_read_graph": (_i32, [ctypes.c_char_p, _i32, _i32, _P(Csr)]),
    "wgm_write_graph": (_i32, [ctypes.c_char_p, _i32, _i32, _P(Csr)]),
    "wgm_format_error": (ctypes.c_char_p, []),
    "wgm_write_solution": (_i32, [ctypes.c_char_p, _i32, _vp, _i32]),
    "wgm_check_solution": (_i32, [ctypes.c_char_p, _i32, _vp, _i32, _P(_i32)]),
    "wgm_write_solution_f32": (_i32, [ctypes.c_char_p, _vp, _i32]),
    "wgm_check_solution_f32": (_i32, [ctypes.c_char_p, _vp, _i32, ctypes.c_float, _P(_i32)]),
    "wgm_free": (None, [_vp]),
    "wgm_batch_spec": (_i32, [_i32, _u64, _i32, _i32, _P(BatchItem)]),
    "wgm_batch_generate": (_i32, [_i32, _P(BatchItem), _P(Csr), _i32]),
    "wgm_create": (_i32, [_i32, _P(_vp)]),
    "wgm_destroy": (_i32, [_vp]),
    "wgm_last_error": (ctypes.c_char_p, [_vp]),
    "wgm_strerror": (ctypes.c_char_p, [_i32]),
    "wgm_get_device_info": (_i32, [_vp, _P(DeviceInfo)]),
    "wgm_occupancy_query_hip": (_i32, [_vp, _i32, _i32, _P(_i32), _P(_i32)]),
    "wgm_discover_occupancy": (_i32, [_vp, _i32, _i32, _i32, _P(_i32)]),
    "wgm_barrier_bench": (_i32, [_vp, _i32, _i32, _i32, _i32, _P(BarrierResult)]),
    "wgm_copy_bench": (_i32, [_vp, _i64, _i32, _P(ctypes.c_double), _P(ctypes.c_double)]),
    "wgm_graph_create": (_i32, [_vp, _P(Csr), _vp, _P(_vp)]),
    "wgm_graph_destroy": (_i32, [_vp]),
    "wgm_graph_bytes": (_i64, [_vp]),
    "wgm_bfs": (_i32, [_vp, _vp, _i32, _P(Launch), _vp, _P(RunStats)]),
    "wgm_sssp": (_i32, [_vp, _vp, _i32, _P(Launch), _vp, _P(RunStats)]),
    "wgm_color": (_i32, [_vp, _vp, _P(Launch), _vp, _P(RunStats)]),
    "wgm_mis": (_i32, [_vp, _vp, _P(Launch), _vp, _P(RunStats)]),
    "wgm_pagerank": (_i32, [_vp, _vp, ctypes.c_float, _i32, _P(Launch), _vp, _P(RunStats)]),
    "wgm_sssp_minplus": (_i32, [_vp, _vp, _i32, _P(Launch), _vp, _P(RunStats)]),
    "wgm_bfs_noalloc": (_i32, [_vp, _vp, _vp, _i32, _i64, _i32, _P(Launch), _vp, _P(RunStats)]),
    "wgm_bc": (_i32, [_vp, _vp, _i32, _P(Launch), _vp, _P(RunStats)]),
    "wgm_mega_create": (_i32, [_vp, _P(NpConfig), _P(_vp)]),
    "wgm_mega_destroy": (_i32, [_vp]),
    "wgm_mega_launch": (_i32, [_vp, _i32, _i32, _i32, _P(_i32)]),
    "wgm_mega_send_persistent": (_i32, [_vp, _vp, _i32, _i32, _i32]),
    "wgm_mega_send_persistent_task": (_i32, [_vp, _vp, _P(PersistentTask), _i32]),
    "wgm_mega_persistent_running": (_i32, [_vp]),
    "wgm_mega_wait_persistent": (_i32, [_vp, ctypes.c_double, _P(ctypes.c_double)]),
    "wgm_mega_send_task": (_i32, [_vp, _i32, _P(TaskTimes)]),
    "wgm_mega_send_task_kind": (_i32, [_vp, _i32, _i32, _P(TaskTimes)]),
    "wgm_mega_quit": (_i32, [_vp]),
    "wgm_mega_get_stats": (_i32, [_vp, _P(MegaStats)]),
    "wgm_mega_result": (_i32, [_vp, _vp, _P(_i32)]),
    "wgm_np_run_solo": (_i32, [_vp, _i32, _i32, _i32, _P(ctypes.c_double), _P(_i32)]),
    "wgm_np_copy": (_i32, [_vp, _i32, _i32, _vp, _i64]),
    "wgm_np_result": (_i32, [_vp, _P(_i32), _P(ctypes.c_double)]),
    "wgm_np_expected": (_i32, [_vp, _i32, _P(_i32), _P(ctypes.c_double), _P(ctypes.c_double)]),
    "wgm_args_create": (_i32, [_P(_vp)]),
    "wgm_args_push": (_i32, [_vp, _vp, _i32]),
    "wgm_args_destroy": (_i32, [_vp]),
    "wgm_launch_standalone": (_i32, [_vp, _vp, _i32, _i32, _i32, _vp, _i32, _P(RunStats)]),
    "wgm_mega_launch_custom": (_i32, [_vp, _vp, _i32, _i32, _i32, _vp, _P(_i32)]),
    "wgm_mega_send_persistent_custom": (_i32, [_vp, _i32]),
    "wgm_mega_send_task_custom": (_i32, [_vp, _i32, _P(TaskTimes)]),
    "wgm_discover_occupancy_custom": (_i32, [_vp, _vp, _i32, _i32, _vp, _i32, _P(_i32)]),
    "wgm_mega_set_query_barrier": (_i32, [_vp, _i32]),
    "wgm_run_persistent": (_i32, [_vp, _P(PersistentApp), _P(SkelConfig), _P(SkelReport)]),
    "wgm_run_merged": (_i32, [_vp, _P(PersistentApp), _P(NpTask), _vp, _P(SkelConfig), _P(SkelReport)]),
    "wgm_serve_create": (_i32, [_vp, _i32, _P(Csr), _P(ServeConfig), _P(_vp)]),
    "wgm_serve_destroy": (_i32, [_vp]),
    "wgm_serve_graph_bytes": (_i64, [_vp]),
    "wgm_serve_plan": (_i32, [_i32, _P(Csr), _P(ServeConfig), _P(_i64), _P(_i64)]),
    "wgm_serve_arena_bytes": (_i64, [_vp]),
    "wgm_serve_start": (_i32, [_vp, _P(_i32)]),
    "wgm_serve_submit": (_i32, [_vp, _i32, _i32, _i32, _P(_i64)]),
    "wgm_serve_poll": (_i32, [_vp, _P(ServeDone), _i32, _P(_i32)]),
    "wgm_serve_outstanding": (_i32, [_vp]),
    "wgm_serve_inflight_bytes": (_i64, [_vp]),
    "wgm_serve_task": (_i32, [_vp, _i32, _i32, _P(TaskTimes)]),
    "wgm_serve_stop": (_i32, [_vp]),
    "wgm_dist_digest": (_u64, [_vp, _i32]),
    "wgm_serve_run_load": (_i32, [_vp, _P(ServeLoad), _P(ServeReport), _P(ServeDone), _P(_i32)]),
    "wgm_batch_create": (_i32, [_vp, _i32, _P(Csr), _vp, _vp, _P(_vp)]),
    "wgm_batch_destroy": (_i32, [_vp]),
    "wgm_batch_bytes": (_i64, [_vp]),
    "wgm_batch_run": (_i32, [_vp, _vp, _P(BatchConfig), _P(BatchStats)]),
    "wgm_batch_run_relaunch": (_i32, [_vp, _vp, _P(BatchConfig), _P(BatchStats)]),
    "wgm_batch_run_standalone": (_i32, [_vp, _vp, _P(Launch), _P(BatchStats)]),
    "wgm_batch_result": (_i32, [_vp, _i32, _vp, _P(_i32), _P(_i64)]),
    "wgm_batch_job_times": (_i32, [_vp, _vp, _vp]),
    "wgm_batch_autotune": (_i32, [_vp, _vp, _i32, _P(BatchConfig), _P(BatchStats)]),
}

_lib = None


def lib():
    """Load libwgm.so (raises ImportError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, rt=None):
    if rc != WGM_OK:
        msg = lib().wgm_last_error(rt)
        raise WgmError(rc, msg.decode() if msg else "")
    return rc


def ptr(a):
    """Raw address of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def i32ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


class Args:
    """wgm_args: kernel arguments in order, each a ctypes scalar or pointer
    (copied into the builder when pushed)."""

    def __init__(self, values=()):
        self._h = ctypes.c_void_p()
        check(lib().wgm_args_create(ctypes.byref(self._h)))
        for v in values:
            self.push(v)

    def push(self, value):
        check(lib().wgm_args_push(self._h, ctypes.cast(ctypes.pointer(value), ctypes.c_void_p), ctypes.sizeof(value)))

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            if self._h:
                lib().wgm_args_destroy(self._h)
                self._h = ctypes.c_void_p()
        except Exception:
            pass
