"""Python mirror of the resizing scheduler's host protocol (CL_Communicator,
src/scheduler_rt/rt_host/cl_communicator.h:19-465) and the merged / solo run
modes of the driver (src/host_skeleton/skel.cpp:168-631), over libwgm."""
import ctypes
import time

import numpy as np

from . import _lib as L

NP_MATMULT_I32, NP_MATMUL_F32, NP_REDUCE_MIN = 0, 1, 2
PERSIST_BFS, PERSIST_SSSP, PERSIST_COLOR, PERSIST_MIS, PERSIST_PAGERANK = 0, 1, 2, 3, 4
_RESULT_DTYPE = {PERSIST_BFS: np.uint32, PERSIST_SSSP: np.uint32, PERSIST_COLOR: np.int32, PERSIST_MIS: np.int32,
                 PERSIST_PAGERANK: np.float32}


def np_config(kind=NP_MATMUL_F32, dim=256, length=1 << 20, seed=0):
    c = L.NpConfig()
    c.kind, c.dim, c.length, c.seed = kind, dim, length, seed
    return c


class Mega:
    """A resident mega-kernel: scheduler workgroup + workers."""

    def __init__(self, rt, np_kind=NP_MATMUL_F32, dim=256, length=1 << 20, seed=0):
        self.rt = rt
        self._h = ctypes.c_void_p()
        cfg = np_config(np_kind, dim, length, seed)
        rt._check(L.lib().wgm_mega_create(rt.handle, ctypes.byref(cfg), ctypes.byref(self._h)))
        self.participating = 0

    def close(self):
        if self._h:
            L.lib().wgm_mega_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        # always drain the resident kernel, also when the block raised
        if self._h:
            L.lib().wgm_mega_quit(self._h)
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def launch(self, wg_size=256, num_wgs=0, query_barrier=False):
        p = ctypes.c_int32()
        self.rt._check(L.lib().wgm_mega_launch(self._h, wg_size, num_wgs, 1 if query_barrier else 0, ctypes.byref(p)))
        self.participating = p.value
        return p.value

    def send_persistent(self, graph, kind=PERSIST_BFS, src=0, groups=None, damping=0.85, iterations=50):
        groups = self.participating if groups is None else groups
        t = L.PersistentTask()
        t.kind, t.src, t.damping, t.iterations = kind, src, damping, iterations
        self._kind = kind
        self.rt._check(L.lib().wgm_mega_send_persistent_task(self._h, graph._h, ctypes.byref(t), groups))

    def persistent_running(self):
        return bool(L.lib().wgm_mega_persistent_running(self._h))

    def wait_persistent(self, timeout_s=60.0):
        ms = ctypes.c_double()
        self.rt._check(L.lib().wgm_mega_wait_persistent(self._h, timeout_s, ctypes.byref(ms)))
        return ms.value

    def send_task(self, groups):
        t = L.TaskTimes()
        rc = L.lib().wgm_mega_send_task(self._h, groups, ctypes.byref(t))
        if rc not in (L.WGM_OK, -6):
            self.rt._check(rc)
        return t.as_dict()

    def quit(self):
        self.rt._check(L.lib().wgm_mega_quit(self._h))

    def stats(self):
        s = L.MegaStats()
        self.rt._check(L.lib().wgm_mega_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def result(self, n):
        """(output, levels / rounds / iterations) of the last persistent task:
        distances, colours, MIS membership or ranks."""
        out = np.empty(n, _RESULT_DTYPE[getattr(self, "_kind", PERSIST_BFS)])
        lv = ctypes.c_int32()
        self.rt._check(L.lib().wgm_mega_result(self._h, L.ptr(out), ctypes.byref(lv)))
        return out, lv.value

    def launch_custom(self, kernel, args=None, wg_size=256, num_wgs=0, lds_bytes=0):
        """Launch the caller's merged mega-kernel (first parameter wgm_mega_args)."""
        p = ctypes.c_int32()
        self.rt._check(L.lib().wgm_mega_launch_custom(self._h, ctypes.c_void_p(kernel), wg_size, num_wgs, lds_bytes,
                                                      args.handle if args is not None else None, ctypes.byref(p)))
        self.participating = p.value
        return p.value

    def send_persistent_custom(self, groups=None):
        groups = self.participating if groups is None else groups
        self.rt._check(L.lib().wgm_mega_send_persistent_custom(self._h, groups))

    def send_task_custom(self, groups):
        t = L.TaskTimes()
        self.rt._check(L.lib().wgm_mega_send_task_custom(self._h, groups, ctypes.byref(t)))
        return t.as_dict()

    def run_solo(self, wg_size=256, num_wgs=8, iterations=10):
        us, errs = ctypes.c_double(), ctypes.c_int32()
        self.rt._check(L.lib().wgm_np_run_solo(self._h, wg_size, num_wgs, iterations, ctypes.byref(us), ctypes.byref(errs)))
        return us.value, errs.value

    def np_copy(self, kind, which, dim=256, length=1 << 20):
        """A task's resident buffer (NP_IN_A / NP_IN_B / NP_OUT_C) as a numpy
        array: int32 or float32 dim x dim, or the reduction's int32 input."""
        if kind == NP_REDUCE_MIN:
            out = np.empty(length, np.int32)
        else:
            out = np.empty((dim, dim), np.float32 if kind == NP_MATMUL_F32 else np.int32)
        self.rt._check(L.lib().wgm_np_copy(self._h, kind, which, L.ptr(out), out.nbytes))
        return out

    def np_result(self):
        """(value, checksum) the last task published."""
        v, c = ctypes.c_int32(), ctypes.c_double()
        self.rt._check(L.lib().wgm_np_result(self._h, ctypes.byref(v), ctypes.byref(c)))
        return v.value, c.value

    def np_expected(self, kind):
        """(value, checksum, tolerance) the library checks a task of `kind` against."""
        v, c, t = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
        self.rt._check(L.lib().wgm_np_expected(self._h, kind, ctypes.byref(v), ctypes.byref(c), ctypes.byref(t)))
        return v.value, c.value, t.value


NP_IN_A, NP_IN_B, NP_OUT_C = 0, 1, 2


def np_groups_for(ratio, occupancy):
    """get_workgroups_for_non_persistent (src/host_skeleton/skel.cpp:387-401):
    -1 -> occupancy - 2, -2 -> 1, k -> occupancy // k."""
    if ratio == -1:
        return max(1, occupancy - 2)
    if ratio == -2:
        return 1
    return max(1, occupancy // ratio)


def run_merged(mega, graph, kind=PERSIST_BFS, src=0, frequency_ms=1.0, np_ratio=4, skip_tasks=False,
               wg_size=256, num_wgs=0, query_barrier=False, max_tasks=100000):
    """execute_merged_iteration (src/host_skeleton/skel.cpp:405-463): launch,
    send the persistent task to every workgroup, then send a non-persistent
    task every frequency_ms until the persistent task ends; quit."""
    occ = mega.launch(wg_size, num_wgs, query_barrier)
    mega.send_persistent(graph, kind, src, occ)
    k = np_groups_for(np_ratio, occ)
    errors = 0
    redlines = 0
    if not skip_tasks:
        begin = time.perf_counter()
        sent = 0
        while mega.persistent_running() and sent < max_tasks:
            now = time.perf_counter()
            if (now - begin) * 1e3 >= frequency_ms:
                redlines += 1
            else:
                while (time.perf_counter() - begin) * 1e3 < frequency_ms:
                    pass
            r = mega.send_task(min(k, occ - 1))
            errors += 0 if r["ok"] else 1
            sent += 1
            begin = time.perf_counter()
    ms = mega.wait_persistent()
    mega.quit()
    st = mega.stats()
    st["np_errors"] = errors
    st["redlines_host"] = redlines
    st["persistent_ms"] = ms
    return st
