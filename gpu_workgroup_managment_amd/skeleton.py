"""Python mirror of the host skeleton (wgm_run_persistent / wgm_run_merged):
the reference's run_persistent and run_merged
(src/host_skeleton/skel.cpp:250-385, 405-631) driving a persistent-app plugin
and a non-persistent-task plugin through their vtables (wgm_persistent_app /
wgm_np_task, filled by the plugins' own C code)."""
import ctypes

from . import _lib as L


def config(threads_per_wg=128, num_wgs=0, iterations=1, non_persistent_wgs=-1, non_persistent_frequency_ms=0.0,
           skip_tasks=False, use_query_barrier=False, lds_bytes=0, max_tasks=0):
    """The reference's flags of these run modes (skel.cpp:21-44)."""
    c = L.SkelConfig()
    c.threads_per_wg, c.num_wgs, c.iterations = threads_per_wg, num_wgs, iterations
    c.non_persistent_wgs, c.non_persistent_frequency_ms = non_persistent_wgs, non_persistent_frequency_ms
    c.skip_tasks, c.use_query_barrier = int(skip_tasks), int(use_query_barrier)
    c.lds_bytes, c.max_tasks = lds_bytes, max_tasks
    return c


def run_persistent(rt, app, cfg=None):
    rep = L.SkelReport()
    cfg = cfg if cfg is not None else config()
    rt._check(L.lib().wgm_run_persistent(rt.handle, ctypes.byref(app), ctypes.byref(cfg), ctypes.byref(rep)))
    return rep.as_dict()


def run_merged(rt, app, np_task, merged_kernel, cfg=None):
    rep = L.SkelReport()
    cfg = cfg if cfg is not None else config()
    rt._check(L.lib().wgm_run_merged(rt.handle, ctypes.byref(app), ctypes.byref(np_task), ctypes.c_void_p(merged_kernel),
                                     ctypes.byref(cfg), ctypes.byref(rep)))
    return rep.as_dict()
