"""Python mirror of the runtime host API (the reference's CL_Communicator /
skel.cpp run modes) over the C ABI of libwgm.so.

Reference interfaces mirrored: occupancy discovery
(src/host_skeleton/skel.cpp:223-247), run_persistent
(src/host_skeleton/skel.cpp:250-385), the os_freq barrier benchmark
(src/test_suite/os_freq/host/os_freq.cpp:262-329).
"""
import ctypes

import numpy as np

from . import _lib as L


class Runtime:
    """One cooperative-workgroup runtime on one GPU (not re-entrant)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        L.check(L.lib().wgm_create(device, ctypes.byref(self._h)))

    @property
    def handle(self):
        if not self._h:
            raise L.WgmError(-8, "runtime closed")
        return self._h

    def close(self):
        if self._h:
            L.lib().wgm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        return L.check(rc, self.handle)

    def device_info(self):
        info = L.DeviceInfo()
        self._check(L.lib().wgm_get_device_info(self.handle, ctypes.byref(info)))
        return {
            "name": info.name.decode(),
            "arch": info.arch.decode(),
            "num_cus": info.num_cus,
            "max_wg_size": info.max_wg_size,
            "wavefront_size": info.wavefront_size,
            "wall_clock_hz": info.wall_clock_hz,
            "total_mem": info.total_mem,
            "lds_per_cu": info.lds_per_cu,
        }

    def occupancy_query_hip(self, kernel, wg_size):
        """(blocks per CU, CU count) from the HIP occupancy API."""
        bpc, cus = ctypes.c_int32(), ctypes.c_int32()
        self._check(L.lib().wgm_occupancy_query_hip(self.handle, kernel, wg_size, ctypes.byref(bpc), ctypes.byref(cus)))
        return bpc.value, cus.value

    def discover_occupancy(self, kernel, wg_size, probe_groups=0):
        """Workgroups found co-resident by the discovery protocol."""
        cnt = ctypes.c_int32()
        self._check(L.lib().wgm_discover_occupancy(self.handle, kernel, wg_size, probe_groups, ctypes.byref(cnt)))
        return cnt.value

    def barrier_bench(self, wg_size=256, num_wgs=0, iterations=1000, variant=L.BARRIER_FLAGS):
        r = L.BarrierResult()
        self._check(L.lib().wgm_barrier_bench(self.handle, variant, wg_size, num_wgs, iterations, ctypes.byref(r)))
        return {f: getattr(r, f) for f, _ in r._fields_}

    def copy_bench(self, nbytes=1 << 30, iterations=20):
        """Plain copy kernel: (ms per copy, GB/s counting read + write)."""
        ms, gbps = ctypes.c_double(), ctypes.c_double()
        self._check(L.lib().wgm_copy_bench(self.handle, nbytes, iterations, ctypes.byref(ms), ctypes.byref(gbps)))
        return ms.value, gbps.value

    def graph(self, csr):
        return DeviceGraph(self, csr)

    def launch_standalone(self, kernel, args=None, wg_size=256, num_wgs=0, lds_bytes=0, handshake=False):
        """Run the caller's standalone persistent kernel (its host handle as an
        int / c_void_p; first parameter wgm_kernel_args) with `args` (L.Args)."""
        st = L.RunStats()
        self._check(L.lib().wgm_launch_standalone(self.handle, ctypes.c_void_p(kernel), wg_size, num_wgs, lds_bytes,
                                                  args.handle if args is not None else None, 1 if handshake else 0,
                                                  ctypes.byref(st)))
        return st.as_dict()


def _launch(wg_size=256, num_wgs=0, handshake=False, mode=0):
    lc = L.Launch()
    lc.wg_size = wg_size
    lc.num_wgs = num_wgs
    lc.handshake = 1 if handshake else 0
    lc.mode = mode
    return lc


class DeviceGraph:
    """A CSR resident in HBM plus its per-run state (kept between runs)."""

    def __init__(self, rt, csr):
        self.rt = rt
        self.csr = csr
        self._h = ctypes.c_void_p()
        c = csr.as_c()
        prio = L.ptr(csr.prio) if csr.prio is not None else None
        rt._check(L.lib().wgm_graph_create(rt.handle, ctypes.byref(c), prio, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            L.lib().wgm_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nbytes(self):
        return L.lib().wgm_graph_bytes(self._h)

    def _run(self, fn, *args, out=None, **lk):
        st = L.RunStats()
        lc = _launch(**lk)
        self.rt._check(fn(self.rt.handle, self._h, *args, ctypes.byref(lc), L.ptr(out), ctypes.byref(st)))
        return out, st.as_dict()

    def bfs(self, src=0, **lk):
        return self._run(L.lib().wgm_bfs, src, out=np.empty(self.csr.n, np.uint32), **lk)

    def sssp(self, src=0, **lk):
        return self._run(L.lib().wgm_sssp, src, out=np.empty(self.csr.n, np.uint32), **lk)

    def color(self, **lk):
        return self._run(L.lib().wgm_color, out=np.empty(self.csr.n, np.int32), **lk)

    def mis(self, **lk):
        return self._run(L.lib().wgm_mis, out=np.empty(self.csr.n, np.int32), **lk)

    def pagerank(self, damping=0.85, iterations=50, **lk):
        return self._run(L.lib().wgm_pagerank, ctypes.c_float(damping), iterations,
                         out=np.empty(self.csr.n, np.float32), **lk)

    def sssp_minplus(self, src=0, **lk):
        """Pannotia SSSP (min-plus sweeps over in-edges): int32 distances,
        L.BIG_NUM where unreached."""
        return self._run(L.lib().wgm_sssp_minplus, src, out=np.empty(self.csr.n, np.int32), **lk)

    def bc(self, num_sources=0, **lk):
        """Pannotia BC (Brandes, FP64) over sources [0, num_sources), all when 0."""
        return self._run(L.lib().wgm_bc, num_sources, out=np.empty(self.csr.n, np.float64), **lk)
