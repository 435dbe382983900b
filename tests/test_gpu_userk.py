"""GPU: the caller's own persistent kernels, written against
include/wgm_device_api.h with the reference's scheduler API names and built
separately (tests/kernels/user_kernels.hip), run through libwgm: a
standalone kernel with global barriers, and the merged form of
kernel_merge's example (a resizable persistent loop + the MY_reduce task)
driven by the resizing scheduler with tasks that take workgroups away."""
import ctypes
import os

import numpy as np
import pytest

from gpu_workgroup_managment_amd import _lib as L
from gpu_workgroup_managment_amd import scheduler as S

pytestmark = pytest.mark.gpu
os.environ.setdefault("WGM_HOST_TIMEOUT_S", "30")

@pytest.mark.parametrize("wg,nw", [(256, 0), (1024, 0), (64, 24)])
def test_standalone_user_kernel(rt, uk, dev, wg, nw):
    slots = dev(np.zeros(8192, np.int32))
    out = dev(np.zeros(4, np.int32))
    args = L.Args([slots.ptr, out.ptr, ctypes.c_int32(20)])
    st = rt.launch_standalone(uk.userk_rounds(), args, wg_size=wg, num_wgs=nw)
    o = out.get()
    print(wg, nw, st, o)
    n = st["participating"]
    assert n > 1 and (nw == 0 or n == nw)
    assert o[0] == 0, f"{o[0]} stale slots seen after global_barrier"
    assert o[1] == n
    assert o[2] == 4 * n  # get_global_id / get_global_size cover exactly the participants
    slots.free()
    out.free()


def test_merged_user_mega_kernel(rt, uk, dev):
    length = 1 << 20
    rng = np.random.default_rng(7)
    buf = rng.integers(1, 1 << 30, length).astype(np.int32)
    buf[123457] = 5
    d_buf = dev(buf)
    d_res = dev(np.full(1, 2**31 - 1, np.int32))
    d_calls = dev(np.zeros(1, np.int32))
    d_iters = dev(np.zeros(1, np.int32))
    target = 20000
    args = L.Args([ctypes.c_int32(length), d_buf.ptr, d_res.ptr, d_calls.ptr, d_iters.ptr, ctypes.c_int32(target)])
    with S.Mega(rt, np_kind=S.NP_REDUCE_MIN) as m:
        p = m.launch_custom(uk.userk_mega(), args, wg_size=256, num_wgs=16)
        assert p == 15
        m.send_persistent_custom(p)
        sizes = []
        while m.persistent_running() and len(sizes) < 30:
            k = 1 + len(sizes) % (p - 1)  # 1 .. p-1: kills and forks the persistent task's workgroups
            m.send_task_custom(k)
            sizes.append(k)
        m.wait_persistent(60)
        m.quit()
        st = m.stats()
    print(len(sizes), st)
    assert d_iters.get()[0] == target  # the loop survived every resize with its context
    assert d_res.get()[0] == 5  # MY_reduce's minimum
    assert d_calls.get()[0] == sum(sizes)  # every task ran on exactly k workgroups
    assert len(sizes) > 0
    for d in (d_buf, d_res, d_calls, d_iters):
        d.free()
