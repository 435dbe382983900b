"""GPU: kernels emitted by the HIP kernel_merge from stub-style sources
(tests/kernel_merge/*.hip -> gen/, built into libwgm_kmerge.so) and run
through libwgm: the standalone forms under wgm_launch_standalone, and the
merged mega-kernels under the resizing scheduler while tasks take
workgroups away (kills) and the persistent task takes them back (forks)."""
import ctypes
import os

import numpy as np
import pytest

from gpu_workgroup_managment_amd import _lib as L
from gpu_workgroup_managment_amd import scheduler as S

pytestmark = pytest.mark.gpu
os.environ.setdefault("WGM_HOST_TIMEOUT_S", "30")

HERE = os.path.dirname(os.path.abspath(__file__))
KERNELS = ("km_mega_barrier", "km_mega_chunks", "km_standalone_barrier", "km_standalone_chunks", "km_mega_octree",
           "km_standalone_octree")


@pytest.fixture(scope="module")
def km():
    path = os.path.join(HERE, "kernel_merge", "libwgm_kmerge.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for f in KERNELS:
        getattr(lib, f).restype = ctypes.c_void_p
    return lib


def chunk_expected(inp, mult):
    """chunk_loop's element function: 16 steps of x*1664525+1013904223, then x*mult+1 (mod 2^32)."""
    x = inp.astype(np.uint64)
    for _ in range(16):
        x = (x * 1664525 + 1013904223) & 0xFFFFFFFF
    return ((x * mult + 1) & 0xFFFFFFFF).astype(np.uint32)


class Chunks:
    def __init__(self, dev, n=1 << 20, chunk=1024, rounds=8, mult=2654435761, seed=3):
        rng = np.random.default_rng(seed)
        self.n, self.chunk, self.rounds, self.mult = n, chunk, rounds, mult
        self.inp = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        self.total = rounds * ((n + chunk - 1) // chunk)
        self.d_in, self.d_out = dev(self.inp), dev(np.zeros(n, np.uint32))
        self.d_hits, self.d_next = dev(np.zeros(self.total, np.int32)), dev(np.full(1, 12345, np.int32))

    def args(self):
        return [self.d_in.ptr, self.d_out.ptr, self.d_hits.ptr, self.d_next.ptr, ctypes.c_int32(self.n),
                ctypes.c_int32(self.chunk), ctypes.c_int32(self.rounds), ctypes.c_uint32(self.mult)]

    def check(self):
        out, hits = self.d_out.get(), self.d_hits.get()
        exp = chunk_expected(self.inp, self.mult)
        bad = np.flatnonzero(out != exp)
        assert bad.size == 0, f"{bad.size} wrong outputs, first at {bad[0]}: {out[bad[0]]} != {exp[bad[0]]}"
        assert np.all(hits == 1), f"chunks claimed != once: {np.unique(hits, return_counts=True)}"
        assert self.d_next.get()[0] >= self.total  # the initialisation reset the counter before any claim

    def free(self):
        for d in (self.d_in, self.d_out, self.d_hits, self.d_next):
            d.free()


def test_emitted_standalone_barrier(rt, km, dev):
    iters = dev(np.zeros(1, np.int32))
    st = rt.launch_standalone(km.km_standalone_barrier(), L.Args([iters.ptr, ctypes.c_int32(500)]), wg_size=256)
    print(st)
    assert st["participating"] > 1
    assert iters.get()[0] == 500
    iters.free()


@pytest.mark.parametrize("wg", [256, 1024])
def test_emitted_standalone_work_queue(rt, km, dev, wg):
    c = Chunks(dev)
    st = rt.launch_standalone(km.km_standalone_chunks(), L.Args(c.args()), wg_size=wg)
    print(wg, st)
    assert st["participating"] > 1
    c.check()
    c.free()


def reduce_buffers(dev, length=1 << 20):
    buf = np.random.default_rng(7).integers(1, 1 << 30, length).astype(np.int32)
    buf[654321] = 9
    return length, dev(buf), dev(np.full(1, 2**31 - 1, np.int32)), dev(np.zeros(1, np.int32))


def drive(m, p, max_tasks=30):
    """Send tasks of 1..p-1 workgroups while the persistent task runs (at least one)."""
    sizes = []
    while len(sizes) < max_tasks and (not sizes or m.persistent_running()):
        k = 1 + len(sizes) % (p - 1)
        m.send_task_custom(k)
        sizes.append(k)
    m.wait_persistent(60)
    return sizes


def test_emitted_mega_barrier(rt, km, dev):
    length, d_buf, d_res, d_calls = reduce_buffers(dev)
    d_iters = dev(np.zeros(1, np.int32))
    target = 20000
    args = L.Args([ctypes.c_int32(length), d_buf.ptr, d_res.ptr, d_calls.ptr, d_iters.ptr, ctypes.c_int32(target)])
    with S.Mega(rt, np_kind=S.NP_REDUCE_MIN) as m:
        p = m.launch_custom(km.km_mega_barrier(), args, wg_size=256, num_wgs=16)
        assert p == 15
        m.send_persistent_custom(p)
        sizes = drive(m, p)
        m.quit()
    print(len(sizes))
    assert d_iters.get()[0] == target  # resumed from its restoration context across every resize
    assert d_res.get()[0] == 9
    assert d_calls.get()[0] == sum(sizes)  # every task ran on exactly its k workgroups
    for d in (d_buf, d_res, d_calls, d_iters):
        d.free()


def test_emitted_mega_work_queue(rt, km, dev):
    length, d_buf, d_res, d_calls = reduce_buffers(dev)
    c = Chunks(dev, rounds=64)
    args = L.Args([ctypes.c_int32(length), d_buf.ptr, d_res.ptr, d_calls.ptr] + c.args())
    with S.Mega(rt, np_kind=S.NP_REDUCE_MIN) as m:
        p = m.launch_custom(km.km_mega_chunks(), args, wg_size=256, num_wgs=16)
        assert p == 15
        m.send_persistent_custom(p)
        sizes = drive(m, p)
        m.quit()
    print(len(sizes))
    # forked workgroups skip the initialisation: their `scale` comes from the
    # restoration context, so a lost context shows as wrong outputs
    c.check()
    assert d_res.get()[0] == 9
    assert d_calls.get()[0] == sum(sizes)
    c.free()
    for d in (d_buf, d_res, d_calls):
        d.free()


# ------------------------------------------------------------------ octree
class Octree:
    """Buffers of tests/kernel_merge/octree.hip: particles uniform in
    [-256, 256)^3, the task pools and the outputs (tree, status, leaves)."""

    def __init__(self, dev, n=200_000, maxchilds=32, seed=5, num_pools=16, pool_size=4096):
        rng = np.random.default_rng(seed)
        p = np.zeros((n, 4), np.float32)
        p[:, :3] = rng.uniform(-256.0, 256.0, (n, 3)).astype(np.float32)
        self.inp, self.n, self.maxchilds = p, n, maxchilds
        self.tree_cap = n
        self.num_pools, self.pool_size = num_pools, pool_size
        self.d_p, self.d_new = dev(p), dev(np.zeros_like(p))
        self.d_tree = dev(np.zeros(self.tree_cap, np.uint32))
        self.d_pools = dev(np.zeros(num_pools * pool_size * 8, np.int32))  # OctTask: 8 words
        self.d_locks, self.d_heads = dev(np.zeros(num_pools, np.int32)), dev(np.zeros(num_pools, np.int32))
        self.d_status = dev(np.full(4, -1, np.int32))
        self.d_range = dev(np.zeros((self.tree_cap, 4), np.int32))
        self.d_box = dev(np.zeros((self.tree_cap, 4), np.float32))

    def args(self):
        return [self.d_p.ptr, self.d_new.ptr, self.d_tree.ptr, ctypes.c_int32(self.tree_cap), ctypes.c_int32(self.n),
                ctypes.c_int32(self.maxchilds), ctypes.c_float(256.0), ctypes.c_float(1.0 / 1024), self.d_pools.ptr,
                self.d_locks.ptr, self.d_heads.ptr, ctypes.c_int32(self.num_pools), ctypes.c_int32(self.pool_size),
                self.d_status.ptr, self.d_range.ptr, self.d_box.ptr]

    def leaves_oracle(self):
        """The leaves (centre, half width, count) of the same subdivision, in
        float32 on the host: a box splits while it holds more than maxchilds."""
        out = []
        stack = [(self.inp[:, :3], np.float32(0), np.float32(0), np.float32(0), np.float32(256.0))]
        while stack:
            pts, cx, cy, cz, h = stack.pop()
            box = (pts[:, 0] >= cx).astype(np.int32) + 2 * (pts[:, 1] >= cy) + 4 * (pts[:, 2] >= cz)
            q = np.float32(h * np.float32(0.5))
            for i in range(8):
                sub = pts[box == i]
                c = (cx + (q if i & 1 else -q), cy + (q if i & 2 else -q), cz + (q if i & 4 else -q))
                c = tuple(np.float32(x) for x in c)
                if len(sub) > self.maxchilds and q >= np.float32(1.0 / 1024):
                    stack.append((sub, *c, q))
                else:
                    out.append((float(c[0]), float(c[1]), float(c[2]), float(q), len(sub)))
        return sorted(out)

    def check(self):
        """The reference's self-check (octree.h:202-241: the leaf counts in the
        tree and the finished-particle counter both equal the particle count),
        plus: the leaves partition [0, n), every particle lies in its leaf's
        box, the particles are a permutation of the input, and the leaves are
        those of the host subdivision."""
        st = self.d_status.get()
        assert st[3] == 0, f"tree or pool overflow ({st[3]})"
        tree = self.d_tree.get()[: st[1]]
        assert int((tree[(tree & 0x80000000) != 0] & 0x7FFFFFFF).sum()) == self.n
        assert st[0] == self.n
        nl = int(st[2])
        rng, box = self.d_range.get()[:nl], self.d_box.get()[:nl]
        order = np.argsort(rng[:, 0], kind="stable")
        rng, box = rng[order], box[order]
        live = rng[:, 1] > 0
        starts, counts = rng[live, 0], rng[live, 1]
        assert starts[0] == 0 and np.all(starts[1:] == starts[:-1] + counts[:-1]) and starts[-1] + counts[-1] == self.n
        parts = self.d_p.get()
        for (b, c, _, _), (cx, cy, cz, h) in zip(rng[live], box[live]):
            p = parts[b:b + c, :3]
            lo = np.array([cx - h, cy - h, cz - h], np.float32)
            hi = np.array([cx + h, cy + h, cz + h], np.float32)
            assert np.all(p >= lo) and np.all(p < hi)
        key = lambda a: a[np.lexsort(a[:, :3].T[::-1])]  # noqa: E731
        assert np.array_equal(key(parts), key(self.inp))
        got = sorted((float(b[0]), float(b[1]), float(b[2]), float(b[3]), int(r[1])) for r, b in zip(rng, box))
        assert got == self.leaves_oracle()

    def free(self):
        for d in (self.d_p, self.d_new, self.d_tree, self.d_pools, self.d_locks, self.d_heads, self.d_status,
                  self.d_range, self.d_box):
            d.free()


@pytest.mark.parametrize("wg", [256, 1024])
def test_emitted_standalone_octree(rt, km, dev, wg):
    o = Octree(dev)
    st = rt.launch_standalone(km.km_standalone_octree(), L.Args(o.args()), wg_size=wg)
    print(wg, st)
    assert st["participating"] > 1
    o.check()
    o.free()


@pytest.mark.parametrize("query", [False, True])
def test_emitted_mega_octree(rt, km, dev, query):
    """The irregular fork/kill app under the resizing scheduler: workgroup 0
    forks at the top of the task loop, the others offer kills there, while
    tasks take workgroups away; the tree is still complete and exact."""
    length, d_buf, d_res, d_calls = reduce_buffers(dev)
    o = Octree(dev, n=400_000, maxchilds=16)
    args = L.Args([ctypes.c_int32(length), d_buf.ptr, d_res.ptr, d_calls.ptr] + o.args())
    with S.Mega(rt, np_kind=S.NP_REDUCE_MIN) as m:
        L.lib().wgm_mega_set_query_barrier(m._h, 1 if query else 0)
        p = m.launch_custom(km.km_mega_octree(), args, wg_size=256, num_wgs=16)
        assert p == 15
        m.send_persistent_custom(p)
        sizes = drive(m, p, max_tasks=60)
        m.quit()
    print(len(sizes))
    o.check()
    assert d_res.get()[0] == 9
    assert d_calls.get()[0] == sum(sizes)
    o.free()
    for d in (d_buf, d_res, d_calls):
        d.free()
