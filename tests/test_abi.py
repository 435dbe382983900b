"""CPU: the C-ABI boundary. libwgm.so loads without a GPU, exports every
function include/wgm.h declares, the ctypes mirror binds exactly those, and
every ctypes struct has the C layout (sizes and field offsets compiled from
the header with gcc). Host-only entry points run; nothing launches a kernel."""
import ctypes
import os
import re
import subprocess

import pytest

from gpu_workgroup_managment_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the C-ABI headers (wgm_device_api.h is HIP device code for the caller's kernels)
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h") and "device" not in f]

STRUCTS = {
    "wgm_csr": L.Csr,
    "wgm_batch_item": L.BatchItem,
    "wgm_device_info": L.DeviceInfo,
    "wgm_barrier_result": L.BarrierResult,
    "wgm_launch": L.Launch,
    "wgm_run_stats": L.RunStats,
    "wgm_batch_config": L.BatchConfig,
    "wgm_batch_stats": L.BatchStats,
    "wgm_np_config": L.NpConfig,
    "wgm_task_times": L.TaskTimes,
    "wgm_mega_stats": L.MegaStats,
}
STRUCTS.update(getattr(L, "EXTRA_STRUCTS", {}))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(wgm_\w+)\s*\(", text, flags=re.M):
            if not text[m.start():m.end()].lstrip().startswith(("typedef", "#", "struct")):
                names.add(m.group(1))
    return names


def test_library_loads_and_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_mirror_binds_exactly_the_header():
    assert set(L.SIGNATURES) == declared_functions()


@pytest.fixture(scope="module")
def layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("abi")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "wgm.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'  printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = d / "probe.c"
    src.write_text("\n".join(lines))
    exe = d / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return dict(line.rsplit(" ", 1) for line in out.splitlines())


@pytest.mark.parametrize("cname", sorted(STRUCTS))
def test_struct_layout_matches_c(layout, cname):
    cls = STRUCTS[cname]
    assert int(layout[cname]) == ctypes.sizeof(cls)
    for f, _ in cls._fields_:
        assert int(layout[f"{cname}.{f}"]) == getattr(cls, f).offset, f


def test_strerror_names_every_status():
    lib = L.lib()
    for code, name in L.ERRORS.items():
        s = lib.wgm_strerror(code).decode()
        assert s and "unknown" not in s.lower(), (code, s)
    assert lib.wgm_strerror(0).decode()


def test_null_arguments_are_rejected_not_dereferenced():
    lib = L.lib()
    assert lib.wgm_destroy(None) == 0
    assert lib.wgm_graph_destroy(None) == 0
    assert lib.wgm_batch_destroy(None) == 0
    assert lib.wgm_mega_destroy(None) == 0
    assert lib.wgm_gen_grid(0, 5, None) < 0
    out = L.Csr()
    assert lib.wgm_gen_grid(-1, 5, ctypes.byref(out)) == -1
    assert lib.wgm_gen_er(10, 1000, 1, ctypes.byref(out)) == -1  # more pairs than exist
    assert lib.wgm_batch_spec(0, 5, 10, 18, None) < 0


def test_dist_digest_host_function():
    import numpy as np

    rng = np.random.default_rng(5)
    dist = rng.integers(0, 1 << 32, 1000, dtype=np.uint64).astype(np.uint32)
    dist[::7] = 1_000_000_000
    x = (np.arange(1000, dtype=np.uint64) << np.uint64(32)) | dist.astype(np.uint64)
    z = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    want = int(z.sum(dtype=np.uint64))
    got = L.lib().wgm_dist_digest(dist.ctypes.data_as(ctypes.c_void_p), len(dist))
    assert got == want
    # order independence is what lets the device sum it in any order
    assert L.lib().wgm_dist_digest(None, 0) == 0


def test_device_api_header_compiles_for_gfx950(tmp_path):
    """The public device API builds a caller's kernel (host compile + gfx950)."""
    src = tmp_path / "k.hip"
    src.write_text('''#include "wgm_device_api.h"
__global__ void k(wgm_kernel_args wa, int* out) {
  DISCOVERY_PROTOCOL(wa);
  if (get_local_id(0) == 0) out[get_group_id(0)] = get_num_groups(0);
  global_barrier();
  WGM_STANDALONE_END(wa);
}
''')
    cmd = ["/opt/rocm/bin/hipcc", "-std=c++17", "--offload-arch=gfx950", "-c", str(src), "-o", str(tmp_path / "k.o"),
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "gpu_workgroup_managment_amd", "csrc")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_serve_plan_keeps_graphs_and_request_state_within_the_budget():
    """The query server's request-state arena is planned on the host: it is
    max_inflight x the largest request's state, cut to what the budget leaves
    after the graphs, so the server never allocates more than its budget."""
    from gpu_workgroup_managment_amd import graphs as G
    from gpu_workgroup_managment_amd import serving as SV

    gs = [G.add_weights(G.grid(40, 40), 1, 255, 3), G.rmat(10, 8), G.add_weights(G.rmat(11, 8), 1, 255, 3)]
    gbytes, arena = SV.plan(gs, SV.config(max_inflight=32))
    # 256-byte aligned row_ptr / col / wt per graph
    al = lambda w: (w + 63) // 64 * 64 * 4  # noqa: E731
    want_g = sum(al(g.n + 1) + al(g.m) + (al(g.m) if g.wt is not None else 0) for g in gs)
    assert gbytes == want_g
    nmax = max(g.n for g in gs)
    largest = (7 * nmax + 4 + 63) // 64 * 64 * 4  # an SSSP request: 7n + 4 words
    assert arena == 32 * largest  # a generous budget: every slot holds the largest request
    for budget in (gbytes + 2 * largest, gbytes + largest + 1000, gbytes + 4096):
        gb, ab = SV.plan(gs, SV.config(max_inflight=32, budget_bytes=budget))
        assert gb == gbytes and gb + ab <= budget and ab >= 0
        assert ab >= min(32 * largest, budget - gbytes - 255)
    with pytest.raises(L.WgmError) as e:
        SV.plan(gs, SV.config(budget_bytes=gbytes - 1))
    assert e.value.code == -9
