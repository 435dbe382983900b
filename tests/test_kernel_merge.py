"""CPU: the HIP kernel_merge emitter (gpu_workgroup_managment_amd.kernel_merge).

The merged form of kernel_merge's own example pair (reduction + simple
barrier, kernel_merge/examples/simple_barrier_with_reduction) is checked
against the structure of the reference's expected output
(expected/merged.cl, expected/restoration_ctx.h) in HIP idiom; every
emitted file matches the committed outputs under tests/kernel_merge/expected;
the rejections of the reference's contract are exercised; and the emitted
kernels compile for gfx950 against include/wgm_device_api.h."""
import ctypes
import os
import re
import subprocess

import pytest

from gpu_workgroup_managment_amd import kernel_merge as km

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
KD = os.path.join(HERE, "kernel_merge")


def read(name):
    with open(os.path.join(KD, name)) as f:
        return f.read()


def squash(s):
    return re.sub(r"\s+", " ", s).strip()


@pytest.fixture(scope="module")
def merged_barrier():
    return km.merge(read("reduction.hip"), "MY_reduce", read("simple_barrier.hip"), "simple_barrier",
                    "mega_barrier", source="reduction.hip + simple_barrier.hip")


def test_restoration_ctx_of_the_example(merged_barrier):
    # expected/restoration_ctx.h: { target; int i; }
    _, rctx = merged_barrier
    assert squash(rctx) == "struct Restoration_ctx { int target; int i; };"


def test_non_persistent_function(merged_barrier):
    # expected/merged.cl:7-35: the __local array becomes a parameter, the
    # kernel context is the last parameter, group ids / sizes come from it
    merged, _ = merged_barrier
    head = re.search(r"__device__ void MY_reduce\(([^)]*)\)", merged).group(1)
    assert squash(head) == ("int length, const int* buffer, int* result, int* calls, int* scratch, "
                            "const Kernel_ctx* __k_ctx")
    body = merged[merged.index("__device__ void MY_reduce"):merged.index("__device__ void simple_barrier")]
    assert "k_get_global_size(__k_ctx)" in body and "k_get_group_id(__k_ctx)" in body
    assert "get_local_id(0)" in body and "get_local_size(0)" in body  # local ids are kept
    assert "__shared__" not in body
    assert "get_group_id(0)" not in body and "get_global_size(0)" not in body


def test_persistent_function(merged_barrier):
    # expected/merged.cl:37-65: restore i when resumed; switch on the target;
    # the loop condition checked at case 0; the barrier saves {1, i} and is
    # followed by case 1; the top-level break becomes return
    merged, _ = merged_barrier
    body = squash(merged[merged.index("__device__ void simple_barrier"):merged.index("__global__")])
    order = ["Restoration_ctx __rc = wgm_restoration_ctx<Restoration_ctx>(__ma);", "int i = 0;",
             "if (__rc.target != 0) { i = __rc.i; }", "while (true) { switch (__rc.target) { case 0:",
             "if (!(true)) return;", "i++;",
             "Restoration_ctx __to_fork; __to_fork.target = 1; __to_fork.i = i; "
             "if (global_barrier_resize(__ma, __k_ctx, __to_fork)) return;",
             "case 1: __rc.target = 0;", "if (i == target_iters) {", "k_get_group_id(__k_ctx) == 0", "return;"]
    pos = 0
    for frag in order:
        j = body.find(frag, pos)
        assert j >= 0, f"missing or out of order: {frag!r}\n{body}"
        pos = j + len(frag)
    assert "break" not in body


def test_mega_kernel(merged_barrier):
    # expected/merged.cl:67-79: non-persistent parameters first, then the
    # persistent ones; the __local array declared in the mega-kernel; the two
    # kernel calls given to the device body
    merged, _ = merged_barrier
    mk = squash(merged[merged.index("__global__"):])
    assert mk.startswith("__global__ void mega_barrier(wgm_mega_args __ma, int length, const int* buffer, "
                         "int* result, int* calls, int* iters_out, int target_iters) {")
    assert "__shared__ int scratch[1024];" in mk
    assert ("WGM_MEGA_BODY(__ma, MY_reduce(length, buffer, result, calls, scratch, np_ctx), "
            "simple_barrier(__ma, p_ctx, iters_out, target_iters));") in mk


GOLDEN = [
    ("merged_barrier.hip", lambda: km.merge(read("reduction.hip"), "MY_reduce", read("simple_barrier.hip"),
                                            "simple_barrier", "mega_barrier",
                                            source="reduction.hip + simple_barrier.hip")[0]),
    ("restoration_ctx_barrier.h", lambda: "#pragma once\n" + km.merge(
        read("reduction.hip"), "MY_reduce", read("simple_barrier.hip"), "simple_barrier")[1]),
    ("merged_chunks.hip", lambda: km.merge(read("reduction.hip"), "MY_reduce", read("chunk_loop.hip"), "chunk_loop",
                                           "mega_chunks", source="reduction.hip + chunk_loop.hip")[0]),
    ("standalone_barrier.hip", lambda: km.standalone(read("simple_barrier.hip"), "simple_barrier",
                                                     source="simple_barrier.hip")),
    ("standalone_chunks.hip", lambda: km.standalone(read("chunk_loop.hip"), "chunk_loop", source="chunk_loop.hip")),
]


@pytest.mark.parametrize("name,make", GOLDEN, ids=[g[0] for g in GOLDEN])
def test_outputs_match_committed(name, make):
    assert make() == read(os.path.join("expected", name))


def test_work_queue_kernel_forms():
    merged, rctx = km.merge(read("reduction.hip"), "MY_reduce", read("chunk_loop.hip"), "chunk_loop", "mega_chunks")
    # restored: the leading scalar declarations; the __shared__ array is hoisted
    assert squash(rctx) == "struct Restoration_ctx { int target; unsigned int scale; int claimed; };"
    p = squash(merged[merged.index("__device__ void chunk_loop"):merged.index("__global__")])
    # initialisation only on a fresh start, its barrier the plain one
    assert "if (__rc.target == 0) { if (k_get_global_id(__k_ctx) == 0) {" in p
    assert "scale = mult; wgm_global_barrier(__ma, __k_ctx); }" in p
    # g is in scope at the fork point inside `if (g == 0)`: hoisted, assigned in place
    assert "int g; while (true)" in p and "g = k_get_group_id(__k_ctx); if (g == 0) {" in p
    assert "wgm_offer_fork(__ma, __k_ctx, __to_fork); } case 1: __rc.target = 0; }" in p
    assert "if (wgm_offer_kill(__ma, __k_ctx)) return;" in p
    # declarations after the last restoration point stay where they are
    assert "int c = s_c[0];" in p and "int chunks = " in p
    # the top-level break returns; the inner loops keep their own code
    assert "if (c >= rounds * chunks) return;" in p
    assert "for (int k = 0; k < 16; k++)" in p
    mk = squash(merged[merged.index("__global__"):])
    assert "__shared__ int scratch[1024]; __shared__ int s_c[1];" in mk
    assert "chunk_loop(__ma, p_ctx, in, out, hits, next, n, chunk, rounds, mult, s_c)" in mk


def test_standalone_form():
    s = squash(km.standalone(read("chunk_loop.hip"), "chunk_loop"))
    assert "__global__ void chunk_loop(wgm_kernel_args __wa, const unsigned* in," in s
    assert s.index("DISCOVERY_PROTOCOL(__wa);") < s.index("unsigned scale = 0;")
    assert s.rstrip().endswith("WGM_STANDALONE_END(__wa); }")
    assert "offer_" not in s and "resizing_global_barrier" not in s
    assert "p_get_global_id() == 0" in s and "p_get_group_id()" in s
    assert s.count("global_barrier()") == 1
    assert "break;" in s  # a standalone loop keeps its break


NESTED = """
__global__ void nested(int* out, int n) {
  int i = 0;
  while (i < n) {
    int before = i * 2;
    for (int k = 0; k < 4; k++) {
      if (k == before) break;
    }
    switch (i) { case 3: break; default: break; }
    resizing_global_barrier();
    i++;
    if (i == 7) break;
    if (i == 8) { return; }
  }
  return;
}
"""
NP_OK = "__global__ void km_np(int* x) { if (get_global_id(0) == 0) x[0] = get_num_groups(0); }\n"


def test_break_depth_and_hoist():
    merged, _ = km.merge(NP_OK, "km_np", NESTED, "nested")
    p = squash(merged[merged.index("__device__ void nested"):merged.index("__global__")])
    assert "if (!(i < n)) return;" in p
    assert "if (k == before) break;" in p  # inside a for: kept
    assert "case 3: break; default: break;" in p  # inside the caller's switch: kept
    assert "if (i == 7) return;" in p
    assert "int before; while (true)" in p and "before = i * 2;" in p
    np_fn = squash(merged[merged.index("__device__ void km_np"):merged.index("__device__ void nested")])
    assert "k_get_global_id(__k_ctx) == 0" in np_fn and "k_get_num_groups(__k_ctx)" in np_fn


@pytest.mark.parametrize("np_src,p_src,msg", [
    ("__global__ void km_np(int* x) { global_barrier(); }\n", NESTED, "non-persistent kernel calls global_barrier"),
    (NP_OK, "__global__ void km_pk(int* y) { while (true) {} while (true) {} }\n", "exactly one top-level while"),
    (NP_OK, "__global__ void km_pk(int* y) { int i = 0; while (i < 3) { for (int k = 0; k < 2; k++) "
            "{ resizing_global_barrier(); } i++; } }\n", "nested loop"),
    (NP_OK, "__global__ void km_pk(int* y) { const int c = 3; while (true) { resizing_global_barrier(); "
            "if (c) break; } }\n", "cannot be const"),
    (NP_OK, "__global__ void km_pk(int* y) { while (true) { break; } y[0] = 1; }\n", "only `return;`"),
    (NP_OK, "__global__ void km_pk(int* y) { offer_fork(); while (true) { break; } }\n", "before the persistent loop"),
    (NP_OK, "__global__ void km_pk(int* y) { while (true) { int a[4]; a[0] = 1; resizing_global_barrier(); "
            "y[0] = a[0]; } }\n", "array a"),
    ("__global__ void km_np(int* y) { y[0] = 0; }\n", "__global__ void km_pk(int* y) { while (true) { break; } }\n",
     "parameter names shared"),
    (NP_OK, "__device__ void km_pk(int* y) { while (true) { break; } }\n", "not a __global__ kernel"),
    (NP_OK, "__global__ void km_pk(int* y) { while (true) { undeclared_call(); } }\n", "could not parse"),
])
def test_contract_rejections(np_src, p_src, msg):
    np_name = re.search(r"void (\w+)\(", np_src).group(1)
    p_name = re.search(r"void (\w+)\(", p_src).group(1)
    with pytest.raises(km.MergeError, match=re.escape(msg)):
        km.merge(np_src, np_name, p_src, p_name)


def test_cli_writes_outputs(tmp_path):
    out, rc = tmp_path / "m.hip", tmp_path / "rc.h"
    r = subprocess.run(["python3", "-m", "gpu_workgroup_managment_amd.kernel_merge", "merge",
                        os.path.join(KD, "reduction.hip"), "MY_reduce", os.path.join(KD, "simple_barrier.hip"),
                        "simple_barrier", "--mega-name", "mega_barrier", "-o", str(out), "--restoration-ctx", str(rc)],
                       cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert out.read_text() == read("expected/merged_barrier.hip")
    assert rc.read_text() == read("expected/restoration_ctx_barrier.h")
    bad = tmp_path / "bad.hip"
    bad.write_text("__global__ void km_pk(int* y) { while (true) {} while (true) {} }\n")
    r = subprocess.run(["python3", "-m", "gpu_workgroup_managment_amd.kernel_merge", "standalone", str(bad), "km_q",
                        "-o", str(tmp_path / "x.hip")], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 1 and "kernel km_q not found" in r.stderr


@pytest.mark.parametrize("name", ["merged_chunks.hip", "standalone_chunks.hip"])
def test_emitted_kernels_compile_for_gfx950(tmp_path, name):
    cmd = ["/opt/rocm/bin/hipcc", "-O3", "-std=c++17", "--offload-arch=gfx950", "--cuda-device-only", "-c",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "gpu_workgroup_managment_amd", "csrc"),
           "-o", str(tmp_path / "k.o"), os.path.join(KD, "expected", name)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_helper_types_and_functions_are_kept(tmp_path):
    """Code outside the kernel -- the task type and the pool helpers of the
    octree example -- goes into both emitted forms ahead of the kernel (and
    of the restoration context, which may hold such types), which then compile
    for gfx950; comment-only surroundings add nothing."""
    merged, _ = km.merge(read("reduction.hip"), "MY_reduce", read("octree.hip"), "octree", "mega_octree")
    alone = km.standalone(read("octree.hip"), "octree")
    for text, kernel in ((merged, "__device__ void octree("), (alone, "__global__ void octree(")):
        assert "struct OctTask {" in text and "__device__ inline bool oct_pop(" in text
        assert text.index("struct OctTask {") < text.index(kernel)
    assert merged.index("struct OctTask {") < merged.index("struct Restoration_ctx")
    assert merged.count("struct OctTask {") == 1
    # the kernel itself appears once, rewritten, not copied into the support code
    assert alone.count("void octree(") == 1
    for name, text in (("m.hip", merged), ("s.hip", alone)):
        src = tmp_path / name
        src.write_text(text)
        cmd = ["/opt/rocm/bin/hipcc", "-O3", "-std=c++17", "--offload-arch=gfx950", "--cuda-device-only", "-c",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "gpu_workgroup_managment_amd", "csrc"),
               "-o", str(tmp_path / (name + ".o")), str(src)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
    assert "struct" not in km.standalone(read("simple_barrier.hip"), "simple_barrier").split("__global__")[0]


def test_fixture_library_exports():
    path = os.path.join(KD, "libwgm_kmerge.so")
    if not os.path.exists(path):
        pytest.fail(f"{path} missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for f in ("km_mega_barrier", "km_mega_chunks", "km_standalone_barrier", "km_standalone_chunks", "km_mega_bfs",
              "km_standalone_bfs", "km_mega_octree", "km_standalone_octree"):
        getattr(lib, f).restype = ctypes.c_void_p
        assert getattr(lib, f)() != 0
