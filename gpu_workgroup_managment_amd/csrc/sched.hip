// sched.hip -- the mega-kernel (scheduler workgroup + workers), the
// resizable persistent BFS / SSSP, the non-persistent tasks, and the host
// side of the scheduler protocol (the reference's CL_Communicator,
// src/scheduler_rt/rt_host/cl_communicator.h:19-465, and run_merged,
// src/host_skeleton/skel.cpp:405-631).

#include "sched_internal.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

using namespace wgm;

namespace wgm {

// Standalone launch of a non-persistent task (run_non_persistent,
// src/host_skeleton/skel.cpp:168-221).
__global__ void __launch_bounds__(1024) k_np_solo(NPDev t, int kind) {
  extern __shared__ int lds[];
  np_task(t, kind, blockIdx.x, gridDim.x, lds);
}

// ================================================== persistent tasks
// Frontier BFS / SSSP with a resizing barrier per level (merged drelax2,
// src/lonestar/graph_apps/bfs/device/merged.cl:310-396). The initial team
// resets the traversal state itself (reset_persistent_task), so nothing
// else has to run on the GPU while the mega-kernel holds it. Work inside a
// level is claimed dynamically, so the set of workgroups may change at every
// barrier; the restoration context carries the resume label and the level.
template <bool SSSP>
__device__ void p_frontier(const SchedDev& s, const MegaApp& a, RCtx rin, int did, int* lds) {
  const int vid = s.p.group_ids[did];
  int level;
  if (rin.target == 0) {
    const int k0 = ld_relaxed(s.bar_groups);
    int lo, hi;
    block_range(a.g.n, vid, k0, &lo, &hi);
    for (int v = lo + (int)threadIdx.x; v < hi; v += blockDim.x) {
      a.f.dist[v] = (v == a.src) ? 0u : kInf;
      if (SSSP) a.f.stamp[v] = 0u;
    }
    if (vid == 0 && threadIdx.x == 0) {
      a.f.q0[0] = a.src;
      st_relaxed(&a.f.counts[0], 1);
      st_relaxed(&a.f.counts[kLine], 0);
      st_relaxed(&a.f.counts[2 * kLine], 0);
      for (int l = 0; l < 3; l++) st_relaxed(&a.f.claims[l * kLine], 0);
    }
    if (resize_barrier(s, vid, RCtx{1, -1})) return;
    level = -1;
  } else {
    level = rin.level;
  }
  while (true) {
    level++;
    if (ld_relaxed(&a.f.counts[(level % 3) * kLine]) == 0) break;
    // the current group count sizes a narrow level's claims
    frontier_level<SSSP>(a.g, a.f, level, vid, max(1, (int)ld_relaxed(s.bar_groups)), lds);
    if (resize_barrier(s, vid, RCtx{1, level})) return;
  }
  if (vid == 0 && threadIdx.x == 0) *a.iters = level;
}

// ------------------------------------------ resizable topology-driven tasks
// Colour, MIS and PageRank under the resizing scheduler (merged color / mis,
// src/pannotia/graph_apps/color/device/merged.cl:73-205,
// src/pannotia/graph_apps/mis/device/merged.cl:75-224). Every phase ends in a
// resizing barrier, the restoration context is {resume label, round}, and
// the vertices of a phase are claimed in chunks from a counter, so a
// workgroup forked in at a barrier or killed at one changes nothing in how
// the phase is split. The per-vertex results do not depend on the split: a
// round's outcome is fixed by the states at its start (colour, MIS), and a
// PageRank sum is taken by one thread in CSR order. Claim counters rotate
// over three lines by phase index (phase p uses line p % 3, and workgroup 0
// clears line (p + 1) % 3, whose last users finished two barriers ago); the
// stop flags rotate the same way by round.
constexpr int kPhaseChunk = 2048;  // vertices per claim

template <class Body>
__device__ __forceinline__ void claim_chunks(int n, int* claim, Body&& body) {
  __shared__ int s_c;
  while (true) {
    if (threadIdx.x == 0) s_c = add_relaxed(claim, 1);
    __syncthreads();
    const long long lo = (long long)__builtin_amdgcn_readfirstlane(s_c) * kPhaseChunk;
    if (lo >= n) break;
    body((int)lo, (int)(lo + kPhaseChunk < n ? lo + kPhaseChunk : n));
    __syncthreads();  // the body ends in a barrier: the next claim by thread 0 stays one loop
  }
  __syncthreads();
}

// reset_persistent_task on the device: static blocks of the initial team
// (no resize happens before the first barrier), and workgroup 0 clears the
// claim and stop lines.
template <class Init>
__device__ __forceinline__ void topo_init(const SchedDev& s, const MegaApp& a, int vid, Init&& init) {
  const int k0 = ld_relaxed(s.bar_groups);
  int lo, hi;
  block_range(a.g.n, vid, k0, &lo, &hi);
  for (int v = lo + (int)threadIdx.x; v < hi; v += blockDim.x) init(v);
  if (vid == 0 && threadIdx.x == 0) {
    for (int l = 0; l < 3; l++) {
      st_relaxed(&a.claims[l * kLine], 0);
      st_relaxed(&a.flags[l * kLine], 0);
    }
  }
}

// Colour: one resize point per round (the fused round of color_round; the
// reference has two, after color1 and after color2). Resume label 1 = right
// after the barrier that ends round `level` (0: the initialisation).
__device__ void p_color(const SchedDev& s, const MegaApp& a, RCtx rin, int did) {
  const int vid = s.p.group_ids[did];
  int round = 0;
  if (rin.target == 0) {
    topo_init(s, a, vid, [&](int v) { a.ia[v] = -1; });
    if (resize_barrier(s, vid, RCtx{1, 0})) return;
  } else {
    round = rin.level;
  }
  while (true) {
    if (round > 0 && ld_relaxed(&a.flags[(round % 3) * kLine]) == 0) break;
    round++;
    bool any = false;
    claim_chunks(a.g.n, &a.claims[(round % 3) * kLine],
                 [&](int lo, int hi) { any = color_round(a.g, a.ia, round, lo, hi) || any; });
    if (__syncthreads_or(any) && threadIdx.x == 0) st_relaxed(&a.flags[(round % 3) * kLine], 1);
    if (vid == 0 && threadIdx.x == 0) {
      st_relaxed(&a.flags[((round + 1) % 3) * kLine], 0);
      st_relaxed(&a.claims[((round + 1) % 3) * kLine], 0);
    }
    if (resize_barrier(s, vid, RCtx{1, round})) return;
  }
  if (vid == 0 && threadIdx.x == 0) *a.iters = round;
}

// MIS: two resize points per round (phase A = mis1 + the join of mis2,
// phase B = the removals of mis2 + mis3; the reference has three). Resume
// label 1 = after phase A of round `level`, 2 = after phase B (0: after the
// initialisation). Phase A of round r is phase 2r - 1, phase B is 2r.
__device__ void p_mis(const SchedDev& s, const MegaApp& a, RCtx rin, int did) {
  const int vid = s.p.group_ids[did];
  int round = 0, at = 2;
  if (rin.target == 0) {
    topo_init(s, a, vid, [&](int v) {
      a.ia[v] = -1;
      a.ib[v] = 0;
    });
    if (resize_barrier(s, vid, RCtx{2, 0})) return;
  } else {
    round = rin.level;
    at = rin.target;
  }
  while (true) {
    if (at == 2) {
      round++;
      bool any = false;
      claim_chunks(a.g.n, &a.claims[((2 * round - 1) % 3) * kLine],
                   [&](int lo, int hi) { any = mis_phase_a(a.g, a.ia, a.ib, round, lo, hi) || any; });
      if (__syncthreads_or(any) && threadIdx.x == 0) st_relaxed(&a.flags[(round % 3) * kLine], 1);
      if (vid == 0 && threadIdx.x == 0) {
        st_relaxed(&a.flags[((round + 1) % 3) * kLine], 0);
        st_relaxed(&a.claims[((2 * round) % 3) * kLine], 0);
      }
      if (resize_barrier(s, vid, RCtx{1, round})) return;
      at = 1;
    } else {
      if (ld_relaxed(&a.flags[(round % 3) * kLine]) == 0) break;
      claim_chunks(a.g.n, &a.claims[((2 * round) % 3) * kLine],
                   [&](int lo, int hi) { mis_phase_b(a.g, a.ia, round, lo, hi); });
      if (vid == 0 && threadIdx.x == 0) st_relaxed(&a.claims[((2 * round + 1) % 3) * kLine], 0);
      if (resize_barrier(s, vid, RCtx{2, round})) return;
      at = 2;
    }
  }
  if (vid == 0 && threadIdx.x == 0) *a.iters = round;
}

// PageRank: one resize point per iteration; resume label 1 with the number
// of iterations done (0: after the initial contributions). Iteration i is
// phase i + 1.
__device__ void p_pagerank(const SchedDev& s, const MegaApp& a, RCtx rin, int did) {
  const int vid = s.p.group_ids[did];
  int it = 0;
  if (rin.target == 0) {
    topo_init(s, a, vid, [&](int v) {
      const int deg = a.g.row_ptr[v + 1] - a.g.row_ptr[v];
      a.fa[v] = deg ? a.r0 / (float)deg : 0.0f;
    });
    if (resize_barrier(s, vid, RCtx{1, 0})) return;
  } else {
    it = rin.level;
  }
  while (it < a.iterations) {
    const float* cin = (it & 1) ? a.fb : a.fa;
    float* cout = (it & 1) ? a.fa : a.fb;
    const bool last = it == a.iterations - 1;
    claim_chunks(a.g.n, &a.claims[((it + 1) % 3) * kLine],
                 [&](int lo, int hi) { pagerank_iter(a.g, cin, cout, a.fc, a.damping, a.base, last, lo, hi); });
    if (vid == 0 && threadIdx.x == 0) st_relaxed(&a.claims[((it + 2) % 3) * kLine], 0);
    it++;
    if (resize_barrier(s, vid, RCtx{1, it})) return;
  }
  if (vid == 0 && threadIdx.x == 0) *a.iters = it;
}

// ======================================================= the mega-kernel
// src/scheduler_rt/rt_device/main_device_body.cl:1-93
__global__ void __launch_bounds__(1024) k_mega(Ctl c, SchedDev s, const MegaApp* app_cell, NPDev np) {
  extern __shared__ int lds[];
  int n;
  const int did = discovery_protocol(c, &n);
  if (did < 0 || c.im_exit) return;
  if (did == 0) {
    if (threadIdx.x == 0) {
      scheduler_init(s, n);
      scheduler_loop(s, n - 1);
    }
    __syncthreads();
    return;
  }
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) {
      st_relaxed(&s.task_array[did * kLine], TASK_WAIT);
      add_relaxed(s.available, 1);
    }
    const int task = get_task(s, did);
    if (task == TASK_QUIT) break;
    if (task == TASK_MULT) {
      np_task(np, ld_relaxed(s.np_kind), s.np.group_ids[did], k_num_groups(s.np), lds, s.np_clock, s.np.executing);
    } else if (task == TASK_PERSIST) {
      const MegaApp app = read_app(app_cell);
      const RCtx r = *rslot<RCtx>(s.r_ctx, did);
      switch (app.kind) {
        case WGM_PERSIST_SSSP: p_frontier<true>(s, app, r, did, lds); break;
        case WGM_PERSIST_COLOR: p_color(s, app, r, did); break;
        case WGM_PERSIST_MIS: p_mis(s, app, r, did); break;
        case WGM_PERSIST_PAGERANK: p_pagerank(s, app, r, did); break;
        default: p_frontier<false>(s, app, r, did, lds); break;
      }
      drain_vmem();
      __syncthreads();
      if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        add_relaxed(s.p.executing, -1);
        sys_add_release(s.persistent_flag, -1);
      }
    }
  }
}

__global__ void k_sched_reset(SchedDev s, int max_groups) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int i = t; i < max_groups; i += gridDim.x * blockDim.x) {
    s.task_array[i * kLine] = TASK_UNINIT;
    s.flags[i * kLine] = 0;
    *(int*)rslot<RCtx>(s.r_ctx, i) = 0;
    s.np.group_ids[i] = -1;
    s.p.group_ids[i] = -1;
  }
  if (t == 0) {
    *s.available = 0;
    *s.pool_lock = 0;
    *s.groups_to_kill = 0;
    *s.check_value = -1;
    *s.bar_groups = 0;
    *s.bar_to_kill = 0;
    *s.np.num_groups = 0;
    *s.np.executing = 0;
    *s.p.num_groups = 0;
    *s.p.executing = 0;
    *s.np_kind = 0;
    *s.serve_on = 0;
    *s.np_epoch = 0;
    *s.np_ticket = 0;
    *s.np_k = 0;
  }
}

__global__ void k_np_reset(NPDev t) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *t.counter = 0;
    if (t.dmin) *t.dmin = INT32_MAX;
  }
}

}  // namespace wgm

#define fail wgm_fail
#define HIPCHK WGM_HIPCHK


static int mega_lds(int wg) { return mega_lds_bytes(wg); }

bool wgm_sched_kernel_info(int kernel, KernelInfo* out) {
  if (kernel == WGM_K_MEGA) {
    *out = KernelInfo{(const void*)&k_mega, mega_lds};
    return true;
  }
  return wgm_serve_kernel_info(kernel, out);
}

int wgm_sched_probe(wgm_runtime* rt, int kernel, int grid, int wg_size) {
  if (kernel != WGM_K_MEGA) return wgm_serve_probe(rt, kernel, grid, wg_size);
  hipLaunchKernelGGL(k_mega, dim3(grid), dim3(wg_size), mega_lds_bytes(wg_size), rt->stream, rt->ctl, SchedDev{},
                     (const MegaApp*)nullptr, NPDev{});
  return WGM_OK;
}

void wgm_runtime_release_extras(wgm_runtime* rt) { (void)rt; }

// ---------------------------------------------------------------- wgm_mega
// init_non_persistent_app for every kind at once: all buffers stay resident
// and the kind is chosen per task.
int mega_np_setup(wgm_mega* m, const wgm_np_config* cfg) {
  wgm_runtime* rt = m->rt;
  m->np_cfg = *cfg;
  if (cfg->kind < WGM_NP_MATMULT_I32 || cfg->kind > WGM_NP_REDUCE_MIN)
    return fail(rt, WGM_ERR_INVALID, "unknown non-persistent task kind %d", cfg->kind);
  m->np_kind = cfg->kind;
  NPDev& t = m->np;
  t = NPDev{};
  t.dim = cfg->dim > 0 ? cfg->dim : 256;
  t.len = cfg->length > 0 ? cfg->length : (1 << 20);
  if (t.dim % 32 != 0) return fail(rt, WGM_ERR_INVALID, "matmul dim must be a multiple of 32");
  // counter and minimum lines, then one FP64 checksum partial per workgroup
  HIPCHK(rt, hipMalloc(&m->np_dev, 4 * kLine * sizeof(int) + (size_t)rt->max_grid * sizeof(double)));
  t.counter = (int*)m->np_dev;
  t.dmin = (int*)m->np_dev + kLine;
  t.partial = (double*)((int*)m->np_dev + 4 * kLine);
  t.h_result = m->h_np_result_dev;
  t.h_checksum = (unsigned long long*)(m->h_np_checksum_dev);
  const size_t nn = (size_t)t.dim * t.dim;
  // matmult: middle-square digit fill (src/non_persistent_kernels/matmult/matmult.h:38-50),
  // A then B from one stream, seed 1234
  std::vector<int32_t> A(nn), B(nn);
  int32_t seed = 1234;
  auto fill = [&](std::vector<int32_t>& M) {
    for (size_t i = 0; i < nn; i++) {
      seed = (int32_t)((uint32_t)seed * (uint32_t)seed);
      seed = (seed / 1000) % 1000000;
      M[i] = seed % 10;
    }
  };
  fill(A);
  fill(B);
  // expected diagonal hash (the self-check of matmult.h:144-149,190-192)
  uint32_t h = 5381;
  for (int i = 0; i < t.dim; i++) {
    uint32_t acc = 0;
    for (int k = 0; k < t.dim; k++) acc += (uint32_t)A[(size_t)i * t.dim + k] * (uint32_t)B[(size_t)k * t.dim + i];
    h = h * 33u + acc;
  }
  m->np_expect_hash = (int32_t)h;
  // FP32 matmul: A, B uniform in [-1, 1) from splitmix64 (seed 6 unless given)
  std::vector<float> Af(nn), Bf(nn);
  uint64_t st = cfg->seed ? (uint64_t)cfg->seed : 6;
  for (size_t i = 0; i < nn; i++) Af[i] = (float)((double)(wgm_splitmix64_next(&st) >> 11) * 0x1p-53 * 2.0 - 1.0);
  for (size_t i = 0; i < nn; i++) Bf[i] = (float)((double)(wgm_splitmix64_next(&st) >> 11) * 0x1p-53 * 2.0 - 1.0);
  // expected weighted checksum u^T (A B) v = sum_k (u^T A)_k (B v)_k in FP64,
  // and its tolerance: the FP32 dot-product bound |fl(x.y) - x.y| <= K u sum|x_k y_k|
  // summed with the same weights (doubled for any accumulation order)
  double sum = 0.0, abs_sum = 0.0;
  for (int k = 0; k < t.dim; k++) {
    double ua = 0.0, bv = 0.0, uaa = 0.0, bva = 0.0;
    for (int i = 0; i < t.dim; i++) {
      ua += np_weight_u(i) * (double)Af[(size_t)i * t.dim + k];
      uaa += np_weight_u(i) * std::fabs((double)Af[(size_t)i * t.dim + k]);
    }
    for (int j = 0; j < t.dim; j++) {
      bv += (double)Bf[(size_t)k * t.dim + j] * np_weight_v(j);
      bva += std::fabs((double)Bf[(size_t)k * t.dim + j]) * np_weight_v(j);
    }
    sum += ua * bv;
    abs_sum += uaa * bva;
  }
  m->np_expect_sum = sum;
  m->np_tol = 2.0 * (double)t.dim * 0x1p-24 * abs_sum + 1e-6 * std::fabs(sum);
  // reduction: values in [1, 2^31 - 1] from splitmix64 seed 7 (the reference
  // uses rand() + 1, src/non_persistent_kernels/reduce/reduce.h:18-35)
  std::vector<int32_t> v(t.len);
  uint64_t rs = 7;
  int32_t mn = INT32_MAX;
  for (int i = 0; i < t.len; i++) {
    v[i] = 1 + (int32_t)((wgm_splitmix64_next(&rs) >> 33) % 2147483646ULL);
    mn = std::min(mn, v[i]);
  }
  m->np_expect_min = mn;
  // one allocation: A B C (int32), Af Bf Cf (float), the reduction input
  const size_t mat = ((nn * 4 + 255) / 256) * 256;
  const size_t bytes = 6 * mat + (size_t)t.len * 4;
  HIPCHK(rt, hipMalloc(&m->np_mem, bytes));
  char* p = (char*)m->np_mem;
  t.A = (const int*)p;
  t.B = (const int*)(p + mat);
  t.C = (int*)(p + 2 * mat);
  t.Af = (const float*)(p + 3 * mat);
  t.Bf = (const float*)(p + 4 * mat);
  t.Cf = (float*)(p + 5 * mat);
  t.buf = (const int*)(p + 6 * mat);
  HIPCHK(rt, hipMemcpy((void*)t.A, A.data(), nn * 4, hipMemcpyHostToDevice));
  HIPCHK(rt, hipMemcpy((void*)t.B, B.data(), nn * 4, hipMemcpyHostToDevice));
  HIPCHK(rt, hipMemcpy((void*)t.Af, Af.data(), nn * 4, hipMemcpyHostToDevice));
  HIPCHK(rt, hipMemcpy((void*)t.Bf, Bf.data(), nn * 4, hipMemcpyHostToDevice));
  HIPCHK(rt, hipMemcpy((void*)t.buf, v.data(), (size_t)t.len * 4, hipMemcpyHostToDevice));
  return WGM_OK;
}

// reset_non_persistent, host part: the result cell. Inside the mega-kernel
// the scheduler workgroup resets the device counters itself; a solo launch
// (no mega-kernel resident) resets them with a kernel on the main stream.
int mega_np_reset(wgm_mega* m, int kind, bool solo) {
  wgm_runtime* rt = m->rt;
  hstore(m->h_np_result, kind == WGM_NP_MATMUL_F32 ? 0 : INT32_MIN);
  hstore(m->h_task_kind, kind);
  if (solo) {
    hipLaunchKernelGGL(k_np_reset, dim3(1), dim3(64), 0, rt->stream, m->np);
    HIPCHK(rt, hipGetLastError());
  }
  return WGM_OK;
}

// check_non_persistent_task: compare the published summary
bool mega_np_check(wgm_mega* m, int kind) {
  const int r = hload(m->h_np_result);
  if (kind == WGM_NP_MATMUL_F32) {
    if (r != 1) return false;
    unsigned long long bits = __hip_atomic_load(m->h_np_checksum, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    double got;
    std::memcpy(&got, &bits, 8);
    return std::fabs(got - m->np_expect_sum) <= m->np_tol;
  }
  return r == (kind == WGM_NP_MATMULT_I32 ? m->np_expect_hash : m->np_expect_min);
}

// Scheduler state reset on the runtime's stream (before a resident launch).
int mega_enqueue_sched_reset(wgm_mega* m) {
  wgm_runtime* rt = m->rt;
  hipLaunchKernelGGL(k_sched_reset, dim3(64), dim3(256), 0, rt->stream, m->sd, rt->max_groups);
  HIPCHK(rt, hipGetLastError());
  return WGM_OK;
}

// The synchronous task protocol shared by the mega-kernel and the query
// server: reset, TO_TASK, GOT_GROUPS, TO_EXECUTE, WAITING, check.
int mega_run_task(wgm_mega* m, int kind, int groups, wgm_task_times* out) {
  wgm_runtime* rt = m->rt;
  const bool builtin = kind >= 0;  // kind < 0: a caller's task, checked by the caller
  if (builtin) {
    int rc = mega_np_reset(m, kind, false);
    if (rc) return rc;
  }
  m->np_groups = groups;
  hstore(m->h_task_size, groups);
  const uint64_t r0 = now_ns();
  hstore(m->h_flag, DEVICE_TO_TASK);
  if (!hwait(m->h_flag, DEVICE_GOT_GROUPS, rt->timeout_s)) return fail(rt, WGM_ERR_TIMEOUT, "no DEVICE_GOT_GROUPS");
  const uint64_t r1 = now_ns();
  hstore(m->h_flag, DEVICE_TO_EXECUTE);
  if (!hwait(m->h_flag, DEVICE_WAITING, rt->timeout_s)) return fail(rt, WGM_ERR_TIMEOUT, "no DEVICE_WAITING");
  const uint64_t r2 = now_ns();
  const bool ok = builtin ? mega_np_check(m, kind) : true;
  if (!ok) m->errors++;
  m->response_us.push_back((double)(r1 - r0) * 1e-3);
  m->exec_us.push_back((double)(r2 - r1) * 1e-3);
  if (out) {
    out->response_us = (double)(r1 - r0) * 1e-3;
    out->execution_us = (double)(r2 - r1) * 1e-3;
    out->ok = ok ? 1 : 0;
    unsigned long long tv[5];
    for (int i = 0; i < 5; i++)
      tv[i] = __hip_atomic_load(&m->h_np_times[i], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    const double us = 1e6 / (double)rt->info.wall_clock_hz;
    auto d = [&](int a, int b) { return tv[b] >= tv[a] ? (double)(tv[b] - tv[a]) * us : 0.0; };
    out->dev_assign_us = d(0, 1);
    out->dev_start_us = d(1, 2);
    out->dev_body_us = d(2, 3);
    out->dev_retire_us = d(3, 4);
  }
  return ok ? WGM_OK : fail(rt, WGM_ERR_MISMATCH, "non-persistent task result check failed");
}

extern "C" {

int wgm_mega_create(wgm_runtime* rt, const wgm_np_config* np, wgm_mega** out) {
  if (!rt || !np || !out) return fail(rt, WGM_ERR_INVALID, "NULL argument");
  wgm_mega* m = new wgm_mega();
  m->rt = rt;
  auto bail = [&](int rc) { wgm_mega_destroy(m); return rc; };
  const size_t line_b = kLine * sizeof(int);
  // host cells: flag, task_size, participating, persistent_flag, stop, np
  // result, np checksum; lines 8..15 hold the persistent task descriptor
  // lines 16..17: the device-clock stamps of the last task
  hipError_t e = hipHostMalloc((void**)&m->h_cells, 24 * line_b, hipHostMallocCoherent | hipHostMallocMapped);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipHostMalloc(scheduler cells)"));
  std::memset(m->h_cells, 0, 24 * line_b);
  static_assert(sizeof(MegaApp) <= 8 * kLine * sizeof(int), "task descriptor exceeds its cells");
  int* hd = nullptr;
  e = hipHostGetDevicePointer((void**)&hd, m->h_cells, 0);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipHostGetDevicePointer"));
  m->h_flag = m->h_cells;
  m->h_task_size = m->h_cells + kLine;
  m->h_participating = m->h_cells + 2 * kLine;
  m->h_persistent = m->h_cells + 3 * kLine;
  m->h_stop = m->h_cells + 4 * kLine;
  m->h_np_result = m->h_cells + 5 * kLine;
  m->h_task_kind = m->h_cells + 7 * kLine;
  m->h_np_checksum = (unsigned long long*)(m->h_cells + 6 * kLine);
  m->h_np_result_dev = hd + 5 * kLine;
  m->h_np_checksum_dev = (unsigned long long*)(hd + 6 * kLine);
  m->h_app = (MegaApp*)(m->h_cells + 8 * kLine);
  m->h_app_dev = (MegaApp*)(hd + 8 * kLine);
  SchedDev& s = m->sd;
  s.scheduler_flag = hd;
  s.task_size = hd + kLine;
  s.participating = hd + 2 * kLine;
  s.persistent_flag = hd + 3 * kLine;
  s.stop = hd + 4 * kLine;
  s.task_kind = hd + 7 * kLine;
  m->h_np_times = (unsigned long long*)(m->h_cells + 16 * kLine);
  s.h_np_times = (unsigned long long*)(hd + 16 * kLine);
  // device arena
  const int G = rt->max_groups;
  const size_t bytes = 16 * line_b + 2 * (size_t)G * line_b + (size_t)G * kRSlotBytes + 2 * (size_t)G * sizeof(int);
  e = hipMalloc(&m->d_arena, bytes);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipMalloc(scheduler arena)"));
  char* p = (char*)m->d_arena;
  auto line = [&]() { int* r = (int*)p; p += line_b; return r; };
  s.available = line();
  s.pool_lock = line();
  s.groups_to_kill = line();
  s.check_value = line();
  s.bar_groups = line();
  s.bar_to_kill = line();
  s.np.num_groups = line();
  s.np.executing = line();
  s.p.num_groups = line();
  s.p.executing = line();
  s.np_kind = line();
  s.serve_on = line();
  s.np_clock = (unsigned long long*)line();
  s.np_epoch = line();
  s.np_ticket = line();
  s.np_k = line();
  s.task_array = (int*)p;
  p += (size_t)G * line_b;
  s.flags = (int*)p;
  p += (size_t)G * line_b;
  s.r_ctx = (unsigned char*)p;
  p += (size_t)G * kRSlotBytes;
  s.np.group_ids = (int*)p;
  p += (size_t)G * sizeof(int);
  s.p.group_ids = (int*)p;
  int rc = mega_np_setup(m, np);
  if (rc) return bail(rc);
  s.np_counter = m->np.counter;
  s.np_dmin = m->np.dmin;
  *out = m;
  return WGM_OK;
}

int wgm_mega_destroy(wgm_mega* m) {
  if (!m) return WGM_OK;
  if (m->running) (void)wgm_mega_quit(m);
  if (m->monitor.joinable()) m->monitor.join();
  void* ptrs[] = {m->d_arena, m->np_dev, m->np_mem, m->claims};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (m->h_cells) (void)hipHostFree(m->h_cells);
  delete m;
  return WGM_OK;
}

// launch_mega_kernel (src/scheduler_rt/rt_host/cl_communicator.h:71-95):
// reset, launch, wait for DEVICE_WAITING, read the participating groups.
int wgm_mega_launch(wgm_mega* m, int32_t wg_size, int32_t num_wgs, int32_t query_barrier, int32_t* participating) {
  if (!m) return fail(nullptr, WGM_ERR_INVALID, "NULL mega");
  wgm_runtime* rt = m->rt;
  if (m->running) return fail(rt, WGM_ERR_STATE, "mega-kernel already running");
  int grid = 0;
  int rc = wgm_resolve_grid(rt, WGM_K_MEGA, wg_size, num_wgs, &grid);
  if (rc) return rc;
  if (grid < 2) return fail(rt, WGM_ERR_INVALID, "the mega-kernel needs at least 2 workgroups");
  m->wg_size = wg_size;
  m->sd.query_barrier = query_barrier ? 1 : 0;
  hstore(m->h_flag, DEVICE_SCHEDULER_INIT);
  hstore(m->h_persistent, -1);
  hstore(m->h_stop, 0);
  hstore(m->h_participating, 0);
  m->response_us.clear();
  m->exec_us.clear();
  m->samples.clear();
  m->errors = 0;
  m->redlines = 0;
  m->persistent_ms = 0;
  m->np_groups = 0;
  rc = wgm_reset_ctl(rt, 0, 0);
  if (rc) return rc;
  rc = mega_enqueue_sched_reset(m);
  if (rc) return rc;
  HIPCHK(rt, hipEventRecord(rt->ev0, rt->stream));
  hipLaunchKernelGGL(k_mega, dim3(grid), dim3(wg_size), mega_lds_bytes(wg_size), rt->stream, rt->ctl, m->sd,
                     (const MegaApp*)m->h_app_dev, m->np);
  HIPCHK(rt, hipGetLastError());
  m->running = true;
  if (!hwait(m->h_flag, DEVICE_WAITING, rt->timeout_s)) return fail(rt, WGM_ERR_TIMEOUT, "mega-kernel never reached DEVICE_WAITING");
  m->participating = hload(m->h_participating);
  if (participating) *participating = m->participating;
  return WGM_OK;
}

// A persistent task: BFS / SSSP (frontier), colour, MIS or PageRank. Its
// descriptor goes to coherent host memory; the initial team resets the
// task's state on the device.
int wgm_mega_send_persistent_task(wgm_mega* m, wgm_graph* g, const wgm_persistent_task* t, int32_t groups) {
  if (!m || !g || !t) return fail(m ? m->rt : nullptr, WGM_ERR_INVALID, "NULL argument");
  wgm_runtime* rt = m->rt;
  if (!m->running) return fail(rt, WGM_ERR_STATE, "mega-kernel not running");
  if (groups < 1 || groups > m->participating)
    return fail(rt, WGM_ERR_TOO_MANY_GROUPS, "cannot send a persistent task with %d workgroups (%d participate)",
                groups, m->participating);
  const int kind = t->kind;
  if (kind < WGM_PERSIST_BFS || kind > WGM_PERSIST_PAGERANK) return fail(rt, WGM_ERR_INVALID, "unknown persistent kind");
  if (kind == WGM_PERSIST_SSSP && !g->wt) return fail(rt, WGM_ERR_INVALID, "SSSP needs weights");
  if ((kind == WGM_PERSIST_COLOR || kind == WGM_PERSIST_MIS) && !g->prio)
    return fail(rt, WGM_ERR_INVALID, "colouring and MIS need vertex priorities");
  if (kind <= WGM_PERSIST_SSSP && (t->src < 0 || t->src >= g->n)) return fail(rt, WGM_ERR_INVALID, "bad source");
  if (kind == WGM_PERSIST_PAGERANK && (t->iterations < 1 || !(t->damping >= 0.0f && t->damping <= 1.0f)))
    return fail(rt, WGM_ERR_INVALID, "PageRank needs iterations >= 1 and damping in [0, 1]");
  if (m->monitor.joinable()) m->monitor.join();
  FrontierDev f{};
  int rc = wgm_graph_frontier_state(g, kind == WGM_PERSIST_SSSP, &f);
  if (!rc) rc = wgm_graph_app_state(g, kind == WGM_PERSIST_COLOR ? 1 : kind == WGM_PERSIST_MIS ? 2 :
                                          kind == WGM_PERSIST_PAGERANK ? 3 : 0);
  if (rc) return rc;
  if (!m->claims) {
    HIPCHK(rt, hipMalloc(&m->claims, 3 * kLine * sizeof(int)));
    HIPCHK(rt, hipMemset(m->claims, 0, 3 * kLine * sizeof(int)));
  }
  f.claims = (int*)m->claims;
  f.chunk = m->wg_size;
  MegaApp app{};
  app.kind = kind;
  app.src = t->src;
  app.g = wgm_graph_view(g);
  app.f = f;
  app.iters = g->iters;
  app.ia = g->i32a;
  app.ib = g->i32b;
  app.fa = g->f0;
  app.fb = g->f1;
  app.fc = g->f2;
  app.flags = f.counts;
  app.claims = (int*)m->claims;
  app.damping = t->damping;
  app.base = (1.0f - t->damping) / (float)g->n;
  app.r0 = 1.0f / (float)g->n;
  app.iterations = t->iterations;
  std::memcpy((void*)m->h_app, &app, sizeof(app));
  std::atomic_thread_fence(std::memory_order_release);
  m->graph = g;
  m->app_kind = kind;
  return mega_start_persistent(m, groups);
}

int wgm_mega_send_persistent(wgm_mega* m, wgm_graph* g, int32_t kind, int32_t src, int32_t groups) {
  const wgm_persistent_task t{kind, src, 0.85f, 50};
  return wgm_mega_send_persistent_task(m, g, &t, groups);
}

}  // extern "C"

// send_persistent_task + monitor_persistent_task
// (src/scheduler_rt/rt_host/cl_communicator.h:168-224): assign the persistent
// task's workgroups, then a monitor thread samples the executing ones.
int mega_start_persistent(wgm_mega* m, int groups) {
  wgm_runtime* rt = m->rt;
  hstore(m->h_task_size, groups);
  hstore(m->h_flag, DEVICE_TO_PERSISTENT_TASK);
  if (!hwait(m->h_flag, DEVICE_GOT_GROUPS, rt->timeout_s)) return fail(rt, WGM_ERR_TIMEOUT, "no DEVICE_GOT_GROUPS");
  m->persist_begin = now_ns();
  m->persist_running = true;
  hstore(m->h_flag, DEVICE_TO_EXECUTE);
  if (!hwait(m->h_flag, DEVICE_WAITING, rt->timeout_s)) return fail(rt, WGM_ERR_TIMEOUT, "no DEVICE_WAITING");
  m->monitor = std::thread([m]() {
    while (true) {
      const int gcount = hload(m->h_persistent);
      m->samples.push_back(std::make_pair((double)(now_ns() - m->persist_begin) * 1e-6, gcount));
      if (gcount == 0) break;
      std::this_thread::sleep_for(std::chrono::microseconds(m->sample_us));
    }
    m->persist_end = now_ns();
    m->persistent_ms = (double)(m->persist_end - m->persist_begin) * 1e-6;
    m->persist_running = false;
  });
  return WGM_OK;
}

extern "C" {

int wgm_mega_persistent_running(wgm_mega* m) { return (m && m->persist_running.load()) ? 1 : 0; }

// send_task_synchronous (src/scheduler_rt/rt_host/cl_communicator.h:123-166):
// response = TO_TASK -> GOT_GROUPS, execution = TO_EXECUTE -> WAITING.
int wgm_mega_send_task_kind(wgm_mega* m, int32_t kind, int32_t groups, wgm_task_times* out) {
  if (!m) return fail(nullptr, WGM_ERR_INVALID, "NULL mega");
  wgm_runtime* rt = m->rt;
  if (!m->running) return fail(rt, WGM_ERR_STATE, "mega-kernel not running");
  if (kind < WGM_NP_MATMULT_I32 || kind > WGM_NP_REDUCE_MIN) return fail(rt, WGM_ERR_INVALID, "unknown task kind %d", kind);
  // the persistent master is never given up, so one workgroup stays with it
  const int limit = m->persist_running ? m->participating - 1 : m->participating;
  if (groups < 1 || groups > limit)
    return fail(rt, WGM_ERR_TOO_MANY_GROUPS, "cannot send a task with %d workgroups (%d available to tasks)", groups,
                limit);
  return mega_run_task(m, kind, groups, out);
}

int wgm_mega_send_task(wgm_mega* m, int32_t groups, wgm_task_times* out) {
  if (!m) return fail(nullptr, WGM_ERR_INVALID, "NULL mega");
  return wgm_mega_send_task_kind(m, m->np_kind, groups, out);
}

int wgm_mega_wait_persistent(wgm_mega* m, double timeout_s, double* persistent_ms) {
  if (!m) return fail(nullptr, WGM_ERR_INVALID, "NULL mega");
  const uint64_t t0 = now_ns();
  while (m->persist_running) {
    if ((double)(now_ns() - t0) * 1e-9 > timeout_s) return fail(m->rt, WGM_ERR_TIMEOUT, "persistent task still running");
    std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
  if (m->monitor.joinable()) m->monitor.join();
  if (persistent_ms) *persistent_ms = m->persistent_ms;
  return WGM_OK;
}

// send_quit_signal + finish.
int wgm_mega_quit(wgm_mega* m) {
  if (!m) return fail(nullptr, WGM_ERR_INVALID, "NULL mega");
  wgm_runtime* rt = m->rt;
  if (!m->running) return WGM_OK;
  hstore(m->h_stop, 1);
  hstore(m->h_flag, DEVICE_TO_QUIT);
  HIPCHK(rt, hipEventRecord(rt->ev1, rt->stream));
  // bounded: a mega-kernel that never drains is reported, not waited on forever
  const uint64_t t0 = now_ns();
  hipError_t q;
  while ((q = hipEventQuery(rt->ev1)) == hipErrorNotReady) {
    if ((double)(now_ns() - t0) * 1e-9 > rt->timeout_s)
      return fail(rt, WGM_ERR_TIMEOUT, "mega-kernel did not drain after DEVICE_TO_QUIT");
    std::this_thread::yield();
  }
  if (q != hipSuccess) return wgm_hip_fail(rt, q, "mega-kernel");
  m->running = false;
  if (m->monitor.joinable()) m->monitor.join();
  return WGM_OK;
}

int wgm_mega_get_stats(wgm_mega* m, wgm_mega_stats* st) {
  if (!m || !st) return fail(m ? m->rt : nullptr, WGM_ERR_INVALID, "NULL argument");
  wgm_mega_stats s{};
  s.participating = m->participating;
  s.tasks = (int32_t)m->response_us.size();
  s.np_groups = m->np_groups;
  s.errors = m->errors;
  s.redlines = m->redlines;
  s.persistent_ms = m->persistent_ms;
  s.response_mean_us = mean(m->response_us);
  s.response_p50_us = pct(m->response_us, 0.5);
  s.response_p99_us = pct(m->response_us, 0.99);
  s.response_max_us = m->response_us.empty() ? 0 : *std::max_element(m->response_us.begin(), m->response_us.end());
  s.exec_mean_us = mean(m->exec_us);
  s.exec_p50_us = pct(m->exec_us, 0.5);
  s.exec_p99_us = pct(m->exec_us, 0.99);
  s.exec_max_us = m->exec_us.empty() ? 0 : *std::max_element(m->exec_us.begin(), m->exec_us.end());
  std::vector<double> tot(m->exec_us.size());
  for (size_t i = 0; i < tot.size(); i++) tot[i] = m->exec_us[i] + m->response_us[i];
  s.total_mean_us = mean(tot);
  s.total_p50_us = pct(tot, 0.5);
  s.total_p99_us = pct(tot, 0.99);
  double gs = 0;
  for (auto& x : m->samples) gs += x.second;
  s.avg_persistent_groups = m->samples.empty() ? 0.0 : gs / (double)m->samples.size();
  *st = s;
  return WGM_OK;
}

// check_persistent_task: the result is read after the mega-kernel has quit
// (a copy may need the GPU, which the resident mega-kernel holds).
int wgm_mega_result(wgm_mega* m, void* out, int32_t* levels) {
  if (!m || !m->graph) return fail(m ? m->rt : nullptr, WGM_ERR_STATE, "no persistent task has run");
  wgm_runtime* rt = m->rt;
  if (m->running) return fail(rt, WGM_ERR_STATE, "read results after wgm_mega_quit");
  const wgm_graph* g = m->graph;
  const void* src = m->app_kind == WGM_PERSIST_COLOR ? (const void*)g->i32a
                    : m->app_kind == WGM_PERSIST_MIS ? (const void*)g->i32b
                    : m->app_kind == WGM_PERSIST_PAGERANK ? (const void*)g->f2 : (const void*)g->dist;
  if (out) HIPCHK(rt, hipMemcpy(out, src, 4 * (size_t)g->n, hipMemcpyDeviceToHost));
  if (levels) HIPCHK(rt, hipMemcpy(levels, m->graph->iters, sizeof(int), hipMemcpyDeviceToHost));
  return WGM_OK;
}

// The tasks' resident buffers, copied to the host so that a caller can check
// them against its own restatement (the reference's host keeps its copies of
// A, B and C for check_non_persistent_task, matmult.h:120-149,
// reduce.h:18-35). The resident kernel must have quit: a copy may need the GPU.
int wgm_np_copy(wgm_mega* m, int32_t kind, int32_t which, void* host, int64_t bytes) {
  if (!m || !host) return fail(m ? m->rt : nullptr, WGM_ERR_INVALID, "NULL argument");
  wgm_runtime* rt = m->rt;
  if (m->running) return fail(rt, WGM_ERR_STATE, "copy task buffers after wgm_mega_quit");
  const NPDev& t = m->np;
  const void* src = nullptr;
  int64_t want = (int64_t)t.dim * t.dim * 4;
  if (kind == WGM_NP_MATMULT_I32) src = which == WGM_NP_IN_A ? (const void*)t.A : which == WGM_NP_IN_B ? (const void*)t.B
                                        : which == WGM_NP_OUT_C ? (const void*)t.C : nullptr;
  else if (kind == WGM_NP_MATMUL_F32) src = which == WGM_NP_IN_A ? (const void*)t.Af : which == WGM_NP_IN_B ? (const void*)t.Bf
                                        : which == WGM_NP_OUT_C ? (const void*)t.Cf : nullptr;
  else if (kind == WGM_NP_REDUCE_MIN && which == WGM_NP_IN_A) {
    src = t.buf;
    want = (int64_t)t.len * 4;
  }
  if (!src) return fail(rt, WGM_ERR_INVALID, "task kind %d has no buffer %d", kind, which);
  if (bytes != want) return fail(rt, WGM_ERR_INVALID, "buffer is %lld bytes, %lld asked", (long long)want, (long long)bytes);
  HIPCHK(rt, hipMemcpy(host, src, (size_t)bytes, hipMemcpyDeviceToHost));
  return WGM_OK;
}

// The summary the last task published (djb2 hash, minimum, or the FP32
// checksum with value = 1), and the values the library checks it against.
int wgm_np_result(wgm_mega* m, int32_t* value, double* checksum) {
  if (!m) return fail(nullptr, WGM_ERR_INVALID, "NULL mega");
  if (value) *value = hload(m->h_np_result);
  if (checksum) {
    unsigned long long bits = __hip_atomic_load(m->h_np_checksum, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    std::memcpy(checksum, &bits, 8);
  }
  return WGM_OK;
}

int wgm_np_expected(wgm_mega* m, int32_t kind, int32_t* value, double* checksum, double* tol) {
  if (!m) return fail(nullptr, WGM_ERR_INVALID, "NULL mega");
  if (kind < WGM_NP_MATMULT_I32 || kind > WGM_NP_REDUCE_MIN) return fail(m->rt, WGM_ERR_INVALID, "unknown task kind %d", kind);
  if (value) *value = kind == WGM_NP_MATMULT_I32 ? m->np_expect_hash : kind == WGM_NP_REDUCE_MIN ? m->np_expect_min : 1;
  if (checksum) *checksum = kind == WGM_NP_MATMUL_F32 ? m->np_expect_sum : 0.0;
  if (tol) *tol = kind == WGM_NP_MATMUL_F32 ? m->np_tol : 0.0;
  return WGM_OK;
}

// run_non_persistent (src/host_skeleton/skel.cpp:168-221): the task alone,
// `iterations` launches of wg_size x num_wgs, each checked.
int wgm_np_run_solo(wgm_mega* m, int32_t wg_size, int32_t num_wgs, int32_t iterations, double* mean_us, int32_t* errors) {
  if (!m || iterations < 1 || num_wgs < 1) return fail(m ? m->rt : nullptr, WGM_ERR_INVALID, "bad arguments");
  wgm_runtime* rt = m->rt;
  int rc = wgm_check_wg(rt, wg_size);
  if (rc) return rc;
  if (m->running) return fail(rt, WGM_ERR_STATE, "the mega-kernel is resident; quit it first");
  double total = 0;
  int errs = 0;
  for (int i = 0; i < iterations; i++) {
    rc = mega_np_reset(m, m->np_kind, true);
    if (rc) return rc;
    HIPCHK(rt, hipStreamSynchronize(rt->stream));
    const uint64_t t0 = now_ns();
    hipLaunchKernelGGL(k_np_solo, dim3(num_wgs), dim3(wg_size), mega_lds_bytes(wg_size), rt->stream, m->np, m->np_kind);
    HIPCHK(rt, hipGetLastError());
    HIPCHK(rt, hipStreamSynchronize(rt->stream));
    total += (double)(now_ns() - t0) * 1e-3;
    if (!mega_np_check(m, m->np_kind)) errs++;
  }
  if (mean_us) *mean_us = total / iterations;
  if (errors) *errors = errs;
  return WGM_OK;
}

}  // extern "C"
