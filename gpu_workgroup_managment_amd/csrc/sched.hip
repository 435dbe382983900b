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

__device__ __forceinline__ int pow2_ceil(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

// ================================================== non-persistent tasks
// matmult (src/non_persistent_kernels/matmult/device/matmult.cl:26-67):
// int32 C = A * B, elements strided over the task's workgroups; the last
// workgroup to finish hashes the diagonal (djb2) into the host cell.
__device__ void np_matmult_i32(const NPDev& t, int vid, int nwg) {
  const int dim = t.dim;
  const int c_size = dim * dim;
  for (int i = vid * blockDim.x + threadIdx.x; i < c_size; i += nwg * blockDim.x) {
    const int r = i / dim, col = i % dim;
    unsigned acc = 0;
    for (int j = 0; j < dim; j++) acc += (unsigned)t.A[r * dim + j] * (unsigned)t.B[j * dim + col];
    t.C[i] = (int)acc;
  }
}

// FP32 square matmul on 32x32 output tiles staged through LDS; each thread
// owns a 2x2 block of its tile (256 threads cover 32x32), accumulating in FP32.
__device__ void np_matmul_f32(const NPDev& t, int vid, int nwg, float* lds) {
  const int dim = t.dim;
  const int tiles = dim / 32;
  float* As = lds;              // [32][33]
  float* Bs = lds + 32 * 33;    // [32][33]
  for (int tile = vid; tile < tiles * tiles; tile += nwg) {
    const int tr = tile / tiles, tc = tile % tiles;
    float acc[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
    const int ty = (threadIdx.x / 16) * 2, tx = (threadIdx.x % 16) * 2;
    for (int k0 = 0; k0 < dim; k0 += 32) {
      for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
        const int r = e / 32, cc = e % 32;
        As[r * 33 + cc] = t.Af[(tr * 32 + r) * dim + k0 + cc];
        Bs[r * 33 + cc] = t.Bf[(k0 + r) * dim + tc * 32 + cc];
      }
      __syncthreads();
      if (threadIdx.x < 256) {
        for (int kk = 0; kk < 32; kk++) {
          const float a0 = As[ty * 33 + kk], a1 = As[(ty + 1) * 33 + kk];
          const float b0 = Bs[kk * 33 + tx], b1 = Bs[kk * 33 + tx + 1];
          acc[0][0] = __builtin_fmaf(a0, b0, acc[0][0]);
          acc[0][1] = __builtin_fmaf(a0, b1, acc[0][1]);
          acc[1][0] = __builtin_fmaf(a1, b0, acc[1][0]);
          acc[1][1] = __builtin_fmaf(a1, b1, acc[1][1]);
        }
      }
      __syncthreads();
    }
    if (threadIdx.x < 256) {
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) t.Cf[(tr * 32 + ty + a) * dim + tc * 32 + tx + b] = acc[a][b];
    }
  }
}

// MY_reduce (src/non_persistent_kernels/reduce/device/reduce.cl:4-41): LDS
// tree minimum per chunk of blockDim values, then one atomic min per chunk.
__device__ void np_reduce_min(const NPDev& t, int vid, int nwg, int* lds) {
  for (int base = vid * blockDim.x; base < t.len; base += nwg * blockDim.x) {
    const int i = base + threadIdx.x;
    lds[threadIdx.x] = i < t.len ? t.buf[i] : INT32_MAX;
    __syncthreads();
    for (int off = pow2_ceil(blockDim.x) / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off && (int)threadIdx.x + off < (int)blockDim.x)
        lds[threadIdx.x] = min(lds[threadIdx.x], lds[threadIdx.x + off]);
      __syncthreads();
    }
    if (threadIdx.x == 0) __hip_atomic_fetch_min(t.dmin, lds[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
  }
}

// Task body + completion: the workgroup that finishes last publishes the
// task's summary (hash / minimum / checksum) to the host cell with a
// system-scope release.
__device__ void np_task(const NPDev& t, int vid, int nwg, int* lds) {
  if (t.kind == WGM_NP_MATMULT_I32) np_matmult_i32(t, vid, nwg);
  else if (t.kind == WGM_NP_MATMUL_F32) np_matmul_f32(t, vid, nwg, (float*)lds);
  else np_reduce_min(t, vid, nwg, lds);
  drain_vmem();
  __syncthreads();
  __shared__ int s_last;
  if (threadIdx.x == 0) {
    fence_release();
    s_last = add_acq_rel(t.counter, 1) == nwg - 1;
  }
  __syncthreads();
  if (!s_last) return;
  if (t.kind == WGM_NP_MATMULT_I32) {
    if (threadIdx.x == 0) {
      unsigned h = 5381;
      for (int i = 0; i < t.dim; i++) h = h * 33u + (unsigned)t.C[i * t.dim + i];
      __hip_atomic_store(t.h_result, (int)h, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  } else if (t.kind == WGM_NP_REDUCE_MIN) {
    if (threadIdx.x == 0)
      __hip_atomic_store(t.h_result, ld_relaxed(t.dmin), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  } else {
    // checksum: sum of all C in FP64
    double* red = (double*)lds;
    double acc = 0.0;
    for (int i = threadIdx.x; i < t.dim * t.dim; i += blockDim.x) acc += (double)t.Cf[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = pow2_ceil(blockDim.x) / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off && (int)threadIdx.x + off < (int)blockDim.x) red[threadIdx.x] += red[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      unsigned long long bits;
      __builtin_memcpy(&bits, &red[0], 8);
      __hip_atomic_store(t.h_checksum, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(t.h_result, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
}

// Standalone launch of a non-persistent task (run_non_persistent,
// src/host_skeleton/skel.cpp:168-221).
__global__ void __launch_bounds__(1024) k_np_solo(NPDev t) {
  extern __shared__ int lds[];
  np_task(t, blockIdx.x, gridDim.x, lds);
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
    frontier_level<SSSP>(a.g, a.f, level, vid, 2, lds);
    if (resize_barrier(s, vid, RCtx{1, level})) return;
  }
  if (vid == 0 && threadIdx.x == 0) *a.iters = level;
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
    RCtx r;
    const int task = get_task(s, did, &r);
    if (task == TASK_QUIT) break;
    if (task == TASK_MULT) {
      np_task(np, s.np.group_ids[did], k_num_groups(s.np), lds);
      __syncthreads();
      if (threadIdx.x == 0) {
        fence_release();
        add_relaxed(s.np.executing, -1);
      }
    } else if (task == TASK_PERSIST) {
      const MegaApp app = read_app(app_cell);
      if (app.kind == WGM_PERSIST_SSSP) p_frontier<true>(s, app, r, did, lds);
      else p_frontier<false>(s, app, r, did, lds);
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
    s.r_ctx[i].target = 0;
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

namespace {

uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

int hload(const int* p) { return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM); }
void hstore(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM); }

// bounded spin on a host cell; yields like the reference's my_yield()
bool hwait(const int* p, int v, double timeout_s) {
  const uint64_t t0 = now_ns();
  uint64_t spins = 0;
  while (hload(p) != v) {
    if ((++spins & 255) == 0) {
      if ((double)(now_ns() - t0) * 1e-9 > timeout_s) return false;
      std::this_thread::yield();
    }
  }
  return true;
}

double pct(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t lo = (size_t)std::floor(pos), hi = (size_t)std::ceil(pos);
  return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

double mean(const std::vector<double>& v) {
  if (v.empty()) return 0.0;
  double s = 0;
  for (double x : v) s += x;
  return s / (double)v.size();
}

}  // namespace

bool wgm_sched_kernel_info(int kernel, KernelInfo* out) {
  if (kernel == WGM_K_MEGA) {
    *out = KernelInfo{(const void*)&k_mega, 1};
    return true;
  }
  return false;
}

int wgm_sched_probe(wgm_runtime* rt, int kernel, int grid, int wg_size) {
  if (kernel != WGM_K_MEGA) return wgm_fail(rt, WGM_ERR_INVALID, "no discovery probe for kernel %d", kernel);
  hipLaunchKernelGGL(k_mega, dim3(grid), dim3(wg_size), mega_lds_bytes(wg_size), rt->stream, rt->ctl, SchedDev{},
                     (const MegaApp*)nullptr, NPDev{});
  return WGM_OK;
}

void wgm_runtime_release_extras(wgm_runtime* rt) { (void)rt; }

// ---------------------------------------------------------------- wgm_mega
int mega_np_setup(wgm_mega* m, const wgm_np_config* cfg) {
  wgm_runtime* rt = m->rt;
  m->np_cfg = *cfg;
  NPDev& t = m->np;
  t = NPDev{};
  t.kind = cfg->kind;
  t.dim = cfg->dim > 0 ? cfg->dim : 256;
  t.len = cfg->length > 0 ? cfg->length : (1 << 20);
  if (t.kind == WGM_NP_MATMUL_F32 && t.dim % 32 != 0) return fail(rt, WGM_ERR_INVALID, "FP32 matmul needs dim % 32 == 0");
  HIPCHK(rt, hipMalloc(&m->np_dev, 4 * kLine * sizeof(int)));
  t.counter = (int*)m->np_dev;
  t.dmin = (int*)m->np_dev + kLine;
  t.h_result = m->h_np_result_dev;
  t.h_checksum = (unsigned long long*)(m->h_np_checksum_dev);
  const size_t nn = (size_t)t.dim * t.dim;
  if (t.kind == WGM_NP_MATMULT_I32) {
    // middle-square digit fill (src/non_persistent_kernels/matmult/matmult.h:38-50), seed 1234
    std::vector<int32_t> A(nn), B(nn), C(nn);
    int32_t seed = cfg->seed ? cfg->seed : 1234;
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
    m->np_expect_i = (int32_t)h;
    HIPCHK(rt, hipMalloc(&m->np_a, nn * 4));
    HIPCHK(rt, hipMalloc(&m->np_b, nn * 4));
    HIPCHK(rt, hipMalloc(&m->np_c, nn * 4));
    HIPCHK(rt, hipMemcpy(m->np_a, A.data(), nn * 4, hipMemcpyHostToDevice));
    HIPCHK(rt, hipMemcpy(m->np_b, B.data(), nn * 4, hipMemcpyHostToDevice));
    t.A = (const int*)m->np_a;
    t.B = (const int*)m->np_b;
    t.C = (int*)m->np_c;
  } else if (t.kind == WGM_NP_MATMUL_F32) {
    // A, B uniform in [-1, 1) from splitmix64 (seed cfg->seed)
    std::vector<float> A(nn), B(nn);
    uint64_t st = cfg->seed ? (uint64_t)cfg->seed : 6;
    for (size_t i = 0; i < nn; i++) A[i] = (float)((double)(wgm_splitmix64_next(&st) >> 11) * 0x1p-53 * 2.0 - 1.0);
    for (size_t i = 0; i < nn; i++) B[i] = (float)((double)(wgm_splitmix64_next(&st) >> 11) * 0x1p-53 * 2.0 - 1.0);
    // expected checksum sum(C) = sum_k (colsum A)_k (rowsum B)_k in FP64
    double sum = 0.0;
    for (int k = 0; k < t.dim; k++) {
      double ca = 0.0, rb = 0.0;
      for (int i = 0; i < t.dim; i++) ca += (double)A[(size_t)i * t.dim + k];
      for (int j = 0; j < t.dim; j++) rb += (double)B[(size_t)k * t.dim + j];
      sum += ca * rb;
    }
    m->np_expect_f = sum;
    HIPCHK(rt, hipMalloc(&m->np_a, nn * 4));
    HIPCHK(rt, hipMalloc(&m->np_b, nn * 4));
    HIPCHK(rt, hipMalloc(&m->np_c, nn * 4));
    HIPCHK(rt, hipMemcpy(m->np_a, A.data(), nn * 4, hipMemcpyHostToDevice));
    HIPCHK(rt, hipMemcpy(m->np_b, B.data(), nn * 4, hipMemcpyHostToDevice));
    t.Af = (const float*)m->np_a;
    t.Bf = (const float*)m->np_b;
    t.Cf = (float*)m->np_c;
  } else if (t.kind == WGM_NP_REDUCE_MIN) {
    // values in [1, 2^31 - 1] from splitmix64 (the reference uses rand() + 1,
    // src/non_persistent_kernels/reduce/reduce.h:18-35)
    std::vector<int32_t> v(t.len);
    uint64_t st = cfg->seed ? (uint64_t)cfg->seed : 7;
    int32_t mn = INT32_MAX;
    for (int i = 0; i < t.len; i++) {
      v[i] = 1 + (int32_t)((wgm_splitmix64_next(&st) >> 33) % 2147483646ULL);
      mn = std::min(mn, v[i]);
    }
    m->np_expect_i = mn;
    HIPCHK(rt, hipMalloc(&m->np_a, (size_t)t.len * 4));
    HIPCHK(rt, hipMemcpy(m->np_a, v.data(), (size_t)t.len * 4, hipMemcpyHostToDevice));
    t.buf = (const int*)m->np_a;
  } else {
    return fail(rt, WGM_ERR_INVALID, "unknown non-persistent task kind %d", t.kind);
  }
  return WGM_OK;
}

// reset_non_persistent, host part: the result cell. Inside the mega-kernel
// the scheduler workgroup resets the device counters itself; a solo launch
// (no mega-kernel resident) resets them with a kernel on the main stream.
int mega_np_reset(wgm_mega* m, bool solo) {
  wgm_runtime* rt = m->rt;
  hstore(m->h_np_result, m->np.kind == WGM_NP_MATMUL_F32 ? 0 : INT32_MIN);
  if (solo) {
    hipLaunchKernelGGL(k_np_reset, dim3(1), dim3(64), 0, rt->stream, m->np);
    HIPCHK(rt, hipGetLastError());
  }
  return WGM_OK;
}

// check_non_persistent_task: compare the published summary
bool mega_np_check(wgm_mega* m) {
  const int r = hload(m->h_np_result);
  if (m->np.kind == WGM_NP_MATMUL_F32) {
    if (r != 1) return false;
    unsigned long long bits = __hip_atomic_load(m->h_np_checksum, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    double got;
    std::memcpy(&got, &bits, 8);
    const double tol = 1e-3 * std::max(1.0, std::fabs(m->np_expect_f)) + 1e-2;
    return std::fabs(got - m->np_expect_f) <= tol;
  }
  return r == m->np_expect_i;
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
  hipError_t e = hipHostMalloc((void**)&m->h_cells, 16 * line_b, hipHostMallocCoherent | hipHostMallocMapped);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipHostMalloc(scheduler cells)"));
  std::memset(m->h_cells, 0, 16 * line_b);
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
  // device arena
  const int G = rt->max_groups;
  const size_t bytes = 16 * line_b + 2 * (size_t)G * line_b + (size_t)G * sizeof(RCtx) + 2 * (size_t)G * sizeof(int);
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
  p += 6 * line_b;
  s.task_array = (int*)p;
  p += (size_t)G * line_b;
  s.flags = (int*)p;
  p += (size_t)G * line_b;
  s.r_ctx = (RCtx*)p;
  p += (size_t)G * sizeof(RCtx);
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
  void* ptrs[] = {m->d_arena, m->np_dev, m->np_a, m->np_b, m->np_c, m->claims};
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
  if (m->np.kind == WGM_NP_MATMUL_F32 && wg_size < 256)
    return fail(rt, WGM_ERR_INVALID, "the FP32 matmul task needs wg_size >= 256");
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
  hipLaunchKernelGGL(k_sched_reset, dim3(64), dim3(256), 0, rt->stream, m->sd, rt->max_groups);
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

// Persistent BFS / SSSP task. Its descriptor goes to coherent host memory;
// the initial team resets the traversal state on the device.
int wgm_mega_send_persistent(wgm_mega* m, wgm_graph* g, int32_t kind, int32_t src, int32_t groups) {
  if (!m || !g) return fail(m ? m->rt : nullptr, WGM_ERR_INVALID, "NULL argument");
  wgm_runtime* rt = m->rt;
  if (!m->running) return fail(rt, WGM_ERR_STATE, "mega-kernel not running");
  if (groups < 1 || groups > m->participating)
    return fail(rt, WGM_ERR_TOO_MANY_GROUPS, "cannot send a persistent task with %d workgroups (%d participate)",
                groups, m->participating);
  if (kind != WGM_PERSIST_BFS && kind != WGM_PERSIST_SSSP) return fail(rt, WGM_ERR_INVALID, "unknown persistent kind");
  if (kind == WGM_PERSIST_SSSP && !g->wt) return fail(rt, WGM_ERR_INVALID, "SSSP needs weights");
  if (src < 0 || src >= g->n) return fail(rt, WGM_ERR_INVALID, "bad source");
  if (m->monitor.joinable()) m->monitor.join();
  FrontierDev f{};
  int rc = wgm_graph_frontier_state(g, kind == WGM_PERSIST_SSSP, &f);
  if (rc) return rc;
  if (!m->claims) HIPCHK(rt, hipMalloc(&m->claims, 3 * kLine * sizeof(int)));
  f.claims = (int*)m->claims;
  f.chunk = m->wg_size;
  MegaApp app{};
  app.kind = kind;
  app.src = src;
  app.g = wgm_graph_view(g);
  app.f = f;
  app.iters = g->iters;
  std::memcpy((void*)m->h_app, &app, sizeof(app));
  std::atomic_thread_fence(std::memory_order_release);
  m->graph = g;
  // monitor_persistent_task (src/scheduler_rt/rt_host/cl_communicator.h:168-205)
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

int wgm_mega_persistent_running(wgm_mega* m) { return (m && m->persist_running.load()) ? 1 : 0; }

// send_task_synchronous (src/scheduler_rt/rt_host/cl_communicator.h:123-166):
// response = TO_TASK -> GOT_GROUPS, execution = TO_EXECUTE -> WAITING.
int wgm_mega_send_task(wgm_mega* m, int32_t groups, wgm_task_times* out) {
  if (!m) return fail(nullptr, WGM_ERR_INVALID, "NULL mega");
  wgm_runtime* rt = m->rt;
  if (!m->running) return fail(rt, WGM_ERR_STATE, "mega-kernel not running");
  // the persistent master is never given up, so one workgroup stays with it
  const int limit = m->persist_running ? m->participating - 1 : m->participating;
  if (groups < 1 || groups > limit)
    return fail(rt, WGM_ERR_TOO_MANY_GROUPS, "cannot send a task with %d workgroups (%d available to tasks)", groups,
                limit);
  int rc = mega_np_reset(m, false);
  if (rc) return rc;
  m->np_groups = groups;
  hstore(m->h_task_size, groups);
  const uint64_t r0 = now_ns();
  hstore(m->h_flag, DEVICE_TO_TASK);
  if (!hwait(m->h_flag, DEVICE_GOT_GROUPS, rt->timeout_s)) return fail(rt, WGM_ERR_TIMEOUT, "no DEVICE_GOT_GROUPS");
  const uint64_t r1 = now_ns();
  hstore(m->h_flag, DEVICE_TO_EXECUTE);
  if (!hwait(m->h_flag, DEVICE_WAITING, rt->timeout_s)) return fail(rt, WGM_ERR_TIMEOUT, "no DEVICE_WAITING");
  const uint64_t r2 = now_ns();
  const bool ok = mega_np_check(m);
  if (!ok) m->errors++;
  m->response_us.push_back((double)(r1 - r0) * 1e-3);
  m->exec_us.push_back((double)(r2 - r1) * 1e-3);
  if (out) {
    out->response_us = (double)(r1 - r0) * 1e-3;
    out->execution_us = (double)(r2 - r1) * 1e-3;
    out->ok = ok ? 1 : 0;
  }
  return ok ? WGM_OK : fail(rt, WGM_ERR_MISMATCH, "non-persistent task result check failed");
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
int wgm_mega_result(wgm_mega* m, uint32_t* dist_out, int32_t* levels) {
  if (!m || !m->graph) return fail(m ? m->rt : nullptr, WGM_ERR_STATE, "no persistent task has run");
  wgm_runtime* rt = m->rt;
  if (m->running) return fail(rt, WGM_ERR_STATE, "read results after wgm_mega_quit");
  if (dist_out) HIPCHK(rt, hipMemcpy(dist_out, m->graph->dist, sizeof(uint32_t) * m->graph->n, hipMemcpyDeviceToHost));
  if (levels) HIPCHK(rt, hipMemcpy(levels, m->graph->iters, sizeof(int), hipMemcpyDeviceToHost));
  return WGM_OK;
}

// run_non_persistent (src/host_skeleton/skel.cpp:168-221): the task alone,
// `iterations` launches of wg_size x num_wgs, each checked.
int wgm_np_run_solo(wgm_mega* m, int32_t wg_size, int32_t num_wgs, int32_t iterations, double* mean_us, int32_t* errors) {
  if (!m || iterations < 1 || num_wgs < 1) return fail(m ? m->rt : nullptr, WGM_ERR_INVALID, "bad arguments");
  wgm_runtime* rt = m->rt;
  int rc = wgm_check_wg(rt, wg_size);
  if (rc) return rc;
  if (m->np.kind == WGM_NP_MATMUL_F32 && wg_size < 256) return fail(rt, WGM_ERR_INVALID, "FP32 matmul needs wg >= 256");
  if (m->running) return fail(rt, WGM_ERR_STATE, "the mega-kernel is resident; quit it first");
  double total = 0;
  int errs = 0;
  for (int i = 0; i < iterations; i++) {
    rc = mega_np_reset(m, true);
    if (rc) return rc;
    HIPCHK(rt, hipStreamSynchronize(rt->stream));
    const uint64_t t0 = now_ns();
    hipLaunchKernelGGL(k_np_solo, dim3(num_wgs), dim3(wg_size), mega_lds_bytes(wg_size), rt->stream, m->np);
    HIPCHK(rt, hipGetLastError());
    HIPCHK(rt, hipStreamSynchronize(rt->stream));
    total += (double)(now_ns() - t0) * 1e-3;
    if (!mega_np_check(m)) errs++;
  }
  if (mean_us) *mean_us = total / iterations;
  if (errors) *errors = errs;
  return WGM_OK;
}

}  // extern "C"
