// runtime.hip -- host side of the wgm runtime and its C ABI (include/wgm.h).
//
// Owns the control arenas (device lines for discovery / barrier / scheduler
// state, coherent host cells the host and the kernels both watch), the launch
// of persistent kernels, occupancy discovery and the barrier microbenchmark.
// Reference counterparts: CL_Execution / CL_Communicator / skel.cpp
// (src/scheduler_rt/rt_host/cl_execution.h:16-174,
// src/scheduler_rt/rt_host/cl_communicator.h:19-465, src/host_skeleton/skel.cpp:127-385).

#include "runtime_internal.hpp"

#include <hip/hip_cooperative_groups.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace wgm {

// ----------------------------------------------------------- reset kernels
__global__ void k_reset_ctl(Ctl c, int flag_lines, int gid_entries) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  for (int i = t; i < flag_lines; i += stride) c.flags[i * kLine] = 0;
  for (int i = t; i < gid_entries; i += stride) c.p_group_ids[i] = -1;
  for (int i = t; i < kNumLines; i += stride) c.lines[i * kLine] = 0;
  for (int i = t; i < kMaxXcd; i += stride) {
    c.xcd_count[i * kLine] = 0;
    c.xflags[i * kLine] = 0;
  }
  if (t == 0) {
    *c.registered = 0;
    *c.poll = 0;
    *c.count = -1;
    *c.bar_groups = 0;
    *c.bar_to_kill = 0;
    *c.sense_counter = 0;
    *c.sense_flag = 0;
    c.clock[0] = 0;
    c.clock[1] = 0;
  }
}

// --------------------------------------------------------- barrier bench
// os_freq (src/test_suite/os_freq/device/os_freq.cl:8-46): iterations x 10
// barriers. With `check`, every workgroup counts its arrivals in one word
// and after barrier b verifies (b+1)*n <= arrivals <= (b+2)*n. Every wait has
// a budget of `budget` spins (kBenchSpins unless WGM_BARRIER_SPIN_BUDGET says
// otherwise): a barrier that does not complete raises *abort and the kernel
// ends (its results are then void) instead of hanging.
constexpr unsigned kBenchSpins = 1u << 22;
__global__ void __launch_bounds__(1024, 8) k_barrier_bench(Ctl c, int iterations, int variant, int check,
                                                        int* errors, int* arrivals, int* abort, unsigned budget) {
  __shared__ int local_sense;
  Team t;
  if (!discover_team(c, &t) || c.im_exit) return;
  const int id = t.id, n = t.n;
  const SpinBounded sp{abort, budget};
  if (id == 0 && threadIdx.x == 0) st_relaxed(c.bar_groups, n);
  if (threadIdx.x == 0) local_sense = 0;
  global_barrier_disc(c, id, n, sp);
  clock_start(c, id);
  int b = 0;
  for (int i = 0; i < iterations; i++) {
    for (int j = 0; j < 10; j++, b++) {
      if (check && threadIdx.x == 0) add_relaxed(arrivals, 1);
      if (variant == 0) global_barrier_disc(c, id, n, sp);
      else if (variant == 2) xcd_barrier<true>(c, t, sp);
      else global_barrier_sense(c.sense_counter, c.sense_flag, c.bar_groups, &local_sense, sp);
      if (check && threadIdx.x == 0) {
        const int a = ld_relaxed(arrivals);
        if (a < (b + 1) * n || a > (b + 2) * n) add_relaxed(errors, 1);
      }
    }
  }
  clock_end(c, id);
}

// HIP's own grid barrier as the reference point: a cooperative launch (every
// workgroup co-resident by construction, no discovery) and
// cooperative_groups' grid sync, 10 x iterations of them, with the same epoch
// check and device-clock stamps as k_barrier_bench.
__global__ void __launch_bounds__(1024) k_coop_barrier(int iterations, int check, int* errors, int* arrivals,
                                                      unsigned long long* clock) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const int n = (int)gridDim.x;
  grid.sync();
  if (blockIdx.x == 0 && threadIdx.x == 0) clock[0] = wall_clock64();
  int b = 0;
  for (int i = 0; i < iterations; i++) {
    for (int j = 0; j < 10; j++, b++) {
      if (check && threadIdx.x == 0) add_relaxed(arrivals, 1);
      grid.sync();
      if (check && threadIdx.x == 0) {
        const int a = ld_relaxed(arrivals);
        if (a < (b + 1) * n || a > (b + 2) * n) add_relaxed(errors, 1);
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) clock[1] = wall_clock64();
}

// Plain streaming copy (16 bytes per lane per step): the calibration kernel
// for HBM traffic counters and the achievable-bandwidth reference.
__global__ void __launch_bounds__(256) k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace wgm

using namespace wgm;

// =============================================================== runtime
namespace {
thread_local char g_err_tls[512];
}  // namespace

int wgm_fail(wgm_runtime* rt, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (rt) rt->last_error = buf;
  std::snprintf(g_err_tls, sizeof(g_err_tls), "%s", buf);
  return code;
}
#define fail wgm_fail
#define HIPCHK WGM_HIPCHK

int wgm_hip_fail(wgm_runtime* rt, hipError_t e, const char* what) {
  return fail(rt, WGM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

const KernelInfo& wgm_kernel_info(int kernel) {
  static KernelInfo table[WGM_K_COUNT];
  static bool init = false;
  if (!init) {
    for (int k = 0; k < WGM_K_COUNT; k++) table[k] = KernelInfo{(const void*)&k_barrier_bench, nullptr};
    for (int k = 0; k < WGM_K_COUNT; k++) {
      KernelInfo ki{};
      if (wgm_app_kernel_info(k, &ki) || wgm_extra_kernel_info(k, &ki)) table[k] = ki;
    }
    init = true;
  }
  return table[kernel];
}

size_t wgm_kernel_lds(int kernel, int wg_size) {
  const KernelInfo& ki = wgm_kernel_info(kernel);
  return ki.lds ? (size_t)ki.lds(wg_size) : 0;
}

int wgm_check_wg(wgm_runtime* rt, int wg_size) {
  if (wg_size < 64 || wg_size > 1024 || wg_size % 64 != 0)
    return fail(rt, WGM_ERR_INVALID, "wg_size %d must be a multiple of 64 in [64, 1024]", wg_size);
  return WGM_OK;
}

int wgm_reset_ctl(wgm_runtime* rt, int im_exit, int handshake) {
  rt->ctl.im_exit = im_exit;
  rt->ctl.handshake = handshake;
  __hip_atomic_store(rt->host_sched_flag(), DEVICE_SCHEDULER_INIT, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(rt->host_persistent_flag(), -1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(rt->host_participating(), 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  hipLaunchKernelGGL(k_reset_ctl, dim3(64), dim3(256), 0, rt->stream, rt->ctl, rt->max_groups, rt->max_grid);
  HIPCHK(rt, hipGetLastError());
  return WGM_OK;
}

// Enqueue the control reset on any stream (used inside hipGraph capture).
int wgm_enqueue_reset_ctl(wgm_runtime* rt, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_ctl, dim3(64), dim3(256), 0, st, rt->ctl, rt->max_groups, rt->max_grid);
  HIPCHK(rt, hipGetLastError());
  return WGM_OK;
}

int wgm_occupancy(wgm_runtime* rt, int kernel, int wg_size, int* per_cu) {
  if (kernel < 0 || kernel >= WGM_K_COUNT) return fail(rt, WGM_ERR_INVALID, "unknown kernel id %d", kernel);
  int rc = wgm_check_wg(rt, wg_size);
  if (rc) return rc;
  const int key = kernel * 17 + wg_size / 64;
  auto it = rt->occ_cache.find(key);
  if (it != rt->occ_cache.end()) {
    *per_cu = it->second;
    return WGM_OK;
  }
  int nb = 0;
  HIPCHK(rt, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, wgm_kernel_info(kernel).fn, wg_size,
                                                           wgm_kernel_lds(kernel, wg_size)));
  rt->occ_cache[key] = nb;
  *per_cu = nb;
  return WGM_OK;
}

// Launch grid for a persistent kernel: the requested count, or the HIP
// occupancy bound (every workgroup co-resident) when none is given.
int wgm_resolve_grid(wgm_runtime* rt, int kernel, int wg_size, int requested, int* grid) {
  int per_cu = 0;
  int rc = wgm_occupancy(rt, kernel, wg_size, &per_cu);
  if (rc) return rc;
  if (per_cu < 1) return fail(rt, WGM_ERR_INVALID, "kernel %d cannot be resident at wg_size %d", kernel, wg_size);
  int g = requested;
  if (g <= 0) {
    // Default: the discovered occupancy (min(flag, occupancy) in the
    // reference, src/host_skeleton/skel.cpp:297-299). The HIP occupancy query
    // can overstate what is co-resident; a grid above the real occupancy would
    // only cost the discovery window.
    const int key = kernel * 17 + wg_size / 64;
    auto it = rt->disc_cache.find(key);
    if (it == rt->disc_cache.end()) {
      int found = 0;
      rc = wgm_discover_occupancy(rt, kernel, wg_size, 0, &found);
      if (rc) return rc;
      if (found < 1) return fail(rt, WGM_ERR_INVALID, "discovery found no workgroups for kernel %d", kernel);
      it = rt->disc_cache.emplace(key, found).first;
    }
    g = it->second;
  }
  if (g > rt->max_grid) return fail(rt, WGM_ERR_INVALID, "grid %d exceeds the discovery table (%d)", g, rt->max_grid);
  *grid = g;
  return WGM_OK;
}

// Bounded host wait on a coherent host cell.
bool wgm_wait_cell(const int* cell, int value, bool equal, double timeout_s) {
  const auto t0 = std::chrono::steady_clock::now();
  uint64_t spins = 0;
  while (true) {
    int v = __hip_atomic_load(cell, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    if ((v == value) == equal) return true;
    if ((++spins & 1023) == 0) {
      double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (el > timeout_s) return false;
    }
  }
}

// After an async launch of a standalone kernel with handshake on: wait for
// DEVICE_WAITING, flip to DEVICE_TO_PERSISTENT_TASK, time until every
// workgroup retired (src/host_skeleton/skel.cpp:336-340). On a timeout the
// flag is still flipped so the kernel can finish.
int wgm_host_handshake(wgm_runtime* rt, double* host_ms) {
  int rc = WGM_OK;
  if (!wgm_wait_cell(rt->host_sched_flag(), DEVICE_WAITING, true, rt->timeout_s))
    rc = fail(rt, WGM_ERR_TIMEOUT, "device never reached DEVICE_WAITING");
  const auto t0 = std::chrono::steady_clock::now();
  __hip_atomic_store(rt->host_sched_flag(), DEVICE_TO_PERSISTENT_TASK, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  if (rc == WGM_OK && !wgm_wait_cell(rt->host_persistent_flag(), 0, true, rt->timeout_s))
    rc = fail(rt, WGM_ERR_TIMEOUT, "persistent task did not retire");
  const auto t1 = std::chrono::steady_clock::now();
  if (host_ms) *host_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return rc;
}

int wgm_finish_launch(wgm_runtime* rt, wgm_run_stats* st) {
  HIPCHK(rt, hipEventRecord(rt->ev1, rt->stream));
  HIPCHK(rt, hipEventSynchronize(rt->ev1));
  float ms = 0.f;
  HIPCHK(rt, hipEventElapsedTime(&ms, rt->ev0, rt->ev1));
  unsigned long long clk[2] = {0, 0};
  HIPCHK(rt, hipMemcpy(clk, rt->ctl.clock, sizeof(clk), hipMemcpyDeviceToHost));
  int count = -1;
  HIPCHK(rt, hipMemcpy(&count, rt->ctl.count, sizeof(int), hipMemcpyDeviceToHost));
  if (st) {
    st->kernel_ms = ms;
    st->device_ms = (clk[1] >= clk[0] && clk[0]) ? (double)(clk[1] - clk[0]) * 1e3 / (double)rt->info.wall_clock_hz : 0.0;
    st->participating = count;
  }
  return WGM_OK;
}

extern "C" {

const char* wgm_strerror(int32_t status) {
  switch (status) {
    case WGM_OK: return "ok";
    case WGM_ERR_INVALID: return "invalid argument";
    case WGM_ERR_NOMEM: return "out of memory";
    case WGM_ERR_HIP: return "HIP error";
    case WGM_ERR_NODEV: return "no GPU";
    case WGM_ERR_TOO_MANY_GROUPS: return "more workgroups requested than discovered";
    case WGM_ERR_MISMATCH: return "result self-check failed";
    case WGM_ERR_TIMEOUT: return "timed out";
    case WGM_ERR_STATE: return "invalid protocol state";
    case WGM_ERR_BUDGET: return "HBM budget exceeded";
    case WGM_ERR_IO: return "file I/O error";
    case WGM_ERR_FORMAT: return "malformed file";
    default: return "unknown status";
  }
}

const char* wgm_last_error(const wgm_runtime* rt) { return rt ? rt->last_error.c_str() : g_err_tls; }

int wgm_create(int32_t device, wgm_runtime** out) {
  if (!out) return fail(nullptr, WGM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, WGM_ERR_NODEV, "no HIP device");
  if (device < 0 || device >= ndev) return fail(nullptr, WGM_ERR_INVALID, "device %d out of range", device);
  wgm_runtime* rt = new wgm_runtime();
  rt->device = device;
  auto bail = [&](int rc) { wgm_destroy(rt); return rc; };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return bail(wgm_hip_fail(nullptr, e, "hipSetDevice"));
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return bail(wgm_hip_fail(nullptr, e, "hipGetDeviceProperties"));
  std::snprintf(rt->info.name, sizeof(rt->info.name), "%s", prop.name);
  std::snprintf(rt->info.arch, sizeof(rt->info.arch), "%s", prop.gcnArchName);
  rt->info.num_cus = prop.multiProcessorCount;
  rt->info.max_wg_size = prop.maxThreadsPerBlock;
  rt->info.wavefront_size = prop.warpSize;
  rt->info.total_mem = (int64_t)prop.totalGlobalMem;
  rt->info.lds_per_cu = (int32_t)prop.maxSharedMemoryPerMultiProcessor;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0) khz = 100000;
  rt->info.wall_clock_hz = (int64_t)khz * 1000;
  if (const char* w = std::getenv("WGM_DISCOVERY_WINDOW_US")) rt->window_us = std::atof(w);
  if (const char* t = std::getenv("WGM_HOST_TIMEOUT_S")) rt->timeout_s = std::atof(t);

  // Up to 32 workgroups per CU can be resident (one wave each); the probe
  // grid may be twice that.
  rt->max_groups = rt->info.num_cus * 32;
  rt->max_grid = rt->max_groups * 2;
  const size_t line_b = kLine * sizeof(int);
  const size_t n_lines_fixed = 8 + kNumLines + 1 + 2 * kMaxXcd + 1;
  rt->d_ctl_bytes = n_lines_fixed * line_b + (size_t)rt->max_groups * line_b + (size_t)rt->max_grid * sizeof(int) +
                    (size_t)kMaxXcd * rt->max_groups * sizeof(int);
  e = hipMalloc(&rt->d_ctl, rt->d_ctl_bytes);
  if (e != hipSuccess) return bail(wgm_hip_fail(nullptr, e, "hipMalloc(control arena)"));
  e = hipMemset(rt->d_ctl, 0, rt->d_ctl_bytes);
  if (e != hipSuccess) return bail(wgm_hip_fail(nullptr, e, "hipMemset(control arena)"));
  char* p = (char*)rt->d_ctl;
  auto take_line = [&]() { int* r = (int*)p; p += line_b; return r; };
  rt->ctl.poll = take_line();
  rt->ctl.count = take_line();
  rt->ctl.bar_groups = take_line();
  rt->ctl.bar_to_kill = take_line();
  rt->ctl.sense_counter = take_line();
  rt->ctl.sense_flag = take_line();
  rt->ctl.clock = (unsigned long long*)take_line();
  take_line();  // spare
  rt->ctl.lines = (int*)p;
  p += kNumLines * line_b;
  take_line();  // guard
  rt->ctl.xcd_count = (int*)p;
  p += kMaxXcd * line_b;
  rt->ctl.xflags = (int*)p;
  p += kMaxXcd * line_b;
  rt->ctl.registered = take_line();
  rt->ctl.flags = (int*)p;
  p += (size_t)rt->max_groups * line_b;
  rt->ctl.xcd_members = (int*)p;
  p += (size_t)kMaxXcd * rt->max_groups * sizeof(int);
  rt->ctl.p_group_ids = (int*)p;
  rt->ctl.max_groups = rt->max_groups;
  rt->ctl.window_ticks = (unsigned)((double)rt->window_us * 1e-6 * (double)rt->info.wall_clock_hz);

  // Host-watched cells: coherent pinned host memory, one line each.
  e = hipHostMalloc((void**)&rt->h_cells, 16 * line_b, hipHostMallocCoherent | hipHostMallocMapped);
  if (e != hipSuccess) return bail(wgm_hip_fail(nullptr, e, "hipHostMalloc(coherent cells)"));
  std::memset(rt->h_cells, 0, 16 * line_b);
  int* hp = nullptr;
  e = hipHostGetDevicePointer((void**)&hp, rt->h_cells, 0);
  if (e != hipSuccess) return bail(wgm_hip_fail(nullptr, e, "hipHostGetDevicePointer"));
  rt->h_cells_dev = hp;
  rt->ctl.h_scheduler_flag = hp;
  rt->ctl.h_persistent_flag = hp + kLine;
  rt->ctl.h_participating = hp + 2 * kLine;

  e = hipStreamCreateWithFlags(&rt->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(wgm_hip_fail(nullptr, e, "hipStreamCreate"));
  // the device address of this GPU's no-allocation scratch array
  const int nrc = wgm_noalloc_init(rt, device);
  if (nrc) return bail(nrc);
  e = hipEventCreate(&rt->ev0);
  if (e == hipSuccess) e = hipEventCreate(&rt->ev1);
  if (e != hipSuccess) return bail(wgm_hip_fail(nullptr, e, "hipEventCreate"));
  *out = rt;
  return WGM_OK;
}

int wgm_destroy(wgm_runtime* rt) {
  if (!rt) return WGM_OK;
  (void)hipSetDevice(rt->device);
  if (rt->stream) {
    // Bounded like every other host wait: a kernel that never finishes
    // (a caller's persistent kernel that deadlocked) must not hang the
    // teardown. Its resources are then left to process exit.
    const auto t0 = std::chrono::steady_clock::now();
    while (hipStreamQuery(rt->stream) == hipErrorNotReady) {
      if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > rt->timeout_s) {
        std::fprintf(stderr, "wgm_destroy: work on the runtime stream did not finish in %.0f s; resources not freed\n",
                     rt->timeout_s);
        return WGM_ERR_TIMEOUT;
      }
      std::this_thread::yield();
    }
  }
  wgm_runtime_release_extras(rt);
  if (rt->ev0) (void)hipEventDestroy(rt->ev0);
  if (rt->ev1) (void)hipEventDestroy(rt->ev1);
  if (rt->stream) (void)hipStreamDestroy(rt->stream);
  if (rt->d_ctl) (void)hipFree(rt->d_ctl);
  if (rt->h_cells) (void)hipHostFree(rt->h_cells);
  delete rt;
  return WGM_OK;
}

int wgm_get_device_info(wgm_runtime* rt, wgm_device_info* out) {
  if (!rt || !out) return fail(rt, WGM_ERR_INVALID, "NULL argument");
  *out = rt->info;
  return WGM_OK;
}

int wgm_occupancy_query_hip(wgm_runtime* rt, int32_t kernel, int32_t wg_size, int32_t* blocks_per_cu,
                            int32_t* num_cus) {
  if (!rt || !blocks_per_cu || !num_cus) return fail(rt, WGM_ERR_INVALID, "NULL argument");
  int per_cu = 0;
  int rc = wgm_occupancy(rt, kernel, wg_size, &per_cu);
  if (rc) return rc;
  *blocks_per_cu = per_cu;
  *num_cus = rt->info.num_cus;
  return WGM_OK;
}

int wgm_discover_occupancy(wgm_runtime* rt, int32_t kernel, int32_t wg_size, int32_t probe_groups,
                           int32_t* count) {
  if (!rt || !count) return fail(rt, WGM_ERR_INVALID, "NULL argument");
  int per_cu = 0;
  int rc = wgm_occupancy(rt, kernel, wg_size, &per_cu);
  if (rc) return rc;
  int grid = probe_groups > 0 ? probe_groups : 2 * per_cu * rt->info.num_cus;
  if (grid > rt->max_grid) grid = rt->max_grid;
  if (grid < 1) return fail(rt, WGM_ERR_INVALID, "empty probe grid");
  rc = wgm_reset_ctl(rt, 1, 0);
  if (rc) return rc;
  // With im_exit set every kernel returns right after discovery; its other
  // arguments are never read.
  switch (kernel) {
    case WGM_K_DISCOVERY:
    case WGM_K_BARRIER_BENCH:
      hipLaunchKernelGGL(k_barrier_bench, dim3(grid), dim3(wg_size), 0, rt->stream, rt->ctl, 0, 0, 0,
                         (int*)nullptr, (int*)nullptr, (int*)nullptr, kBenchSpins);
      break;
    default: {
      int rc2 = wgm_app_probe(rt, kernel, grid, wg_size);
      if (rc2 == WGM_ERR_INVALID) rc2 = wgm_discover_extra(rt, kernel, grid, wg_size);
      if (rc2) return rc2;
      break;
    }
  }
  HIPCHK(rt, hipGetLastError());
  HIPCHK(rt, hipStreamSynchronize(rt->stream));
  int c = -1;
  HIPCHK(rt, hipMemcpy(&c, rt->ctl.count, sizeof(int), hipMemcpyDeviceToHost));
  *count = c;
  return WGM_OK;
}

int wgm_barrier_bench(wgm_runtime* rt, int32_t variant, int32_t wg_size, int32_t num_wgs, int32_t iterations,
                      wgm_barrier_result* out) {
  if (!rt || !out) return fail(rt, WGM_ERR_INVALID, "NULL argument");
  if (variant != WGM_BARRIER_FLAGS && variant != WGM_BARRIER_SENSE && variant != WGM_BARRIER_XCD &&
      variant != WGM_BARRIER_COOP)
    return fail(rt, WGM_ERR_INVALID, "unknown barrier variant %d", variant);
  if (iterations < 0) return fail(rt, WGM_ERR_INVALID, "negative iterations");
  int grid = 0;
  int rc = wgm_resolve_grid(rt, WGM_K_BARRIER_BENCH, wg_size, num_wgs, &grid);
  if (rc) return rc;
  int* d_aux = nullptr;
  HIPCHK(rt, hipMalloc(&d_aux, 3 * sizeof(int)));  // errors, arrivals, abort
  if (variant == WGM_BARRIER_COOP) {
    // The grid is never above what the discovery probe found co-resident for
    // k_barrier_bench (which uses more registers and LDS than this kernel),
    // and HIP rejects a cooperative grid it cannot make co-resident.
    int per_cu = 0;
    rc = wgm_occupancy(rt, WGM_K_BARRIER_BENCH, wg_size, &per_cu);
    if (rc == WGM_OK && num_wgs > 0) {
      int found = 0;
      rc = wgm_resolve_grid(rt, WGM_K_BARRIER_BENCH, wg_size, 0, &found);
      if (rc == WGM_OK && grid > found) rc = fail(rt, WGM_ERR_TOO_MANY_GROUPS, "cooperative grid %d above the %d found co-resident", grid, found);
    }
    auto run_coop = [&](int check, int iters, wgm_barrier_result* r) -> int {
      HIPCHK(rt, hipMemsetAsync(d_aux, 0, 2 * sizeof(int), rt->stream));
      HIPCHK(rt, hipMemsetAsync(rt->ctl.clock, 0, 2 * sizeof(unsigned long long), rt->stream));
      int* errs_p = d_aux;
      int* arr_p = d_aux + 1;
      unsigned long long* clk = rt->ctl.clock;
      void* params[] = {&iters, &check, &errs_p, &arr_p, &clk};
      HIPCHK(rt, hipEventRecord(rt->ev0, rt->stream));
      HIPCHK(rt, hipLaunchCooperativeKernel((const void*)&k_coop_barrier, dim3(grid), dim3(wg_size), params, 0,
                                            rt->stream));
      HIPCHK(rt, hipEventRecord(rt->ev1, rt->stream));
      HIPCHK(rt, hipEventSynchronize(rt->ev1));
      float ms = 0.f;
      HIPCHK(rt, hipEventElapsedTime(&ms, rt->ev0, rt->ev1));
      unsigned long long c[2] = {0, 0};
      HIPCHK(rt, hipMemcpy(c, rt->ctl.clock, sizeof(c), hipMemcpyDeviceToHost));
      int errs = 0;
      HIPCHK(rt, hipMemcpy(&errs, d_aux, sizeof(int), hipMemcpyDeviceToHost));
      r->participating = grid;
      r->barriers = iters * 10;
      r->epoch_errors = errs;
      r->kernel_ms = ms;
      r->device_ms = c[1] >= c[0] ? (double)(c[1] - c[0]) * 1e3 / (double)rt->info.wall_clock_hz : 0.0;
      r->ns_per_barrier = r->barriers ? r->device_ms * 1e6 / r->barriers : 0.0;
      return WGM_OK;
    };
    wgm_barrier_result checked{}, timed{};
    if (rc == WGM_OK) rc = run_coop(1, iterations < 10 ? iterations : 10, &checked);
    if (rc == WGM_OK) rc = run_coop(0, iterations, &timed);
    (void)hipFree(d_aux);
    if (rc) return rc;
    timed.epoch_errors = checked.epoch_errors;
    *out = timed;
    return WGM_OK;
  }
  unsigned budget = kBenchSpins;
  if (const char* b = std::getenv("WGM_BARRIER_SPIN_BUDGET"))
    if (std::atol(b) >= 1) budget = (unsigned)std::atol(b);
  auto run = [&](int check, int iters, wgm_barrier_result* r) -> int {
    HIPCHK(rt, hipMemsetAsync(d_aux, 0, 3 * sizeof(int), rt->stream));
    int rc2 = wgm_reset_ctl(rt, 0, 0);
    if (rc2) return rc2;
    HIPCHK(rt, hipEventRecord(rt->ev0, rt->stream));
    hipLaunchKernelGGL(k_barrier_bench, dim3(grid), dim3(wg_size), 0, rt->stream, rt->ctl, iters, variant, check,
                       d_aux, d_aux + 1, d_aux + 2, budget);
    HIPCHK(rt, hipGetLastError());
    wgm_run_stats st{};
    rc2 = wgm_finish_launch(rt, &st);
    if (rc2) return rc2;
    int aux[3] = {0, 0, 0};
    HIPCHK(rt, hipMemcpy(aux, d_aux, sizeof(aux), hipMemcpyDeviceToHost));
    if (aux[2])
      return fail(rt, WGM_ERR_TIMEOUT, "barrier variant %d: a wait exceeded %u spins (%d participants)", variant,
                  budget, st.participating);
    const int errs = aux[0];
    r->participating = st.participating;
    r->barriers = iters * 10;
    r->epoch_errors = errs;
    r->kernel_ms = st.kernel_ms;
    r->device_ms = st.device_ms;
    r->ns_per_barrier = r->barriers ? st.device_ms * 1e6 / r->barriers : 0.0;
    return WGM_OK;
  };
  wgm_barrier_result checked{}, timed{};
  rc = run(1, iterations < 10 ? iterations : 10, &checked);
  if (rc == WGM_OK) rc = run(0, iterations, &timed);
  (void)hipFree(d_aux);
  if (rc) return rc;
  timed.epoch_errors = checked.epoch_errors;
  *out = timed;
  return WGM_OK;
}

}  // extern "C"

extern "C" int wgm_copy_bench(wgm_runtime* rt, int64_t bytes, int32_t iterations, double* ms_per_copy, double* gbps) {
  if (!rt || bytes < 16 || iterations < 1) return wgm_fail(rt, WGM_ERR_INVALID, "bad copy bench arguments");
  const long long n = bytes / 16;
  void *src = nullptr, *dst = nullptr;
  WGM_HIPCHK(rt, hipMalloc(&src, (size_t)n * 16));
  hipError_t e = hipMalloc(&dst, (size_t)n * 16);
  if (e != hipSuccess) {
    (void)hipFree(src);
    return wgm_hip_fail(rt, e, "hipMalloc(copy dst)");
  }
  e = hipMemsetAsync(src, 1, (size_t)n * 16, rt->stream);
  const int grid = rt->info.num_cus * 16;
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, rt->stream, (const uint4*)src, (uint4*)dst, n);  // warm
  if (e == hipSuccess) e = hipEventRecord(rt->ev0, rt->stream);
  for (int i = 0; i < iterations && e == hipSuccess; i++)
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, rt->stream, (const uint4*)src, (uint4*)dst, n);
  if (e == hipSuccess) e = hipEventRecord(rt->ev1, rt->stream);
  if (e == hipSuccess) e = hipEventSynchronize(rt->ev1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, rt->ev0, rt->ev1);
  (void)hipFree(src);
  (void)hipFree(dst);
  if (e != hipSuccess) return wgm_hip_fail(rt, e, "copy bench");
  const double per = ms / iterations;
  if (ms_per_copy) *ms_per_copy = per;
  if (gbps) *gbps = per > 0 ? 2.0 * (double)n * 16 / (per * 1e6) : 0.0;
  return WGM_OK;
}
