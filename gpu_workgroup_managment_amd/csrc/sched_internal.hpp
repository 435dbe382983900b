// sched_internal.hpp -- mega-kernel descriptors and host-side state.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>
#include <utility>
#include <vector>

#include "apps_internal.hpp"
#include "np_tasks.hpp"
#include "sched_device.hpp"

namespace wgm {

// Persistent task descriptor, written by the host into coherent host memory
// before DEVICE_TO_PERSISTENT_TASK and read by each workgroup it is given to.
struct MegaApp {
  int kind;  // WGM_PERSIST_*
  int src;
  GraphDev g;
  FrontierDev f;
  int* iters;
  // topology-driven tasks (colour, MIS, PageRank): outputs and per-vertex
  // state, stop-flag lines (3, by round % 3) and claim lines (3, by phase % 3)
  int32_t* ia;   // colour / MIS state
  int32_t* ib;   // MIS membership (2 = in the set, 0 = not)
  float* fa;     // PageRank contributions, even iterations
  float* fb;     // PageRank contributions, odd iterations
  float* fc;     // PageRank result
  int* flags;
  int* claims;
  float damping, base, r0;
  int iterations;
};

// Copy the descriptor out of host memory with system-scope loads.
__device__ __forceinline__ MegaApp read_app(const MegaApp* cell) {
  MegaApp a;
  const unsigned long long* src = (const unsigned long long*)cell;
  unsigned long long* dst = (unsigned long long*)&a;
  for (size_t i = 0; i < sizeof(MegaApp) / 8; i++)
    dst[i] = __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  return a;
}
static_assert(sizeof(MegaApp) % 8 == 0, "descriptor copied in 8-byte words");

// Dynamic LDS of the mega-kernel: the larger of the frontier layout and the
// non-persistent tasks' tiles / reductions.
__host__ __device__ constexpr int mega_lds_bytes(int bd) {
  return frontier_lds_bytes(bd) > np_lds_bytes(bd) ? frontier_lds_bytes(bd) : np_lds_bytes(bd);
}

// ----------------------------------------------- host protocol helpers
inline uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

inline int hload(const int* p) { return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM); }
inline void hstore(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM); }

// bounded spin on a host cell; yields like the reference's my_yield()
inline bool hwait(const int* p, int v, double timeout_s) {
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

inline double pct(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t lo = (size_t)std::floor(pos), hi = (size_t)std::ceil(pos);
  return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

inline double mean(const std::vector<double>& v) {
  if (v.empty()) return 0.0;
  double s = 0;
  for (double x : v) s += x;
  return s / (double)v.size();
}


}  // namespace wgm

struct wgm_mega {
  wgm_runtime* rt = nullptr;
  wgm::SchedDev sd{};
  wgm::NPDev np{};
  wgm_np_config np_cfg{};
  int np_kind = 0;               // kind of wgm_mega_send_task
  int32_t np_expect_hash = 0;    // matmult djb2 diagonal hash
  int32_t np_expect_min = 0;     // reduction minimum
  double np_expect_sum = 0.0;    // FP32 matmul weighted checksum (FP64)
  double np_tol = 0.0;           // its FP32 error bound
  void* d_arena = nullptr;
  void* np_dev = nullptr;
  void* np_mem = nullptr;        // all task buffers
  void* claims = nullptr;
  int* h_cells = nullptr;
  int* h_flag = nullptr;
  int* h_task_size = nullptr;
  int* h_participating = nullptr;
  int* h_persistent = nullptr;
  int* h_stop = nullptr;
  int* h_task_kind = nullptr;
  unsigned long long* h_np_times = nullptr;
  int* h_np_result = nullptr;
  unsigned long long* h_np_checksum = nullptr;
  int* h_np_result_dev = nullptr;
  unsigned long long* h_np_checksum_dev = nullptr;
  wgm::MegaApp* h_app = nullptr;
  wgm::MegaApp* h_app_dev = nullptr;
  wgm_graph* graph = nullptr;
  int app_kind = 0;              // WGM_PERSIST_* of the last persistent task
  int custom_query_barrier = 0;  // barrier form of a caller's mega-kernel
  int wg_size = 256;
  bool running = false;
  int participating = 0;
  // persistent-task monitor (the reference's detached monitor thread)
  std::thread monitor;
  std::atomic<bool> persist_running{false};
  uint64_t persist_begin = 0, persist_end = 0;
  double persistent_ms = 0.0;
  int sample_us = 1000;
  std::vector<std::pair<double, int>> samples;  // (ms since start, executing persistent workgroups)
  // non-persistent task latencies
  std::vector<double> response_us, exec_us;
  int np_groups = 0;
  int errors = 0;
  int redlines = 0;
};

// shared with the query server (serve.hip)
int mega_run_task(wgm_mega* m, int kind, int groups, wgm_task_times* out);
int mega_enqueue_sched_reset(wgm_mega* m);
int mega_start_persistent(wgm_mega* m, int groups);
