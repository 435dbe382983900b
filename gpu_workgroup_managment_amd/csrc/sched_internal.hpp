// sched_internal.hpp -- mega-kernel descriptors and host-side state.
#pragma once

#include <atomic>
#include <thread>
#include <utility>
#include <vector>

#include "apps_internal.hpp"
#include "sched_device.hpp"

namespace wgm {

// Non-persistent task descriptor (fixed for the life of a wgm_mega).
struct NPDev {
  int kind;  // WGM_NP_*
  int dim;
  int len;
  const int* A;
  const int* B;
  int* C;
  const float* Af;
  const float* Bf;
  float* Cf;
  const int* buf;
  int* dmin;
  int* counter;
  int* h_result;                   // coherent host cell
  unsigned long long* h_checksum;  // coherent host cell
};

// Persistent task descriptor, written by the host into coherent host memory
// before DEVICE_TO_PERSISTENT_TASK and read by each workgroup it is given to.
struct MegaApp {
  int kind;  // WGM_PERSIST_*
  int src;
  GraphDev g;
  FrontierDev f;
  int* iters;
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

// Dynamic LDS of the mega-kernel: the larger of the frontier layout, the
// FP32 matmul tiles (2 x 32 x 33 floats) and the FP64 checksum reduction.
__host__ __device__ constexpr int mega_lds_bytes(int bd) {
  return frontier_lds_bytes(bd) > 2 * 32 * 33 * 4
             ? (frontier_lds_bytes(bd) > bd * 8 ? frontier_lds_bytes(bd) : bd * 8)
             : (2 * 32 * 33 * 4 > bd * 8 ? 2 * 32 * 33 * 4 : bd * 8);
}

}  // namespace wgm

struct wgm_mega {
  wgm_runtime* rt = nullptr;
  wgm::SchedDev sd{};
  wgm::NPDev np{};
  wgm_np_config np_cfg{};
  int32_t np_expect_i = 0;
  double np_expect_f = 0.0;
  void* d_arena = nullptr;
  void* np_dev = nullptr;
  void* np_a = nullptr;
  void* np_b = nullptr;
  void* np_c = nullptr;
  void* claims = nullptr;
  int* h_cells = nullptr;
  int* h_flag = nullptr;
  int* h_task_size = nullptr;
  int* h_participating = nullptr;
  int* h_persistent = nullptr;
  int* h_stop = nullptr;
  int* h_np_result = nullptr;
  unsigned long long* h_np_checksum = nullptr;
  int* h_np_result_dev = nullptr;
  unsigned long long* h_np_checksum_dev = nullptr;
  wgm::MegaApp* h_app = nullptr;
  wgm::MegaApp* h_app_dev = nullptr;
  wgm_graph* graph = nullptr;
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
