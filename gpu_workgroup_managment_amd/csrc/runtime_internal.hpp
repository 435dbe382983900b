// runtime_internal.hpp -- host-side runtime state shared by the .hip units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>

#include "graph_apps.hpp"
#include "wgm.h"

struct wgm_extras;  // scheduler / batched / serving state (other units)

struct wgm_runtime {
  int device = 0;
  wgm_device_info info{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void* d_ctl = nullptr;
  size_t d_ctl_bytes = 0;
  int* h_cells = nullptr;      // host view of the coherent cells
  int* h_cells_dev = nullptr;  // device view of the same cells
  int max_groups = 0;          // flag lines (participants)
  int max_grid = 0;            // p_group_ids entries (launch grid)
  // Discovery quiescence window: only a grid larger than what is resident
  // (a probe) waits it out, once; a short window under-counts when the
  // dispatch brings workgroups up slowly (first use of a kernel, profiling).
  double window_us = 1000.0;
  double timeout_s = 60.0;
  wgm::Ctl ctl{};              // device view, passed to kernels
  std::map<int, int> occ_cache;   // HIP occupancy per (kernel, wg)
  std::map<int, int> disc_cache;  // discovered occupancy per (kernel, wg)
  std::string last_error;
  wgm_extras* extras = nullptr;

  int* host_sched_flag() { return h_cells; }
  int* host_persistent_flag() { return h_cells + wgm::kLine; }
  int* host_participating() { return h_cells + 2 * wgm::kLine; }
};

struct KernelInfo {
  const void* fn;
  int (*lds)(int wg_size);  // dynamic LDS bytes at a workgroup size (nullptr: none)
};

inline int wgm_frontier_lds(int wg) { return wgm::frontier_lds_bytes(wg); }

const KernelInfo& wgm_kernel_info(int kernel);
size_t wgm_kernel_lds(int kernel, int wg_size);
int wgm_hip_fail(wgm_runtime* rt, hipError_t e, const char* what);
int wgm_fail(wgm_runtime* rt, int code, const char* fmt, ...);
int wgm_check_wg(wgm_runtime* rt, int wg_size);
int wgm_reset_ctl(wgm_runtime* rt, int im_exit, int handshake);
int wgm_occupancy(wgm_runtime* rt, int kernel, int wg_size, int* per_cu);
int wgm_resolve_grid(wgm_runtime* rt, int kernel, int wg_size, int requested, int* grid);
int wgm_host_handshake(wgm_runtime* rt, double* host_ms);
int wgm_finish_launch(wgm_runtime* rt, wgm_run_stats* st);
bool wgm_wait_cell(const int* cell, int value, bool equal, double timeout_s);
// defined by the app unit (apps.hip): true / WGM_OK when `kernel` is one of its kernels
bool wgm_app_kernel_info(int kernel, KernelInfo* out);
int wgm_app_probe(wgm_runtime* rt, int kernel, int grid, int wg_size);
int wgm_enqueue_reset_ctl(wgm_runtime* rt, hipStream_t st);
// defined by the batched unit (batched.hip), which forwards to the scheduler unit
bool wgm_extra_kernel_info(int kernel, KernelInfo* out);
int wgm_discover_extra(wgm_runtime* rt, int kernel, int grid, int wg_size);
// defined by the scheduler unit (sched.hip), which forwards to the server unit
bool wgm_sched_kernel_info(int kernel, KernelInfo* out);
int wgm_sched_probe(wgm_runtime* rt, int kernel, int grid, int wg_size);
// defined by the query-server unit (serve.hip)
bool wgm_serve_kernel_info(int kernel, KernelInfo* out);
int wgm_serve_probe(wgm_runtime* rt, int kernel, int grid, int wg_size);
void wgm_runtime_release_extras(wgm_runtime* rt);
// defined by the no-allocation unit (noalloc.hip): per-device scratch address
int wgm_noalloc_init(wgm_runtime* rt, int device);

#define WGM_HIPCHK(rt, call)                                    \
  do {                                                          \
    hipError_t e_ = (call);                                     \
    if (e_ != hipSuccess) return wgm_hip_fail(rt, e_, #call);   \
  } while (0)
