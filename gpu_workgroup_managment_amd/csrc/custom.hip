// custom.hip -- launching the caller's own persistent kernels, written
// against include/wgm_device_api.h: the standalone form and the merged
// mega-kernel form, with arguments appended in order through wgm_args
// (in place of the reference's cl::Kernel::setArg with a running index,
// src/lonestar/graph_apps/bfs/bfs.h:265-283 and skel.cpp:300-320).

#include "sched_internal.hpp"

#include <cstring>
#include <vector>

#include "../../include/wgm_device_api.h"

using namespace wgm;

struct wgm_args {
  std::vector<std::vector<unsigned char>> values;
};

#define fail wgm_fail
#define HIPCHK WGM_HIPCHK

namespace {

// Occupancy-bounded grid for a caller's kernel: the requested count, or the
// workgroups that a discovery probe of the kernel itself finds co-resident.
int custom_grid(wgm_runtime* rt, const void* kernel, int wg_size, int lds, int requested, void** params, int* grid) {
  int rc = wgm_check_wg(rt, wg_size);
  if (rc) return rc;
  int per_cu = 0;
  HIPCHK(rt, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, wg_size, lds));
  if (per_cu < 1) return fail(rt, WGM_ERR_INVALID, "the kernel cannot be resident at wg_size %d", wg_size);
  if (requested > 0) {
    if (requested > rt->max_grid) return fail(rt, WGM_ERR_INVALID, "grid %d exceeds %d", requested, rt->max_grid);
    *grid = requested;
    return WGM_OK;
  }
  int probe = std::min(2 * per_cu * rt->info.num_cus, rt->max_grid);
  rc = wgm_reset_ctl(rt, 1, 0);
  if (rc) return rc;
  HIPCHK(rt, hipLaunchKernel(kernel, dim3(probe), dim3(wg_size), params, lds, rt->stream));
  HIPCHK(rt, hipStreamSynchronize(rt->stream));
  int c = -1;
  HIPCHK(rt, hipMemcpy(&c, rt->ctl.count, sizeof(int), hipMemcpyDeviceToHost));
  if (c < 1) return fail(rt, WGM_ERR_INVALID, "discovery found no workgroups for the kernel");
  *grid = c;
  return WGM_OK;
}

// bounded wait for the runtime stream
int wait_stream(wgm_runtime* rt, const char* what) {
  HIPCHK(rt, hipEventRecord(rt->ev1, rt->stream));
  const uint64_t t0 = now_ns();
  hipError_t q;
  while ((q = hipEventQuery(rt->ev1)) == hipErrorNotReady) {
    if ((double)(now_ns() - t0) * 1e-9 > rt->timeout_s) return fail(rt, WGM_ERR_TIMEOUT, "%s did not finish", what);
    std::this_thread::yield();
  }
  if (q != hipSuccess) return wgm_hip_fail(rt, q, what);
  return WGM_OK;
}

}  // namespace

extern "C" {

int wgm_args_create(wgm_args** out) {
  if (!out) return WGM_ERR_INVALID;
  *out = new wgm_args();
  return WGM_OK;
}

int wgm_args_push(wgm_args* a, const void* value, int32_t bytes) {
  if (!a || !value || bytes < 1 || bytes > 4096) return WGM_ERR_INVALID;
  const unsigned char* p = (const unsigned char*)value;
  a->values.emplace_back(p, p + bytes);
  return WGM_OK;
}

int wgm_args_destroy(wgm_args* a) {
  delete a;
  return WGM_OK;
}

int wgm_launch_standalone(wgm_runtime* rt, const void* kernel, int32_t wg_size, int32_t num_wgs, int32_t lds_bytes,
                          const wgm_args* args, int32_t handshake, wgm_run_stats* st) {
  if (!rt || !kernel) return fail(rt, WGM_ERR_INVALID, "NULL argument");
  wgm_kernel_args wa{};
  std::vector<void*> params;
  params.push_back(&wa);
  if (args)
    for (const auto& v : args->values) params.push_back((void*)v.data());
  wa.ctl = rt->ctl;
  wa.ctl.im_exit = 1;
  int grid = 0;
  int rc = custom_grid(rt, kernel, wg_size, lds_bytes, num_wgs, params.data(), &grid);
  if (rc) return rc;
  rc = wgm_reset_ctl(rt, 0, handshake ? 1 : 0);
  if (rc) return rc;
  wa.ctl = rt->ctl;
  HIPCHK(rt, hipEventRecord(rt->ev0, rt->stream));
  HIPCHK(rt, hipLaunchKernel(kernel, dim3(grid), dim3(wg_size), params.data(), lds_bytes, rt->stream));
  double host_ms = 0.0;
  if (handshake) {
    rc = wgm_host_handshake(rt, &host_ms);
    if (rc) return rc;
  }
  rc = wait_stream(rt, "standalone kernel");
  if (rc) return rc;
  wgm_run_stats s{};
  rc = wgm_finish_launch(rt, &s);
  if (rc) return rc;
  s.host_ms = host_ms;
  if (st) *st = s;
  return WGM_OK;
}

int wgm_discover_occupancy_custom(wgm_runtime* rt, const void* kernel, int32_t wg_size, int32_t lds_bytes,
                                  const wgm_args* args, int32_t mega, int32_t* count) {
  if (!rt || !kernel || !count) return fail(rt, WGM_ERR_INVALID, "NULL argument");
  wgm_kernel_args wa{};
  wgm_mega_args ma{};
  std::vector<void*> params;
  params.push_back(mega ? (void*)&ma : (void*)&wa);
  if (args)
    for (const auto& v : args->values) params.push_back((void*)v.data());
  wa.ctl = rt->ctl;
  wa.ctl.im_exit = 1;
  ma.ctl = wa.ctl;  // im_exit: every workgroup returns right after discovery
  return custom_grid(rt, kernel, wg_size, lds_bytes, 0, params.data(), count);
}

int wgm_mega_set_query_barrier(wgm_mega* m, int32_t query_barrier) {
  if (!m) return fail(nullptr, WGM_ERR_INVALID, "NULL mega");
  if (m->running) return fail(m->rt, WGM_ERR_STATE, "set the barrier form before the launch");
  m->custom_query_barrier = query_barrier ? 1 : 0;
  return WGM_OK;
}

int wgm_mega_launch_custom(wgm_mega* m, const void* kernel, int32_t wg_size, int32_t num_wgs, int32_t lds_bytes,
                           const wgm_args* args, int32_t* participating) {
  if (!m || !kernel) return fail(m ? m->rt : nullptr, WGM_ERR_INVALID, "NULL argument");
  wgm_runtime* rt = m->rt;
  if (m->running) return fail(rt, WGM_ERR_STATE, "mega-kernel already running");
  wgm_mega_args ma{};
  std::vector<void*> params;
  params.push_back(&ma);
  if (args)
    for (const auto& v : args->values) params.push_back((void*)v.data());
  ma.ctl = rt->ctl;
  ma.ctl.im_exit = 1;
  ma.sched = m->sd;
  int grid = 0;
  int rc = custom_grid(rt, kernel, wg_size, lds_bytes, num_wgs, params.data(), &grid);
  if (rc) return rc;
  if (grid < 2) return fail(rt, WGM_ERR_INVALID, "a mega-kernel needs at least 2 workgroups");
  m->wg_size = wg_size;
  m->sd.query_barrier = m->custom_query_barrier;
  m->sd.serve = 0;
  hstore(m->h_flag, DEVICE_SCHEDULER_INIT);
  hstore(m->h_persistent, -1);
  hstore(m->h_stop, 0);
  hstore(m->h_participating, 0);
  m->response_us.clear();
  m->exec_us.clear();
  m->samples.clear();
  m->errors = 0;
  m->persistent_ms = 0;
  rc = wgm_reset_ctl(rt, 0, 0);
  if (rc) return rc;
  rc = mega_enqueue_sched_reset(m);
  if (rc) return rc;
  ma.ctl = rt->ctl;
  ma.sched = m->sd;
  HIPCHK(rt, hipLaunchKernel(kernel, dim3(grid), dim3(wg_size), params.data(), lds_bytes, rt->stream));
  m->running = true;
  if (!hwait(m->h_flag, DEVICE_WAITING, rt->timeout_s)) return fail(rt, WGM_ERR_TIMEOUT, "mega-kernel never reached DEVICE_WAITING");
  m->participating = hload(m->h_participating);
  if (participating) *participating = m->participating;
  return WGM_OK;
}

// send_persistent_task for a caller's persistent task (its own arguments
// were given at launch).
int wgm_mega_send_persistent_custom(wgm_mega* m, int32_t groups) {
  if (!m) return fail(nullptr, WGM_ERR_INVALID, "NULL mega");
  wgm_runtime* rt = m->rt;
  if (!m->running) return fail(rt, WGM_ERR_STATE, "mega-kernel not running");
  if (groups < 1 || groups > m->participating)
    return fail(rt, WGM_ERR_TOO_MANY_GROUPS, "cannot send a persistent task with %d workgroups (%d participate)",
                groups, m->participating);
  if (m->monitor.joinable()) m->monitor.join();
  return mega_start_persistent(m, groups);
}

// send_task_synchronous for a caller's non-persistent task: no built-in
// result check (ok = 1); the caller checks its own outputs.
int wgm_mega_send_task_custom(wgm_mega* m, int32_t groups, wgm_task_times* out) {
  if (!m) return fail(nullptr, WGM_ERR_INVALID, "NULL mega");
  wgm_runtime* rt = m->rt;
  if (!m->running) return fail(rt, WGM_ERR_STATE, "mega-kernel not running");
  const int limit = m->persist_running ? m->participating - 1 : m->participating;
  if (groups < 1 || groups > limit)
    return fail(rt, WGM_ERR_TOO_MANY_GROUPS, "cannot send a task with %d workgroups (%d available to tasks)", groups,
                limit);
  return mega_run_task(m, -1, groups, out);
}

}  // extern "C"
