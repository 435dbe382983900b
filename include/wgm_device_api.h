/*
 * wgm_device_api.h -- device API for persistent HIP kernels run by libwgm.
 *
 * A kernel written against the reference's scheduler stubs
 * (kernel_merge/stubs/stubs.h:4-7 and the OpenCL work-item builtins) keeps
 * its calls; this header gives them their meaning on gfx950. Two forms:
 *
 * Standalone persistent kernel (the form kernel_merge's standalone emitter
 * produces, kernel_merge/src/ProcessStandalonePersistentKernelVisitor.cpp):
 *
 *   __global__ void k(wgm_kernel_args wa, int* data, int n) {
 *     DISCOVERY_PROTOCOL(wa);                 // non-participants return here
 *     for (int i = get_global_id(0); i < n; i += get_global_size(0)) ...
 *     global_barrier();                       // every participant, acquire/release
 *     ...
 *     WGM_STANDALONE_END(wa);
 *   }
 *
 * launched with wgm_launch_standalone (wgm.h). Ids and sizes count the
 * discovered participants only (p_get_group_id & co.,
 * src/scheduler_rt/rt_device/discovery.cl:10-24); resizing_global_barrier()
 * is global_barrier() and offer_kill()/offer_fork() do nothing, as in the
 * standalone emitter.
 *
 * Merged mega-kernel (the form of kernel_merge/src/KernelMerge.cpp:52-67 and
 * src/scheduler_rt/rt_device/main_device_body.cl:1-93): the kernel's first
 * argument is wgm_mega_args; its body is WGM_MEGA_BODY(ma, NP_CALL, P_CALL),
 * where NP_CALL runs the non-persistent task with a Kernel_ctx (k_get_*) and
 * P_CALL runs the persistent task, which calls
 * global_barrier_resize(ma, k_ctx, &to_fork) at its resizing points and
 * resumes from its restoration context (first member: int target). Launched
 * with wgm_mega_launch_custom and driven by the wgm_mega_* host protocol.
 *
 * Compile with hipcc -I include -I gpu_workgroup_managment_amd/csrc for
 * --offload-arch=gfx950; the implementation is header-only device code.
 */
#ifndef WGM_DEVICE_API_H
#define WGM_DEVICE_API_H

#include <hip/hip_runtime.h>

#include "sched_device.hpp"
#include "wgm.h"

/* First argument of a standalone persistent kernel (filled by the runtime). */
struct wgm_kernel_args {
  wgm::Ctl ctl;
};

/* First argument of a merged mega-kernel (filled by the runtime). */
struct wgm_mega_args {
  wgm::Ctl ctl;
  wgm::SchedDev sched;
};

/* Per-task virtual workgroup ids (Kernel_ctx). */
typedef wgm::KCtx Kernel_ctx;

namespace wgm_api {

/* Per-workgroup state of a standalone kernel, set by DISCOVERY_PROTOCOL. */
struct Standalone {
  wgm::Ctl ctl;
  wgm::Team team;
};
__device__ inline Standalone& standalone_state() {
  __shared__ Standalone st;
  return st;
}

__device__ inline bool begin_standalone(const wgm_kernel_args& wa) {
  wgm::Team t;  // per thread: the barrier advances every thread's own epoch copy
  const bool in = wgm::standalone_begin(wa.ctl, t);
  if (in && threadIdx.x == 0) {
    // Word by word: a struct copy into LDS goes through a private temporary,
    // which the backend then promotes to LDS once per lane of the largest
    // workgroup (32 KB per workgroup, a cap on occupancy).
    static_assert(sizeof(wgm::Ctl) % 8 == 0, "Ctl is copied in 8-byte words");
    const unsigned long long* src = (const unsigned long long*)&wa.ctl;
    unsigned long long* dst = (unsigned long long*)&standalone_state().ctl;
    for (int i = 0; i < (int)(sizeof(wgm::Ctl) / 8); i++) dst[i] = src[i];
    standalone_state().team = t;
  }
  __syncthreads();
  return in;
}

}  // namespace wgm_api

/* ------------------------------------------------------------ standalone */
#define DISCOVERY_PROTOCOL(wa)                  \
  do {                                          \
    if (!wgm_api::begin_standalone(wa)) return; \
  } while (0)

/* End of a standalone kernel: stop the device clock, retire from the
 * persistent flag the host watches (standalone.cl:316-319). */
#define WGM_STANDALONE_END(wa)                                                      \
  do {                                                                              \
    wgm::drain_vmem();                                                              \
    __syncthreads();                                                                \
    wgm::clock_end((wa).ctl, wgm_api::standalone_state().team.id);                  \
    wgm::standalone_retire((wa).ctl);                                               \
  } while (0)

__device__ inline int p_get_group_id() { return wgm_api::standalone_state().team.id; }
__device__ inline int p_get_num_groups() { return wgm_api::standalone_state().team.n; }
__device__ inline int p_get_global_id() { return p_get_group_id() * (int)blockDim.x + (int)threadIdx.x; }
__device__ inline int p_get_global_size() { return p_get_num_groups() * (int)blockDim.x; }

/* The work-item builtins of the stubs, with dimension 0 only. */
__device__ inline int get_group_id(int) { return p_get_group_id(); }
__device__ inline int get_num_groups(int) { return p_get_num_groups(); }
__device__ inline int get_global_id(int) { return p_get_global_id(); }
__device__ inline int get_global_size(int) { return p_get_global_size(); }
__device__ inline int get_local_id(int) { return (int)threadIdx.x; }
__device__ inline int get_local_size(int) { return (int)blockDim.x; }

/* Inter-workgroup barrier over the participants (global_barrier_disc,
 * src/scheduler_rt/rt_device/iw_barrier.cl:63-109): the two-level XCD barrier. */
__device__ inline void global_barrier() {
  wgm_api::Standalone& st = wgm_api::standalone_state();
  wgm::Team t = st.team;
  wgm::xcd_barrier(st.ctl, t);
  __syncthreads();
  if (threadIdx.x == 0) st.team.epoch = t.epoch;
  __syncthreads();
}
__device__ inline void resizing_global_barrier() { global_barrier(); }
__device__ inline void offer_kill() {}
__device__ inline void offer_fork() {}

/* ----------------------------------------------------------- merged form */
/* k_get_* (src/scheduler_rt/rt_device/kernel_ctx.cl:6-28): virtual ids of
 * the task the workgroup currently serves. */
__device__ inline int k_get_num_groups(const Kernel_ctx* k) { return wgm::ld_relaxed(k->num_groups); }
__device__ inline int k_get_global_size(const Kernel_ctx* k) { return k_get_num_groups(k) * (int)blockDim.x; }

namespace wgm_api {
/* The discovered id of this workgroup inside a mega-kernel. */
__device__ inline int& mega_did() {
  __shared__ int did;
  return did;
}
}  // namespace wgm_api

__device__ inline int k_get_group_id(const Kernel_ctx* k) { return k->group_ids[wgm_api::mega_did()]; }
__device__ inline int k_get_global_id(const Kernel_ctx* k) {
  return k_get_group_id(k) * (int)blockDim.x + (int)threadIdx.x;
}

/* Resizing barrier of the persistent task (global_barrier_resize,
 * src/scheduler_rt/rt_device/iw_barrier.cl:165-227,307-311). Returns true when
 * this workgroup must leave the persistent task (it was killed): the caller
 * returns. Workgroups forked in at this barrier start from `to_fork`. */
template <typename R>
__device__ inline bool global_barrier_resize(const wgm_mega_args& ma, const Kernel_ctx* k, const R& to_fork) {
  return wgm::resize_barrier<R>(ma.sched, k_get_group_id(k), to_fork);
}

/* Plain barrier over the persistent task's current workgroups, with no kill
 * and no fork (global_barrier of a merged kernel, the form kernel_merge
 * gives it in ProcessPersistentKernelVisitor.cpp:96-101; the group count is
 * re-read while waiting, so it stays correct across resizes). */
__device__ inline void wgm_global_barrier(const wgm_mega_args& ma, const Kernel_ctx* k) {
  wgm::resize_barrier<wgm::RCtx>(ma.sched, k_get_group_id(k), wgm::RCtx{0, 0}, false);
}

/* offer_kill (src/scheduler_rt/rt_device/scheduler_1.cl:59-81,138): true when
 * the scheduler wants workgroups back and this one, the task's highest id,
 * must leave the persistent task (the caller returns). */
__device__ inline bool wgm_offer_kill(const wgm_mega_args& ma, const Kernel_ctx* k) {
  return wgm::offer_kill(ma.sched, k_get_group_id(k));
}

/* offer_fork (scheduler_1.cl:10-57): free workgroups join the persistent
 * task and start from `to_fork`. Only the task's workgroup 0 forks: it also
 * leads every barrier, so a fork never runs while a barrier is counting. */
template <typename R>
__device__ inline void wgm_offer_fork(const wgm_mega_args& ma, const Kernel_ctx* k, const R& to_fork) {
  const bool lead = k_get_group_id(k) == 0;
  if (lead) {
    wgm::drain_vmem();
    __syncthreads();
    if (threadIdx.x == 0) {
      wgm::fence_release();
      int after;
      wgm::offer_fork_thread0<R>(ma.sched, to_fork, &after);
    }
    __syncthreads();
  }
}

/* The restoration context this workgroup was started or forked with. */
template <typename R>
__device__ inline R wgm_restoration_ctx(const wgm_mega_args& ma) {
  return *wgm::rslot<R>(ma.sched.r_ctx, wgm_api::mega_did());
}

/* Mega-kernel body (main_device_body.cl): discovery; workgroup 0 runs the
 * scheduler; every other participant loops on its mailbox and runs
 * NP_CALL (non-persistent task; Kernel_ctx* np_ctx in scope) or P_CALL
 * (persistent task; Kernel_ctx* p_ctx in scope) until DEVICE_TO_QUIT.
 * A kernel using it takes dynamic LDS of at least its own needs. */
#define WGM_MEGA_BODY(ma, NP_CALL, P_CALL)                                            \
  do {                                                                                \
    int wgm_n_;                                                                       \
    const int wgm_did_ = wgm::discovery_protocol((ma).ctl, &wgm_n_);                  \
    if (wgm_did_ < 0 || (ma).ctl.im_exit) return;                                     \
    if (threadIdx.x == 0) wgm_api::mega_did() = wgm_did_;                             \
    __syncthreads();                                                                  \
    if (wgm_did_ == 0) {                                                              \
      if (threadIdx.x == 0) {                                                         \
        wgm::scheduler_init((ma).sched, wgm_n_);                                      \
        wgm::scheduler_loop((ma).sched, wgm_n_ - 1);                                  \
      }                                                                               \
      __syncthreads();                                                                \
      return;                                                                         \
    }                                                                                 \
    const Kernel_ctx* np_ctx = &(ma).sched.np;                                        \
    const Kernel_ctx* p_ctx = &(ma).sched.p;                                          \
    (void)np_ctx;                                                                     \
    (void)p_ctx;                                                                      \
    while (true) {                                                                    \
      __syncthreads();                                                                \
      if (threadIdx.x == 0) {                                                         \
        wgm::st_relaxed(&(ma).sched.task_array[wgm_did_ * wgm::kLine], wgm::TASK_WAIT); \
        wgm::add_relaxed((ma).sched.available, 1);                                    \
      }                                                                               \
      const int wgm_task_ = wgm::get_task((ma).sched, wgm_did_);                      \
      if (wgm_task_ == wgm::TASK_QUIT) break;                                         \
      if (wgm_task_ == wgm::TASK_MULT) {                                              \
        NP_CALL;                                                                      \
        wgm::drain_vmem();                                                            \
        __syncthreads();                                                              \
        if (threadIdx.x == 0) {                                                       \
          wgm::fence_release();                                                       \
          wgm::add_relaxed((ma).sched.np.executing, -1);                              \
        }                                                                             \
      } else if (wgm_task_ == wgm::TASK_PERSIST) {                                    \
        P_CALL;                                                                       \
        wgm::drain_vmem();                                                            \
        __syncthreads();                                                              \
        if (threadIdx.x == 0) {                                                       \
          wgm::fence_release();                                                       \
          wgm::add_relaxed((ma).sched.p.executing, -1);                               \
          wgm::sys_add_release((ma).sched.persistent_flag, -1);                       \
        }                                                                             \
      }                                                                               \
    }                                                                                 \
  } while (0)

#endif /* WGM_DEVICE_API_H */
