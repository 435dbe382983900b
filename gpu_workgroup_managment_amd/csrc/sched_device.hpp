// sched_device.hpp -- the resizing scheduler on gfx950.
//
// Discovered workgroup 0 is the scheduler (one thread polls the host flag in
// coherent host memory); every other participant loops on its task mailbox.
// A persistent task gives workgroups up at resizing barriers (offer_kill:
// only the highest virtual id may leave) and takes free ones back there
// (offer_fork, carrying a restoration context). Non-persistent tasks take
// free workgroups from the top of the discovered ids, the persistent task
// from the bottom. State machine, id mapping and protocol follow
// src/scheduler_rt/rt_device/cl_scheduler.cl:8-171,
// src/scheduler_rt/rt_device/scheduler_1.cl:10-138,
// src/scheduler_rt/rt_device/main_device_body.cl:1-93 and
// src/scheduler_rt/rt_device/iw_barrier.cl:165-305.
#pragma once

#include "wgm_device.hpp"

namespace wgm {

// Restoration context: resume label + live variables of a persistent task.
// A frontier traversal keeps all other state in device memory, so the level
// is its only live variable; 8 bytes, copied in one store, kept in registers
// (a larger context passed by reference ends up in per-lane scratch).
struct RCtx {
  int target;  // 0: start from the top; 1: resume right after the resizing barrier
  int level;   // level / round / phase counter
};
static_assert(sizeof(RCtx) == 8, "restoration context is one 8-byte word");

// Every workgroup owns one restoration slot of kRSlotBytes in device memory.
// A persistent task stores there whatever context type it uses (its first
// member is the int resume target); the runtime's own tasks use RCtx.
constexpr int kRSlotBytes = 64;
template <typename R>
__device__ __forceinline__ R* rslot(unsigned char* slots, int did) {
  static_assert(sizeof(R) <= kRSlotBytes, "restoration context exceeds its slot");
  return (R*)(slots + (size_t)did * kRSlotBytes);
}

// Per-task virtual workgroup ids (Kernel_ctx,
// src/scheduler_rt/rt_common/kernel_ctx.h:6-10).
struct KCtx {
  int* num_groups;  // line
  int* executing;   // line
  int* group_ids;   // [max_groups]: discovered id -> virtual id
};

struct SchedDev {
  // coherent host cells (device view)
  int* scheduler_flag;
  int* task_size;
  int* participating;
  int* persistent_flag;
  int* stop;            // persistent query server: host asks it to drain and return
  // device lines
  int* task_array;      // mailbox of discovered workgroup i at task_array[i * kLine]
  int* available;       // free workgroups
  int* pool_lock;
  int* groups_to_kill;
  int* check_value;
  int* flags;           // resizable barrier flags, by persistent virtual id (lines)
  int* bar_groups;      // bar->num_groups: group count of the current phase
  int* bar_to_kill;     // query variant: kills decided by the master
  unsigned char* r_ctx;  // [max_groups] restoration slots of kRSlotBytes
  KCtx np;
  KCtx p;
  int* np_counter;      // non-persistent task: finished-workgroup counter (reset per task)
  int* np_dmin;         // non-persistent task: running minimum (reset per task)
  int* task_kind;       // host cell: kind of the next non-persistent task
  int* np_kind;         // device line: kind of the task being executed
  int* serve_on;        // device line: 1 while a query server is the persistent task
  unsigned long long* np_clock;    // device line: [0] first task begin (min), [1] last share end (max)
  int* np_epoch;        // query server: published non-persistent tasks (device line)
  int* np_ticket;       // query server: tickets taken for the current task (device line)
  int* np_k;            // query server: workgroups of the current task (device line)
  unsigned long long* h_np_times;  // host cells: exec seen, assigned, begin, end, done (device clock)
  int query_barrier;    // 1: __global_barrier_resize_query semantics
  int serve;            // 1: the persistent task is a query server (any member may leave)
  int reserve;          // query server: free workgroups kept back for non-persistent tasks
};

// ------------------------------------------------------------ pool lock
__device__ __forceinline__ bool try_lock(int* m) { return xchg_acquire(m, 1) == 0; }
__device__ __forceinline__ void sched_lock(int* m) {
  while (xchg_acquire(m, 1) != 0) {
  }
}
__device__ __forceinline__ void sched_unlock(int* m) { st_release(m, 0); }

__device__ __forceinline__ int k_num_groups(const KCtx& k) { return ld_relaxed(k.num_groups); }

// ------------------------------------------------------------- get_task
// Thread 0 spins on this workgroup's mailbox; the restoration context is
// copied through LDS to every thread (src/scheduler_rt/rt_device/cl_scheduler.cl:31-50).
// The restoration context is left in the workgroup's slot (rslot); a task
// reads it from there after the acquire above.
__device__ __forceinline__ int get_task(const SchedDev& s, int did) {
  __shared__ int s_task;
  if (threadIdx.x == 0) {
    int t;
    while ((t = ld_relaxed(&s.task_array[did * kLine])) == TASK_WAIT) {
    }
    fence_acquire();
    s_task = t;
  }
  __syncthreads();
  const int t = s_task;
  __syncthreads();
  return t;
}

// ---------------------------------------------------------- offer_kill
// Only the workgroup with the highest virtual id may leave; it takes one
// unit of groups_to_kill and releases num_groups - 1
// (src/scheduler_rt/rt_device/scheduler_1.cl:59-81). Returns true if this
// workgroup must leave the persistent task.
__device__ __forceinline__ bool offer_kill(const SchedDev& s, int vid) {
  __shared__ int s_k;
  if (threadIdx.x == 0) {
    int killed = 0;
    if (vid == k_num_groups(s.p) - 1) {
      fence_acquire();
      int to_kill = ld_relaxed(s.groups_to_kill);
      // compare-exchange: the scheduler zeroes the counter once its gather
      // is complete, and a stale decrement must not bring it back
      if (to_kill > 0 && __hip_atomic_compare_exchange_strong(s.groups_to_kill, &to_kill, to_kill - 1,
                                                              __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT)) {
        // Leave only while still the highest id: a fork can take the pool
        // lock once the scheduler's gather is complete, and it grows
        // num_groups from the value it reads. If it got in first, the kill
        // unit is dropped (the scheduler no longer waits for it).
        int expect = vid + 1;
        killed = __hip_atomic_compare_exchange_strong(s.p.num_groups, &expect, vid, __ATOMIC_RELEASE,
                                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                     ? 1
                     : 0;
      }
    }
    s_k = killed;
  }
  __syncthreads();
  const bool k = s_k != 0;
  __syncthreads();
  return k;
}

// ---------------------------------------------------------- offer_fork
// Master thread: if workgroups are free and the pool lock is free, hand the
// next discovered ids (just above the persistent range) the restoration
// context and TASK_PERSIST, then grow num_groups / executing /
// persistent_flag (src/scheduler_rt/rt_device/scheduler_1.cl:10-57).
// Returns the group count before the fork; *after = count after.
template <typename R>
__device__ __forceinline__ int offer_fork_thread0(const SchedDev& s, const R& to_fork, int* after) {
  const int pre = k_num_groups(s.p);
  *after = pre;
  if (ld_relaxed(s.available) > 0 && try_lock(s.pool_lock)) {
    const int groups = k_num_groups(s.p);
    const int snapshot = ld_relaxed(s.available);
    for (int i = groups; i < groups + snapshot; i++) {
      s.p.group_ids[i + 1] = i;
      while (ld_relaxed(&s.task_array[(i + 1) * kLine]) != TASK_WAIT) {
      }
      *rslot<R>(s.r_ctx, i + 1) = to_fork;
      st_release(&s.task_array[(i + 1) * kLine], TASK_PERSIST);
    }
    add_relaxed(s.p.num_groups, snapshot);
    add_relaxed(s.p.executing, snapshot);
    __hip_atomic_fetch_add(s.persistent_flag, snapshot, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    add_relaxed(s.available, -snapshot);
    sched_unlock(s.pool_lock);
    *after = groups + snapshot;
    return groups;
  }
  return pre;
}

// ------------------------------------------------- resizing barrier (a9/a10)
// Master (virtual id 0) waits for every peer below the live group count
// (re-read on every spin, so peers that left drop out), lets free
// workgroups fork in, publishes the next phase's group count in bar_groups
// and releases the peers that were there before the fork. Slaves first offer
// to leave. Returns true if the caller must leave the persistent task.
// Query variant: the master reads groups_to_kill and shrinks the team
// itself; condemned slaves (highest ids) leave after the release.
// resizable = false: the plain barrier over the persistent task's current
// workgroups (global_barrier in a merged kernel): no kill, no fork.
template <typename R>
__device__ __forceinline__ bool resize_barrier(const SchedDev& s, int vid, const R& to_fork, bool resizable = true) {
  __shared__ int s_b[2];
  drain_vmem();
  __syncthreads();
  if (vid == 0) {
    for (int p = 1 + (int)threadIdx.x; p < k_num_groups(s.p); p += blockDim.x) {
      while (ld_relaxed(&s.flags[p * kLine]) == 0 && p < k_num_groups(s.p)) {
      }
    }
    fence_acquire();
    __syncthreads();
    if (threadIdx.x == 0) {
      const int to_kill = s.query_barrier && resizable ? ld_relaxed(s.groups_to_kill) : 0;
      int former, after, kill = 0;
      if (!resizable) {
        former = after = k_num_groups(s.p);
      } else if (to_kill <= 0) {
        former = offer_fork_thread0(s, to_fork, &after);
      } else {
        former = k_num_groups(s.p);
        kill = to_kill < former - 1 ? to_kill : former - 1;  // the master stays
        add_relaxed(s.groups_to_kill, -kill);  // may go below zero after a reset: read as no kill
        __hip_atomic_fetch_sub(s.p.num_groups, kill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        after = former - kill;
      }
      st_relaxed(s.bar_to_kill, kill);
      st_relaxed(s.bar_groups, after);
      s_b[0] = former;
    }
    __syncthreads();
    const int former = s_b[0];
    fence_release();
    for (int p = 1 + (int)threadIdx.x; p < former; p += blockDim.x) st_relaxed(&s.flags[p * kLine], 0);
    __syncthreads();
    return false;
  }
  if (resizable && !s.query_barrier && offer_kill(s, vid)) return true;
  if (threadIdx.x == 0) {
    int* f = &s.flags[vid * kLine];
    fence_release();
    st_relaxed(f, 1);
    while (ld_relaxed(f) != 0) {
    }
    fence_acquire();
    s_b[1] = s.query_barrier ? (vid >= ld_relaxed(s.bar_groups) ? 1 : 0) : 0;
  }
  __syncthreads();
  const bool leave = s_b[1] != 0;
  __syncthreads();
  return leave;
}

// ------------------------------------------------- scheduler (WG 0, thread 0)
// src/scheduler_rt/rt_device/cl_scheduler.cl:54-64
__device__ __forceinline__ void scheduler_init(const SchedDev& s, int count) {
  sys_st_release(s.participating, count - 1);
  st_release(s.pool_lock, 0);
  st_release(s.np.executing, 0);
  st_release(s.p.executing, 0);
  sys_st_release(s.persistent_flag, 0);
  sys_st_release(s.scheduler_flag, DEVICE_WAITING);
}

// Non-persistent tasks take discovered ids lpg, lpg-1, ... as virtual ids
// 0, 1, ... (src/scheduler_rt/rt_device/scheduler_1.cl:83-105).
__device__ __forceinline__ void assign_np(const SchedDev& s, int k, int lpg) {
  if (s.serve) {
    // Query server: any member may have left, so the free workgroups are not
    // necessarily the top ids. Take the first k free mailboxes from the top;
    // the gather guaranteed k of them (a mailbox stays TASK_WAIT until this
    // thread assigns it).
    int found = 0;
    for (int i = lpg; found < k; i = (i == 1) ? lpg : i - 1) {
      if (ld_relaxed(&s.task_array[i * kLine]) != TASK_WAIT) continue;
      s.np.group_ids[i] = found++;
      fence_release();  // group id (and the task's counters) before the mailbox
      st_relaxed(&s.task_array[i * kLine], TASK_MULT);
      add_relaxed(s.available, -1);
    }
    return;
  }
  for (int i = 0; i < k; i++) {
    s.np.group_ids[lpg - i] = i;
    while (ld_relaxed(&s.task_array[(lpg - i) * kLine]) != TASK_WAIT) {
    }
    st_release(&s.task_array[(lpg - i) * kLine], TASK_MULT);
    add_relaxed(s.available, -1);
  }
}

// The persistent task takes discovered ids 1..k as virtual ids 0..k-1
// (src/scheduler_rt/rt_device/scheduler_1.cl:107-135).
__device__ __forceinline__ void assign_p(const SchedDev& s, int k) {
  for (int i = 0; i < k; i++) {
    s.p.group_ids[i + 1] = i;
    while (ld_relaxed(&s.task_array[(i + 1) * kLine]) != TASK_WAIT) {
    }
    *(int*)rslot<RCtx>(s.r_ctx, i + 1) = 0;  // target 0: start from the top
    st_release(&s.task_array[(i + 1) * kLine], TASK_PERSIST);
    add_relaxed(s.available, -1);
  }
}

__device__ __forceinline__ int host_flag(const SchedDev& s) { return sys_ld_relaxed(s.scheduler_flag); }

// Wait until `available` reaches k. Returns false if the host asked to quit
// meanwhile (the gather is then abandoned: nothing has been assigned yet).
__device__ __forceinline__ bool gather(const SchedDev& s, int k) {
  unsigned spins = 0;
  while (ld_relaxed(s.available) < k) {
    if ((++spins & 1023u) == 0 && host_flag(s) == DEVICE_TO_QUIT) return false;
  }
  fence_acquire();
  return true;
}

// Wait for DEVICE_TO_EXECUTE; false if the host sent DEVICE_TO_QUIT instead.
__device__ __forceinline__ bool await_execute(const SchedDev& s) {
  int f;
  while ((f = host_flag(s)) != DEVICE_TO_EXECUTE) {
    if (f == DEVICE_TO_QUIT) return false;
  }
  fence_acquire();
  return true;
}

// Query server only (no counterpart in the reference, whose persistent task
// regains workgroups only at its own barriers): while the host is idle, free
// workgroups beyond `reserve` rejoin the server. The server has no global
// barrier, so a member needs no restoration context or dense id.
// One mailbox per call (the position is kept by the caller), so the host
// flag is polled between any two mailbox loads.
__device__ __forceinline__ void serve_fork_back(const SchedDev& s, int lpg, int* pos) {
  if (ld_relaxed(s.available) <= s.reserve || ld_relaxed(s.serve_on) == 0) return;
  const int i = *pos;
  *pos = i >= lpg ? 1 : i + 1;
  if (ld_relaxed(&s.task_array[i * kLine]) != TASK_WAIT) return;
  add_relaxed(s.p.executing, 1);
  sys_add_release(s.persistent_flag, 1);
  add_relaxed(s.available, -1);
  st_release(&s.task_array[i * kLine], TASK_PERSIST);
}

// src/scheduler_rt/rt_device/cl_scheduler.cl:66-171
__device__ __forceinline__ void scheduler_loop(const SchedDev& s, int lpg) {
  int fork_pos = 1;
  while (true) {
    const int flag = host_flag(s);
    if (s.serve && flag == DEVICE_WAITING) serve_fork_back(s, lpg, &fork_pos);
    if (flag == DEVICE_TO_QUIT) {
      fence_acquire();
      if (s.serve) st_relaxed(s.serve_on, 0);
      while (ld_relaxed(s.np.executing) != 0) {
      }
      while (ld_relaxed(s.p.executing) != 0) {
      }
      fence_acquire();
      sched_lock(s.pool_lock);
      for (int i = 1; i < lpg + 1; i++) {
        while (ld_acquire(&s.task_array[i * kLine]) != TASK_WAIT) {
        }
        st_release(&s.task_array[i * kLine], TASK_QUIT);
        add_relaxed(s.available, -1);
      }
      sched_unlock(s.pool_lock);
      break;
    }
    if (flag == DEVICE_TO_TASK) {
      fence_acquire();
      const int k = sys_ld_relaxed(s.task_size);
      st_relaxed(s.np.num_groups, k);
      st_relaxed(s.np.executing, k);
      sched_lock(s.pool_lock);
      // Clamped at zero: with enough free workgroups, k - available is
      // negative, and a negative kill count must not reach the barriers
      // (the reference stores it unclamped, cl_scheduler.cl:112-113).
      const int deficit = k - ld_relaxed(s.available);
      st_release(s.groups_to_kill, deficit > 0 ? deficit : 0);
      const bool got = gather(s, k);
      xchg_acquire(s.groups_to_kill, 0);  // a persistent group that ended may have freed a slot meanwhile
      if (!got || (sys_st_release(s.scheduler_flag, DEVICE_GOT_GROUPS), !await_execute(s))) {
        st_relaxed(s.np.num_groups, 0);
        st_release(s.np.executing, 0);
        sched_unlock(s.pool_lock);
        continue;  // the quit branch runs on the next pass
      }
      // reset_non_persistent, done on the device (nothing else can run on
      // the GPU while the mega-kernel holds it)
      const unsigned long long t_seen = wall_clock64();
      st_relaxed(s.np_kind, sys_ld_relaxed(s.task_kind));
      if (s.np_counter) st_relaxed(s.np_counter, 0);
      if (s.np_dmin) st_relaxed(s.np_dmin, 0x7fffffff);
      if (s.np_clock) {
        __hip_atomic_store(&s.np_clock[0], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&s.np_clock[1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (s.serve) {
        // publish: k tickets, then the epoch. A free workgroup whose epoch is
        // stale may take a ticket as soon as the ticket counter is reset, before
        // the epoch moves, so the task's kind, counters and k are released
        // before that reset (and the taker acquires after its ticket RMW).
        st_relaxed(s.np_k, k);
        fence_release();
        st_relaxed(s.np_ticket, 0);
        fence_release();
        add_relaxed(s.np_epoch, 1);
      } else {
        assign_np(s, k, lpg);
      }
      const unsigned long long t_assigned = wall_clock64();
      sched_unlock(s.pool_lock);
      while (ld_relaxed(s.np.executing) != 0) {
      }
      fence_acquire();
      if (s.h_np_times) {
        const unsigned long long t_done = wall_clock64();
        const unsigned long long tv[5] = {
            t_seen, t_assigned, __hip_atomic_load(&s.np_clock[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
            __hip_atomic_load(&s.np_clock[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), t_done};
        for (int i = 0; i < 5; i++) __hip_atomic_store(&s.h_np_times[i], tv[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
      sys_st_release(s.scheduler_flag, DEVICE_WAITING);
    }
    if (flag == DEVICE_TO_PERSISTENT_TASK) {
      fence_acquire();
      const int k = sys_ld_relaxed(s.task_size);
      st_relaxed(s.p.num_groups, k);
      st_relaxed(s.bar_groups, k);
      st_relaxed(s.p.executing, k);
      sys_st_release(s.persistent_flag, k);
      sched_lock(s.pool_lock);
      if (!gather(s, k) || (sys_st_release(s.scheduler_flag, DEVICE_GOT_GROUPS), !await_execute(s))) {
        st_relaxed(s.p.executing, 0);
        sys_st_release(s.persistent_flag, 0);
        sched_unlock(s.pool_lock);
        continue;
      }
      assign_p(s, k);
      if (s.serve) st_relaxed(s.serve_on, 1);
      sched_unlock(s.pool_lock);
      sys_st_release(s.scheduler_flag, DEVICE_WAITING);
    }
  }
}

}  // namespace wgm
