// wgm_device.hpp -- device-side cooperative workgroup runtime for gfx950.
//
// Occupancy discovery, inter-workgroup barriers, per-task virtual workgroup
// ids, and the resizing scheduler's offer_kill / offer_fork, written for a
// wave64 CDNA4 part with eight XCDs. Device-internal flags are agent-scope
// atomics in HBM, one 128-byte line per workgroup; cells the host watches are
// system-scope atomics in coherent host memory.
//
// Reference counterparts (behaviour kept, mechanism redesigned):
//   discovery      src/scheduler_rt/rt_device/discovery.cl:26-72
//   flag barrier   src/scheduler_rt/rt_device/iw_barrier.cl:63-109
//   resize barrier src/scheduler_rt/rt_device/iw_barrier.cl:165-227
//   sense barrier  src/scheduler_rt/rt_device/iw_barrier.cl:315-351
//   kernel ctx     src/scheduler_rt/rt_device/kernel_ctx.cl:6-28
//   scheduler      src/scheduler_rt/rt_device/cl_scheduler.cl:8-171,
//                  src/scheduler_rt/rt_device/scheduler_1.cl:10-138
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wgm {

constexpr int kLine = 32;  // ints per 128-byte line

// ------------------------------------------------------------------ atomics
__device__ __forceinline__ int ld_relaxed(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_acquire(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int add_relaxed(int* p, int v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int add_acq_rel(int* p, int v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int or_acq_rel(int* p, int v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int xchg_acquire(int* p, int v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }

// System scope: cells shared with the host through coherent host memory.
__device__ __forceinline__ int sys_ld_acquire(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ int sys_ld_relaxed(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void sys_st_release(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ int sys_add_release(int* p, int v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void sys_st_u64(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ void spin_pause() { __builtin_amdgcn_s_sleep(1); }

// ---------------------------------------------------- shared control layout
// Host-side set-up (runtime.hip) allocates and resets these; kernels receive
// the struct by value. Every int that is spun on lives in its own line.
struct Ctl {
  // discovery (a1-a5)
  int* poll;         // bit 31: poll closed; low bits: joins
  int* count;        // participant count, -1 until the first closer publishes it
  int* p_group_ids;  // hardware workgroup id -> discovered id, -1 if not participating
  // barrier (a6-a11)
  int* flags;        // flag of discovered workgroup i at flags[i * kLine]
  int* bar_groups;   // resizable barrier group count (bar->num_groups)
  int* bar_to_kill;  // query-barrier kill count (bar->to_kill)
  int* sense_counter;
  int* sense_flag;
  // XCD-aware barrier (two levels: members -> XCD leader, leaders <-> leaders)
  int* xcd_count;    // kMaxXcd lines: participants per XCD
  int* registered;   // participants that have recorded their XCD
  int* xcd_members;  // [kMaxXcd * max_groups]: discovered ids by (XCD, rank in XCD)
  int* xflags;       // kMaxXcd lines: leader of XCD x arrived at epoch
  // per-launch scratch lines (level counters, stop flags, ...)
  int* lines;        // kNumLines lines
  unsigned long long* clock;  // [0] start tick, [1] end tick
  // host-watched cells (coherent host memory)
  int* h_scheduler_flag;
  int* h_persistent_flag;
  int* h_participating;
  int max_groups;
  int im_exit;
  unsigned int window_ticks;  // discovery quiescence window (device clock)
  int handshake;              // standalone host handshake on/off
};
constexpr int kNumLines = 16;
constexpr int kMaxXcd = 16;

// scheduler flag values (src/scheduler_rt/rt_common/cl_scheduler.h:15-21)
constexpr int DEVICE_SCHEDULER_INIT = 0;
constexpr int DEVICE_WAITING = 1;
constexpr int DEVICE_TO_QUIT = 2;
constexpr int DEVICE_TO_TASK = 3;
constexpr int DEVICE_GOT_GROUPS = 4;
constexpr int DEVICE_TO_EXECUTE = 5;
constexpr int DEVICE_TO_PERSISTENT_TASK = 6;
// task mailbox values (src/scheduler_rt/rt_common/constants.h:4-8)
constexpr int TASK_UNINIT = -2;
constexpr int TASK_WAIT = -1;
constexpr int TASK_QUIT = 0;
constexpr int TASK_MULT = 1;
constexpr int TASK_PERSIST = 2;

constexpr int kPollClosed = 0x40000000;
constexpr int kPollMask = 0x3fffffff;

// Hardware XCC (XCD) id of the executing workgroup: HW_REG_XCC_ID (id 20), bits [3:0].
__device__ __forceinline__ int hw_xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (kMaxXcd - 1); }

// Wait until this wave's vector memory operations have completed
// (s_waitcnt vmcnt(0); gfx9 encoding: vmcnt 0, expcnt 7, lgkmcnt 15). A
// non-tgsplit __syncthreads() does not drain stores, so every wave does this
// before a barrier whose release is performed by one thread.
__device__ __forceinline__ void drain_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// The participating workgroups as seen by one workgroup.
struct Team {
  int id;     // discovered id, 0..n-1
  int n;      // participants
  int xcc;    // XCD this workgroup runs on
  int xrank;  // rank among participants on that XCD (0 = XCD leader)
  int xm;     // participants on that XCD
  int epoch;  // barriers passed (same on every participant)
};

// --------------------------------------------------------------- discovery
// One fetch_add joins the poll; the join order gives dense ids 0..count-1.
// The poll closes with one fetch_or once every launched workgroup has joined
// or no workgroup has joined for the quiescence window; the first closer
// publishes the count. A workgroup whose join sees the poll closed was
// not co-resident with the participants and leaves. Replaces the ticket-lock
// poll with its 12 serialised lock rounds per workgroup
// (src/scheduler_rt/rt_device/discovery.cl:26-57).
// Returns the discovered id, or -1 for a non-participant. *count_out gets the
// participant count (valid for participants).
__device__ __forceinline__ int discovery_protocol(const Ctl& c, int* count_out) {
  __shared__ int s_disc[2];
  if (threadIdx.x == 0) {
    int id = -1, count = 0;
    int old = add_relaxed(c.poll, 1);
    // Joins beyond the flag table are turned away like late arrivals.
    if (!(old & kPollClosed) && (old & kPollMask) < c.max_groups) {
      id = old & kPollMask;
      c.p_group_ids[blockIdx.x] = id;
      unsigned long long t0 = wall_clock64();
      const int launched = (int)gridDim.x;
      int seen = (old & kPollMask) + 1;
      while (true) {
        int v = ld_relaxed(c.poll);
        if ((v & kPollClosed) || (v & kPollMask) >= launched) break;
        // quiescence: the window restarts whenever another workgroup joins,
        // so a slow dispatch ramp is not mistaken for the end of residency
        if ((v & kPollMask) != seen) {
          seen = v & kPollMask;
          t0 = wall_clock64();
        } else if (wall_clock64() - t0 > c.window_ticks) {
          break;
        }
        spin_pause();
      }
      int before = or_acq_rel(c.poll, kPollClosed);
      if (!(before & kPollClosed)) {
        count = before & kPollMask;
        if (count > c.max_groups) count = c.max_groups;
        st_release(c.count, count);
      } else {
        while ((count = ld_acquire(c.count)) < 0) spin_pause();
      }
    }
    s_disc[0] = id;
    s_disc[1] = count;
  }
  __syncthreads();
  *count_out = s_disc[1];
  return s_disc[0];
}

// Discovery plus XCD registration for the two-level barrier: each
// participant records its XCD and its rank there, then waits until every
// participant has registered so the per-XCD member counts are final.
// Returns false for a non-participant.
__device__ __forceinline__ bool discover_team(const Ctl& c, Team* t) {
  __shared__ int s_team[3];
  int n;
  const int id = discovery_protocol(c, &n);
  if (id < 0) return false;
  if (threadIdx.x == 0) {
    const int x = hw_xcc_id();
    const int r = add_relaxed(&c.xcd_count[x * kLine], 1);
    c.xcd_members[x * c.max_groups + r] = id;
    fence_release();
    add_relaxed(c.registered, 1);
    while (ld_relaxed(c.registered) < n) spin_pause();
    fence_acquire();
    s_team[0] = x;
    s_team[1] = r;
    s_team[2] = ld_relaxed(&c.xcd_count[x * kLine]);
  }
  __syncthreads();
  t->id = id;
  t->n = n;
  t->xcc = s_team[0];
  t->xrank = s_team[1];
  t->xm = s_team[2];
  t->epoch = 0;
  return true;
}

// ------------------------------------------------------------ spin policies
// How long a barrier waits. Production kernels wait as long as it takes
// (SpinForever, which compiles to the bare loop). The barrier microbenchmark
// gives every wait a budget of spins (SpinBounded): a workgroup that exhausts
// it raises the abort line, every other wait sees the line within 64 spins
// and gives up too, so a barrier that never completes ends the kernel with an
// error instead of hanging it.
struct SpinForever {
  __device__ __forceinline__ bool go(unsigned&) const { return true; }
};
struct SpinBounded {
  int* abort;       // device line, nonzero once any wait gave up
  unsigned budget;  // spins per wait
  __device__ __forceinline__ bool go(unsigned& k) const {
    ++k;
    if ((k & 63u) == 0u && ld_relaxed(abort)) return false;
    if (k >= budget) {
      st_relaxed(abort, 1);
      return false;
    }
    return true;
  }
};

// ------------------------------------------------------------ flag barrier
// Master/slave barrier over discovered ids 0..n-1 (global_barrier_disc).
// Slaves publish arrival in their own line and spin there; the master polls
// up to eight peers per thread with all loads in flight at once, then
// releases every slave. One agent-scope release fence before each batch of
// relaxed flag stores and one acquire fence after each spin.
template <class Spin>
__device__ __forceinline__ void master_collect(int* flags, int lo, int hi, const Spin& sp) {
  // wait for flags[p] == 1 for p in [lo, hi)
  const int tid = threadIdx.x, bd = blockDim.x;
  unsigned spins = 0;
  if (hi - lo <= 8 * bd) {
    unsigned pending = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (lo + tid + k * bd < hi) pending |= 1u << k;
    while (pending && sp.go(spins)) {
      int v[8];
#pragma unroll
      for (int k = 0; k < 8; k++)
        v[k] = ((pending >> k) & 1u) ? ld_relaxed(&flags[(lo + tid + k * bd) * kLine]) : 1;
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (v[k] != 0) pending &= ~(1u << k);
    }
  } else {
    for (int p = lo + tid; p < hi; p += bd)
      while (ld_relaxed(&flags[p * kLine]) == 0 && sp.go(spins)) {
      }
  }
}

__device__ __forceinline__ void master_release(int* flags, int lo, int hi) {
  fence_release();
  for (int p = lo + (int)threadIdx.x; p < hi; p += blockDim.x) st_relaxed(&flags[p * kLine], 0);
}

template <class Spin>
__device__ __forceinline__ void slave_arrive_wait(int* flags, int id, const Spin& sp) {
  int* f = &flags[id * kLine];
  fence_release();
  st_relaxed(f, 1);
  unsigned spins = 0;
  while (ld_relaxed(f) != 0 && sp.go(spins)) {
  }
  fence_acquire();
}

template <class Spin = SpinForever>
__device__ __forceinline__ void global_barrier_disc(const Ctl& c, int id, int n, const Spin& sp = Spin()) {
  drain_vmem();
  if (n <= 1) {
    __syncthreads();
    return;
  }
  if (id == 0) {
    master_collect(c.flags, 1, n, sp);
    fence_acquire();
    __syncthreads();
    master_release(c.flags, 1, n);
  } else {
    __syncthreads();
    if (threadIdx.x == 0) slave_arrive_wait(c.flags, id, sp);
    __syncthreads();
  }
}

// --------------------------------------------------- XCD-aware barrier
// Barrier over the participants of a fixed team, two levels deep. A member
// drains its stores into its XCD's L2, publishes arrival 2e-1 in its own line
// and spins there for 2e. The leader of each XCD (rank 0 there) collects its
// members, writes back its XCD's L2 once (agent release) -- the L2 is shared
// by the XCD, so this publishes every member's stores -- meets the other XCD
// leaders through one line each, and releases its members; every workgroup
// then performs its own agent acquire. Per barrier that is one L2 writeback
// per XCD instead of one per workgroup, and no workgroup polls more than one
// XCD's worth of lines. Same contract as global_barrier_disc
// (src/scheduler_rt/rt_device/iw_barrier.cl:63-109).
// Measured on MI355X: the XCD L2s are not coherent with each other (without
// the leaders' writeback, readers on other XCDs see stale data) and an
// L1-only invalidate (buffer_inv sc0) does not act outside threadgroup-split
// mode, so members keep the full agent-scope acquire.
// MAINT = false drops the write-back and the acquires: for callers whose
// shared state goes only through agent-scope atomics (run_team), where the
// barrier need only order arrivals after this workgroup's drained accesses.
template <bool MAINT = true, class Spin = SpinForever>
__device__ __forceinline__ void xcd_barrier(const Ctl& c, Team& t, const Spin& sp = Spin()) {
  drain_vmem();
  if (t.n <= 1) {
    __syncthreads();
    return;
  }
  t.epoch++;
  const int arrive = 2 * t.epoch - 1, release = 2 * t.epoch;
  const int* members = c.xcd_members + t.xcc * c.max_groups;
  unsigned spins = 0;
  __syncthreads();
  if (t.xrank != 0) {
    if (threadIdx.x == 0) {
      int* f = &c.flags[t.id * kLine];
      st_relaxed(f, arrive);
      while (ld_relaxed(f) != release && sp.go(spins)) {
      }
      if (MAINT) fence_acquire();
    }
  } else {
    for (int r = 1 + (int)threadIdx.x; r < t.xm; r += blockDim.x) {
      const int* f = &c.flags[members[r] * kLine];
      while (ld_relaxed(f) != arrive && sp.go(spins)) {
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      if (MAINT) fence_release();
      st_relaxed(&c.xflags[t.xcc * kLine], t.epoch);
    }
    if (threadIdx.x < kMaxXcd && ld_relaxed(&c.xcd_count[threadIdx.x * kLine]) > 0) {
      while (ld_relaxed(&c.xflags[threadIdx.x * kLine]) < t.epoch && sp.go(spins)) {
      }
    }
    __syncthreads();
    if (MAINT && threadIdx.x == 0) fence_acquire();
    __syncthreads();
    for (int r = 1 + (int)threadIdx.x; r < t.xm; r += blockDim.x) st_relaxed(&c.flags[members[r] * kLine], release);
  }
  __syncthreads();
}

// -------------------------------------------------- sense-reversal barrier
// Counter + sense barrier whose group count is re-read on every spin
// (global_barrier_robust_to_resizing); `*groups` may change between calls.
template <class Spin = SpinForever>
__device__ __forceinline__ void global_barrier_sense(int* counter, int* sense_word, const int* groups,
                                                     int* local_sense, const Spin& sp = Spin()) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const int s = !(*local_sense);
    *local_sense = s;
    unsigned spins = 0;
    if (add_acq_rel(counter, 1) == 0) {
      while (sp.go(spins)) {
        const int arrived = ld_acquire(counter);
        const int n = ld_relaxed(groups);
        if (arrived == n) {
          st_relaxed(counter, 0);
          st_release(sense_word, s);
          break;
        }
      }
    } else {
      while (ld_relaxed(sense_word) != s && sp.go(spins)) {
      }
      fence_acquire();
    }
  }
  __syncthreads();
}

// ---------------------------------------------------- standalone prologue
// Group 0 publishes persistent_flag = count and DEVICE_WAITING, then waits for
// DEVICE_TO_PERSISTENT_TASK (src/lonestar/graph_apps/bfs/device/standalone.cl:277-290).
__device__ __forceinline__ void standalone_handshake(const Ctl& c, int id, int n) {
  if (c.handshake && id == 0 && threadIdx.x == 0) {
    *c.h_participating = n;
    sys_st_release(c.h_persistent_flag, n);
    sys_st_release(c.h_scheduler_flag, DEVICE_WAITING);
    while (sys_ld_relaxed(c.h_scheduler_flag) != DEVICE_TO_PERSISTENT_TASK) spin_pause();
    fence_acquire();
  }
}

// Epilogue: each workgroup retires from the persistent task
// (src/lonestar/graph_apps/bfs/device/standalone.cl:316-319).
__device__ __forceinline__ void standalone_retire(const Ctl& c) {
  if (c.handshake && threadIdx.x == 0) sys_add_release(c.h_persistent_flag, -1);
}

__device__ __forceinline__ void clock_start(const Ctl& c, int id) {
  if (id == 0 && threadIdx.x == 0) c.clock[0] = wall_clock64();
}
__device__ __forceinline__ void clock_end(const Ctl& c, int id) {
  if (id == 0 && threadIdx.x == 0) c.clock[1] = wall_clock64();
}

// Prologue of a standalone persistent kernel: discovery (+XCD registration),
// optional host handshake, a first barrier, start of the device clock.
// Returns false for a workgroup that does not take part.
__device__ __forceinline__ bool standalone_begin(const Ctl& c, Team& t) {
  if (!discover_team(c, &t) || c.im_exit) return false;
  standalone_handshake(c, t.id, t.n);
  xcd_barrier(c, t);
  clock_start(c, t.id);
  return true;
}

// --------------------------------------------------------- workgroup scans
// Exclusive prefix sum over the workgroup (blockDim.x a multiple of 64).
// s_wave must hold 17 ints. Returns the exclusive prefix; *total = sum.
__device__ __forceinline__ int wg_exclusive_scan(int x, int* s_wave, int* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int inc = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    int w = lane < nw ? s_wave[lane] : 0;
    int winc = w;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
      int y = __shfl_up(winc, d, 64);
      if (lane >= d) winc += y;
    }
    if (lane < nw) s_wave[lane] = winc - w;
    if (lane == nw - 1) s_wave[16] = winc;
  }
  __syncthreads();
  const int excl = inc - x + s_wave[wid];
  *total = s_wave[16];
  __syncthreads();
  return excl;
}

// Wave-aggregated push: one atomic per wave, lanes write consecutive slots.
__device__ __forceinline__ void wave_push(bool want, int item, int* queue, int* count) {
  const unsigned long long mask = __ballot(want);
  if (mask == 0ull) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (lane == leader) base = add_relaxed(count, __popcll(mask));
  base = __builtin_amdgcn_readlane(base, leader);  // leader is wave-uniform: no LDS round trip
  if (want) queue[base + __popcll(mask & ((1ull << lane) - 1ull))] = item;
}

}  // namespace wgm
