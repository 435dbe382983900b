// serve.hip -- the query server: a resident kernel whose persistent task
// serves BFS / SSSP requests on resident graphs while non-persistent tasks
// borrow workgroups through the resizing scheduler.
//
// No counterpart in the reference, whose driver pairs one persistent task
// with a fixed-period stream of non-persistent tasks
// (src/host_skeleton/skel.cpp:405-463). The scheduler protocol, the
// non-persistent tasks and the frontier relaxation are the same as in the
// mega-kernel (sched.hip) and the batched kernel (batched.hip).
//
// Device side. Discovered workgroup 0 runs the scheduler (wave 0) and a
// doorbell (wave 1) that mirrors the host's ring tail and stop cells into
// device lines, so serving workgroups never poll across the host link. A
// serving workgroup takes, in order: a kill request (it leaves the server),
// a preempted query to resume, a new query from the ring. One workgroup runs
// one query, level by level, with the frontier counters in LDS; at every
// level boundary it may yield: the level and frontier size go to device
// memory and the query goes to the resume queue. Finished queries publish a
// completion record (levels, reached, edges scanned, digest, device clock
// stamps) into coherent host memory.
//
// Host side. Requests are admitted into slots while their state (distances,
// stamps, two frontier queues) fits the HBM budget next to the resident
// graphs, and are published through the ring; completions are harvested by
// polling the slots' records.

#include "sched_internal.hpp"
#include "team_frontier.hpp"

#include <atomic>
#include <cmath>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <thread>
#include <vector>

using namespace wgm;

namespace wgm {

struct ServeGraphDev {
  const int32_t* row;
  const int32_t* col;
  const int32_t* wt;
  int n;
  int m;
  uint32_t delta;  // SSSP near-far band width
  int pad;
};

// SSSP band widths of the server's queries, as tuned for the batched launch
// (weights in [1, 255]): wide on graphs of out-degree <= 4 (long, narrow
// traversals), narrow on the others (short, wide ones).
constexpr uint32_t kServeDelta = 512;
constexpr uint32_t kServeDeltaHighDegree = 8;

// Written by the host into its slot before the slot id is published.
struct ServeReq {
  int graph, query, src, seq;
  long long state_off;  // 4-byte words into the state arena
  long long pad;
};
static_assert(sizeof(ServeReq) == 32, "request descriptor is four 8-byte words");

// Written by the device when the query in the slot finishes; seq last.
struct ServeRec {
  int seq, levels;
  long long reached, scanned;
  unsigned long long digest;
  long long t0, t1;
  int preempted, pad;
  long long pad2;
};
static_assert(sizeof(ServeRec) == 64, "completion record: one 64-byte line per slot");

constexpr int kSaved = 5;  // ints saved per slot at a yield

struct ServeDev {
  const ServeGraphDev* graphs;
  const ServeReq* reqs;        // host memory, device view
  ServeRec* recs;              // host memory, device view
  const int* ring;             // host memory: published slot ids
  const int* ring_tail_host;   // host cell: entries published
  int* ring_tail;              // device line: mirror of ring_tail_host
  int* ring_head;              // device line: entries claimed
  int* stop_dev;               // device line: mirror of the host stop cell
  int* quit_dev;               // device line: the scheduler has quit
  int* resume;                 // [slots]: slot + 1, 0 when empty
  int* resume_head;            // device line
  int* resume_tail;            // device line
  int* saved;                  // [slots * kSaved]: level, frontier size, band threshold, far queue, its count
  long long* saved_t0;         // [slots]: device clock at the first level
  int* saved_pre;              // [slots]: times yielded
  uint32_t* arena;             // request state
  int ring_cap;
  int slots;
};

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Any serving workgroup may leave: the server has no global barrier.
__device__ __forceinline__ bool take_kill(const SchedDev& s) {
  int k = ld_relaxed(s.groups_to_kill);
  while (k > 0) {
    if (__hip_atomic_compare_exchange_strong(s.groups_to_kill, &k, k - 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return true;
  }
  return false;
}

__device__ __forceinline__ void push_resume(const ServeDev& v, int slot) {
  const int p = add_relaxed(v.resume_tail, 1);
  st_release(&v.resume[p % v.slots], slot + 1);
}

__device__ __forceinline__ int pop_resume(const ServeDev& v) {
  int h = ld_relaxed(v.resume_head);
  while (h < ld_relaxed(v.resume_tail)) {
    if (__hip_atomic_compare_exchange_strong(v.resume_head, &h, h + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)) {
      int* cell = &v.resume[h % v.slots];
      int x;
      while ((x = ld_relaxed(cell)) == 0) {
      }
      st_relaxed(cell, 0);
      fence_acquire();
      return x - 1;
    }
  }
  return -1;
}

// Idle workgroups poll here: relaxed loads only (an acquire in a spin would
// invalidate the XCD's L2 on every iteration); one acquire on success.
__device__ __forceinline__ int claim_ring(const ServeDev& v) {
  int h = ld_relaxed(v.ring_head);
  while (h < ld_relaxed(v.ring_tail)) {
    if (__hip_atomic_compare_exchange_strong(v.ring_head, &h, h + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)) {
      fence_acquire();
      return sys_ld_relaxed(&v.ring[h % v.ring_cap]);
    }
  }
  return -1;
}

__device__ __forceinline__ ServeReq read_req(const ServeReq* cell) {
  ServeReq r;
  const unsigned long long* src = (const unsigned long long*)cell;
  unsigned long long* dst = (unsigned long long*)&r;
  for (int i = 0; i < 4; i++) dst[i] = __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  return r;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned lo = __shfl_xor((unsigned)x, off, 64);
    const unsigned hi = __shfl_xor((unsigned)(x >> 32), off, 64);
    x += ((unsigned long long)hi << 32) | lo;
  }
  return x;
}

// One query on this workgroup, from the top or from a saved level. Returns
// true if the workgroup yielded (the query is in the resume queue).
// SSSP runs near-far bands as the batched single-workgroup step does: a
// relaxation at or past the band's threshold defers the target to the far
// queue (once, under its flag); when the near frontier empties, the band
// moves to the smallest deferred distance still at or above the old
// threshold + delta, and the far queue is split. The request state is dist,
// stamps, two frontier queues and, for SSSP, two far queues and the flags.
// A yield happens only at a level's end, with the band's threshold, far queue
// and its count saved beside the level.
template <bool SSSP>
__device__ bool serve_query(const SchedDev& s, const ServeDev& v, int slot, bool resume, const ServeReq& rq,
                            int* lds, int* s_cnt, int* s_misc) {
  const ServeGraphDev gd = v.graphs[rq.graph];
  __shared__ int s_far[3];        // far counts of far(0), far(1); band minimum
  __shared__ uint32_t s_band[2];  // threshold, current far queue
  GraphDev g{};
  g.row_ptr = gd.row;
  g.col = gd.col;
  g.wt = gd.wt;
  g.n = gd.n;
  g.m = gd.m;
  FrontierDev f{};
  f.dist = v.arena + rq.state_off;
  f.stamp = f.dist + gd.n;
  f.q0 = (int32_t*)(f.stamp + gd.n);
  f.q1 = f.q0 + gd.n + 1;
  const bool near_far = SSSP && gd.delta > 0;
  NearFar nf{};
  if (near_far) {
    nf.q[0] = f.q1 + gd.n + 1;
    nf.q[1] = nf.q[0] + gd.n + 1;
    nf.flag = (uint32_t*)(nf.q[1] + gd.n + 1);
    nf.delta = gd.delta;
  }
  if (!resume) {
    for (int x = threadIdx.x; x < gd.n; x += blockDim.x) {
      f.dist[x] = (x == rq.src) ? 0u : kInf;
      if (SSSP) f.stamp[x] = 0u;
      if (near_far) nf.flag[x] = 0u;
    }
    if (threadIdx.x == 0) {
      f.q0[0] = rq.src;
      s_cnt[0] = 1;
      s_cnt[1] = 0;
      s_cnt[2] = 0;
      s_misc[0] = 0;
      s_far[0] = 0;
      s_far[1] = 0;
      s_far[2] = INT32_MAX;
      s_band[0] = near_far ? nf.delta : 0xffffffffu;
      s_band[1] = 0u;
      v.saved_t0[slot] = (long long)wall_clock64();
      v.saved_pre[slot] = 0;
    }
  } else if (threadIdx.x == 0) {
    const int* sv = v.saved + (size_t)slot * kSaved;
    const int level = sv[0];
    s_cnt[level % 3] = sv[1];
    s_cnt[(level + 1) % 3] = 0;
    s_cnt[(level + 2) % 3] = 0;
    s_misc[0] = level;
    s_band[0] = (uint32_t)sv[2];
    s_band[1] = (uint32_t)sv[3];
    s_far[sv[3]] = sv[4];
    s_far[1 - sv[3]] = 0;
    s_far[2] = INT32_MAX;
  }
  drain_vmem();
  __syncthreads();
  int level = s_misc[0];
  uint32_t thr = s_band[0];
  int fi = (int)s_band[1];
  const int bd = blockDim.x;
  while (true) {
    while (true) {
      const int in_n = s_cnt[level % 3];
      if (in_n == 0) break;
      const int32_t* in_q = (level & 1) ? f.q1 : f.q0;
      int32_t* out_q = (level & 1) ? f.q0 : f.q1;
      for (int a = 0; a < in_n; a += bd)
        frontier_batch<SSSP, true>(g, f, in_q, a, min(a + bd, in_n), out_q, &s_cnt[(level + 1) % 3],
                                   (uint32_t)level + 1u, lds, thr, near_far ? nf.far(fi) : nullptr, &s_far[fi],
                                   nf.flag);
      if (threadIdx.x == 0) {
        s_cnt[(level + 2) % 3] = 0;
        s_misc[1] = take_kill(s) ? 1 : 0;
      }
      drain_vmem();
      __syncthreads();
      level++;
      if (s_misc[1]) {
        // yield: the level is complete, its successor's frontier is in out_q
        if (threadIdx.x == 0) {
          int* sv = v.saved + (size_t)slot * kSaved;
          sv[0] = level;
          sv[1] = s_cnt[level % 3];
          sv[2] = (int)thr;
          sv[3] = fi;
          sv[4] = s_far[fi];
          v.saved_pre[slot] += 1;
          fence_release();
          push_resume(v, slot);
        }
        __syncthreads();
        return true;
      }
    }
    if (!near_far) break;
    const int fn = s_far[fi];
    if (fn == 0) break;
    // band split: far entries already below the finished band's threshold
    // were queued near when they fell below it (relaxed in the band) and are
    // dropped; the new minimum is over the rest
    const uint32_t t_old = thr;
    int mn = INT32_MAX;
    for (int i = threadIdx.x; i < fn; i += bd) {
      const uint32_t d = f.dist[nf.far(fi)[i]];
      if (d >= t_old) mn = min(mn, (int)d);
    }
    mn = wave_min_int(mn);
    if ((threadIdx.x & 63) == 0) atomicMin(&s_far[2], mn);
    __syncthreads();
    thr = (uint32_t)s_far[2] + nf.delta;
    int32_t* lq = (level & 1) ? f.q1 : f.q0;  // this level's (empty) queue
    for (int base = 0; base < fn; base += bd) {
      const int i = base + threadIdx.x;
      bool to_near = false, to_far = false;
      int x = 0;
      if (i < fn) {
        x = nf.far(fi)[i];
        const uint32_t dx = f.dist[x];
        if (dx < thr) {
          nf.flag[x] = 0u;  // each deferred vertex is in one far queue once: it moves once
          to_near = dx >= t_old;
        } else {
          to_far = true;  // stays flagged in the other far queue
        }
      }
      push_item<false>(to_near, x, lq, &s_cnt[level % 3], __HIP_MEMORY_SCOPE_WORKGROUP);
      push_item<false>(to_far, x, nf.far(1 - fi), &s_far[1 - fi], __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      s_far[fi] = 0;
      s_far[2] = INT32_MAX;
    }
    fi = 1 - fi;
    drain_vmem();
    __syncthreads();
  }
  // completion: digest, reached vertices and scanned edges in one pass
  fence_acquire();
  unsigned long long dig = 0, reached = 0, scanned = 0;
  for (int x = threadIdx.x; x < gd.n; x += blockDim.x) {
    const uint32_t d = f.dist[x];
    dig += mix64(((unsigned long long)(unsigned)x << 32) | d);
    if (d != kInf) {
      reached++;
      scanned += (unsigned long long)(gd.row[x + 1] - gd.row[x]);
    }
  }
  dig = wave_sum_u64(dig);
  reached = wave_sum_u64(reached);
  scanned = wave_sum_u64(scanned);
  unsigned long long* red = (unsigned long long*)lds;
  const int wave = threadIdx.x / 64, nw = (bd + 63) / 64;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[wave] = dig;
    red[16 + wave] = reached;
    red[32 + wave] = scanned;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, b = 0, c = 0;
    for (int w = 0; w < nw; w++) {
      a += red[w];
      b += red[16 + w];
      c += red[32 + w];
    }
    ServeRec* r = &v.recs[slot];
    unsigned long long* rw = (unsigned long long*)r;
    // words: [seq|levels] [reached] [scanned] [digest] [t0] [t1] [preempted|pad]
    const unsigned long long t1 = wall_clock64();
    __hip_atomic_store(rw + 1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(rw + 2, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(rw + 3, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(rw + 4, (unsigned long long)v.saved_t0[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(rw + 5, t1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(rw + 6, (unsigned long long)(unsigned)v.saved_pre[slot], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
    const unsigned long long head = ((unsigned long long)(unsigned)level << 32) | (unsigned)rq.seq;
    __hip_atomic_store(rw + 0, head, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  __syncthreads();
  return false;
}

// The persistent task: serve until killed or stopped.
__device__ void serve_loop(const SchedDev& s, const ServeDev& v, int* lds) {
  __shared__ int s_cmd[2];
  __shared__ int s_cnt[3];
  __shared__ int s_misc[2];
  __shared__ ServeReq s_req;
  while (true) {
    if (threadIdx.x == 0) {
      int cmd = 0, slot = -1;  // 0 idle, 1 start, 2 resume, 3 leave
      if (take_kill(s)) {
        cmd = 3;
      } else if ((slot = pop_resume(v)) >= 0) {
        cmd = 2;
      } else if ((slot = claim_ring(v)) >= 0) {
        cmd = 1;
      } else if (ld_relaxed(v.stop_dev)) {
        cmd = 3;
      }
      if (cmd == 1 || cmd == 2) s_req = read_req(&v.reqs[slot]);
      s_cmd[0] = cmd;
      s_cmd[1] = slot;
    }
    __syncthreads();
    const int cmd = s_cmd[0], slot = s_cmd[1];
    const ServeReq rq = s_req;
    __syncthreads();
    if (cmd == 3) return;
    if (cmd == 0) {
      __builtin_amdgcn_s_sleep(2);
      continue;
    }
    const bool left = rq.query == WGM_QUERY_SSSP
                          ? serve_query<true>(s, v, slot, cmd == 2, rq, lds, s_cnt, s_misc)
                          : serve_query<false>(s, v, slot, cmd == 2, rq, lds, s_cnt, s_misc);
    if (left) return;
  }
}

// Wave 1 of the scheduler workgroup: host cells -> device lines.
__device__ void doorbell_loop(const SchedDev& s, const ServeDev& v) {
  int last_tail = 0, last_stop = 0;
  while (ld_relaxed(v.quit_dev) == 0) {
    const int t = sys_ld_relaxed(v.ring_tail_host);
    if (t != last_tail) {
      // the ring entries and descriptors are read by the claimer with
      // system-scope loads after its own acquire; order the mirror after
      // this load of the tail
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      st_release(v.ring_tail, t);
      last_tail = t;
    }
    const int st = sys_ld_relaxed(s.stop);
    if (st != last_stop) {
      st_release(v.stop_dev, st);
      last_stop = st;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

__host__ __device__ constexpr int serve_lds_bytes(int bd) {
  return mega_lds_bytes(bd) > 48 * 8 ? mega_lds_bytes(bd) : 48 * 8;
}

__global__ void __launch_bounds__(1024) k_serve(Ctl c, SchedDev s, ServeDev v, NPDev np) {
  extern __shared__ int lds[];
  int n;
  const int did = discovery_protocol(c, &n);
  if (did < 0 || c.im_exit) return;
  if (did == 0) {
    if (threadIdx.x == 0) {
      scheduler_init(s, n);
      scheduler_loop(s, n - 1);
      st_release(v.quit_dev, 1);
    } else if (threadIdx.x == 64) {
      doorbell_loop(s, v);
    }
    __syncthreads();
    return;
  }
  __shared__ int s_task[2];
  int seen_epoch = 0;  // thread 0: last non-persistent task epoch looked at
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) {
      st_relaxed(&s.task_array[did * kLine], TASK_WAIT);
      add_relaxed(s.available, 1);
      // Wait for the mailbox (server / quit) or a ticket of a published
      // non-persistent task: the scheduler publishes a task with one store
      // and the free workgroups take its k tickets in parallel.
      int t, vid = -1;
      while (true) {
        t = ld_relaxed(&s.task_array[did * kLine]);
        if (t != TASK_WAIT) break;
        const int e = ld_relaxed(s.np_epoch);
        if (e != seen_epoch) {
          seen_epoch = e;
          const int tk = add_relaxed(s.np_ticket, 1);
          fence_acquire();  // pairs with the release before the ticket reset: np_k / np_kind are the new task's
          if (tk < ld_relaxed(s.np_k)) {
            add_relaxed(s.available, -1);
            t = TASK_MULT;
            vid = tk;
            break;
          }
        }
        __builtin_amdgcn_s_sleep(1);
      }
      fence_acquire();
      s_task[0] = t;
      s_task[1] = vid;
    }
    __syncthreads();
    const int task = s_task[0], np_vid = s_task[1];
    if (task == TASK_QUIT) break;
    if (task == TASK_MULT) {
      np_task(np, ld_relaxed(s.np_kind), np_vid, ld_relaxed(s.np_k), lds, s.np_clock, s.np.executing);
    } else if (task == TASK_PERSIST) {
      serve_loop(s, v, lds);
      drain_vmem();
      __syncthreads();
      if (threadIdx.x == 0) {
        fence_release();
        add_relaxed(s.p.executing, -1);
        sys_add_release(s.persistent_flag, -1);
      }
    }
  }
}

__global__ void k_serve_reset(ServeDev v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int i = t; i < v.slots; i += gridDim.x * blockDim.x) {
    v.resume[i] = 0;
    for (int k = 0; k < kSaved; k++) v.saved[i * kSaved + k] = 0;
    v.saved_t0[i] = 0;
    v.saved_pre[i] = 0;
  }
  if (t == 0) {
    *v.ring_tail = 0;
    *v.ring_head = 0;
    *v.stop_dev = 0;
    *v.quit_dev = 0;
    *v.resume_head = 0;
    *v.resume_tail = 0;
  }
}

}  // namespace wgm

#define fail wgm_fail
#define HIPCHK WGM_HIPCHK

namespace {

uint64_t mix64_host(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct Pending {
  int64_t id;
  int graph, query, src;
  uint64_t t_submit;
};

struct Inflight {
  int64_t id = -1;
  int graph = 0, query = 0, src = 0, seq = 0;
  uint64_t t_submit = 0, t_publish = 0;
  int64_t bytes = 0;
  int64_t off = 0;  // 4-byte words into the state arena
};

int serve_lds(int wg) { return serve_lds_bytes(wg); }

wgm_serve_config serve_defaults(const wgm_serve_config* in) {
  wgm_serve_config cfg = in ? *in : wgm_serve_config{};
  if (cfg.wg_size <= 0) cfg.wg_size = 1024;
  if (cfg.max_inflight <= 0) cfg.max_inflight = 256;
  if (cfg.reserve == 0) cfg.reserve = 16;
  if (cfg.reserve < 0) cfg.reserve = 0;
  if (cfg.budget_bytes <= 0) cfg.budget_bytes = (int64_t)64 << 30;
  return cfg;
}

}  // namespace

struct wgm_serve {
  wgm_runtime* rt = nullptr;
  wgm_mega* mega = nullptr;
  wgm_serve_config cfg{};
  int count = 0;
  std::vector<int32_t> gn;
  std::vector<int64_t> gm;
  std::vector<char> has_wt;
  int64_t graph_bytes = 0;
  int nmax = 0;
  int64_t arena_words = 0;   // request-state arena: graphs + arena <= budget
  std::map<int64_t, int64_t> free_ranges;  // free arena words: offset -> length, coalesced
  void* d_graphs = nullptr;  // pooled CSR arrays
  void* d_gtab = nullptr;    // ServeGraphDev[count]
  void* d_lines = nullptr;   // lines + per-slot arrays
  void* d_arena = nullptr;   // arena_words
  char* h_mem = nullptr;     // reqs, recs, ring, tail (coherent host memory)
  char* h_mem_dev = nullptr;
  ServeReq* h_reqs = nullptr;
  ServeRec* h_recs = nullptr;
  int* h_ring = nullptr;
  int* h_ring_tail = nullptr;
  ServeDev dev{};
  int participating = 0;
  bool started = false;
  // host bookkeeping (submit / poll thread)
  int64_t next_id = 0;
  int seq = 0;
  int tail = 0;
  std::deque<Pending> pending;
  std::vector<Inflight> inflight;
  std::vector<int> free_slots;
  int64_t inflight_bytes = 0;
  int64_t max_inflight_bytes = 0;
  std::deque<wgm_serve_done> done;
  int64_t outstanding = 0;
  double clock_hz = 1e8;
};

namespace {

// Request state: dist, stamps and two frontier queues (4 n + 2 words); SSSP
// adds two far queues and the far flags (3 n + 2).
int64_t state_words(int n, bool sssp) {
  const int64_t w = (int64_t)4 * n + 2 + (sssp ? (int64_t)3 * n + 2 : 0);
  return (w + 63) / 64 * 64;
}

// Request state is carved out of the arena first-fit; a freed range merges
// with its free neighbours. With nothing in flight the arena is one free
// range, so any request that fits the arena at all is admitted eventually.
int64_t arena_take(wgm_serve* s, int64_t words) {
  for (auto it = s->free_ranges.begin(); it != s->free_ranges.end(); ++it) {
    if (it->second < words) continue;
    const int64_t off = it->first, len = it->second;
    s->free_ranges.erase(it);
    if (len > words) s->free_ranges[off + words] = len - words;
    return off;
  }
  return -1;
}

void arena_give(wgm_serve* s, int64_t off, int64_t words) {
  auto next = s->free_ranges.lower_bound(off);
  if (next != s->free_ranges.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == off) {
      off = prev->first;
      words += prev->second;
      s->free_ranges.erase(prev);
    }
  }
  if (next != s->free_ranges.end() && off + words == next->first) {
    words += next->second;
    s->free_ranges.erase(next);
  }
  s->free_ranges[off] = words;
}

// Harvest finished slots, then admit queued requests.
void serve_progress(wgm_serve* s) {
  const uint64_t now = now_ns();
  for (int slot = 0; slot < (int)s->inflight.size(); slot++) {
    Inflight& f = s->inflight[slot];
    if (f.id < 0) continue;
    const unsigned long long head =
        __hip_atomic_load((unsigned long long*)&s->h_recs[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    if ((int)(head & 0xffffffffu) != f.seq) continue;
    const ServeRec& r = s->h_recs[slot];
    wgm_serve_done d{};
    d.id = f.id;
    d.graph = f.graph;
    d.query = f.query;
    d.src = f.src;
    d.levels = (int32_t)(head >> 32);
    d.reached = r.reached;
    d.edges_scanned = r.scanned;
    d.digest = r.digest;
    d.queue_us = (double)(f.t_publish - f.t_submit) * 1e-3;
    d.service_us = (double)(now - f.t_publish) * 1e-3;
    d.device_us = (double)(r.t1 - r.t0) / s->clock_hz * 1e6;
    d.preempted = r.preempted;
    s->done.push_back(d);
    s->inflight_bytes -= f.bytes;
    arena_give(s, f.off, f.bytes / 4);
    f.id = -1;
    s->free_slots.push_back(slot);
  }
  while (!s->pending.empty() && !s->free_slots.empty()) {
    const Pending& p = s->pending.front();
    const int64_t words = state_words(s->gn[p.graph], p.query == WGM_QUERY_SSSP);
    const int64_t off = arena_take(s, words);
    if (off < 0) break;  // admission: wait for memory to free up
    const int64_t bytes = words * 4;
    const int slot = s->free_slots.back();
    s->free_slots.pop_back();
    Inflight& f = s->inflight[slot];
    f.id = p.id;
    f.graph = p.graph;
    f.query = p.query;
    f.src = p.src;
    f.seq = ++s->seq;
    f.t_submit = p.t_submit;
    f.bytes = bytes;
    f.off = off;
    s->inflight_bytes += bytes;
    if (s->inflight_bytes > s->max_inflight_bytes) s->max_inflight_bytes = s->inflight_bytes;
    ServeReq rq{};
    rq.graph = p.graph;
    rq.query = p.query;
    rq.src = p.src;
    rq.seq = f.seq;
    rq.state_off = (long long)off;
    std::memcpy((void*)&s->h_reqs[slot], &rq, sizeof(rq));
    s->h_ring[s->tail % s->dev.ring_cap] = slot;
    s->tail++;
    f.t_publish = now_ns();
    hstore(s->h_ring_tail, s->tail);
    s->pending.pop_front();
  }
}

struct ServePlan {
  int64_t graph_words = 0;  // pooled CSR arrays
  int64_t arena_words = 0;  // request-state arena
  int nmax = 0;
};

// Memory plan: the graph pool (per graph row_ptr, col and wt if present, each
// 256-byte aligned) and the request-state arena, as large as max_inflight of
// the largest request but never beyond the budget left by the graphs.
// Returns an error message, or nullptr.
const char* serve_plan(int count, const wgm_csr* graphs, const wgm_serve_config& cfg, ServePlan* out,
                       int64_t* off_row, int64_t* off_col, int64_t* off_wt) {
  auto align = [](int64_t w) { return (w + 63) / 64 * 64; };
  ServePlan p;
  bool any_wt = false;
  for (int i = 0; i < count; i++) {
    const wgm_csr& g = graphs[i];
    if (g.n < 1 || g.m < 0 || g.m > INT32_MAX || !g.row_ptr || (g.m > 0 && !g.col)) return "a graph is malformed";
    p.nmax = std::max(p.nmax, g.n);
    any_wt = any_wt || g.wt != nullptr;
    if (off_row) off_row[i] = p.graph_words;
    p.graph_words += align((int64_t)g.n + 1);
    if (off_col) off_col[i] = p.graph_words;
    p.graph_words += align(std::max<int64_t>(g.m, 1));
    if (off_wt) off_wt[i] = g.wt ? p.graph_words : -1;
    if (g.wt) p.graph_words += align(std::max<int64_t>(g.m, 1));
  }
  *out = p;
  if (p.graph_words * 4 > cfg.budget_bytes) return "the resident graphs exceed the budget";
  // SSSP (the larger state) needs a weighted graph
  const int64_t largest = state_words(p.nmax, any_wt);
  const int64_t room = (cfg.budget_bytes / 4 - p.graph_words) / 64 * 64;
  out->arena_words = std::min<int64_t>((int64_t)cfg.max_inflight * largest, room);
  return nullptr;
}

}  // namespace

bool wgm_serve_kernel_info(int kernel, KernelInfo* out) {
  if (kernel == WGM_K_SERVE) {
    *out = KernelInfo{(const void*)&k_serve, serve_lds};
    return true;
  }
  return false;
}

int wgm_serve_probe(wgm_runtime* rt, int kernel, int grid, int wg_size) {
  if (kernel != WGM_K_SERVE) return wgm_fail(rt, WGM_ERR_INVALID, "no discovery probe for kernel %d", kernel);
  hipLaunchKernelGGL(k_serve, dim3(grid), dim3(wg_size), serve_lds_bytes(wg_size), rt->stream, rt->ctl, SchedDev{},
                     ServeDev{}, NPDev{});
  return WGM_OK;
}

extern "C" {

uint64_t wgm_dist_digest(const uint32_t* dist, int32_t n) {
  uint64_t d = 0;
  for (int32_t v = 0; v < n; v++) d += mix64_host(((uint64_t)(uint32_t)v << 32) | dist[v]);
  return d;
}

int wgm_serve_create(wgm_runtime* rt, int32_t count, const wgm_csr* graphs, const wgm_serve_config* cfg_in,
                     wgm_serve** out) {
  if (!rt || !graphs || !out || count < 1) return fail(rt, WGM_ERR_INVALID, "bad arguments");
  const wgm_serve_config cfg = serve_defaults(cfg_in);
  int rc = wgm_check_wg(rt, cfg.wg_size);
  if (rc) return rc;
  if (cfg.wg_size < 128) return fail(rt, WGM_ERR_INVALID, "the server needs wg_size >= 128");
  wgm_serve* s = new wgm_serve();
  s->rt = rt;
  s->cfg = cfg;
  s->count = count;
  auto bail = [&](int code) { wgm_serve_destroy(s); return code; };
  std::vector<int64_t> off_row(count), off_col(count), off_wt(count);
  ServePlan plan;
  const char* why = serve_plan(count, graphs, cfg, &plan, off_row.data(), off_col.data(), off_wt.data());
  if (why) return bail(fail(rt, plan.graph_words * 4 > cfg.budget_bytes ? WGM_ERR_BUDGET : WGM_ERR_INVALID, "%s", why));
  for (int i = 0; i < count; i++) {
    s->gn.push_back(graphs[i].n);
    s->gm.push_back(graphs[i].m);
    s->has_wt.push_back(graphs[i].wt != nullptr);
  }
  s->nmax = plan.nmax;
  const int64_t words = plan.graph_words;
  s->graph_bytes = words * 4;
  hipError_t e = hipMalloc(&s->d_graphs, (size_t)words * 4);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipMalloc(graph pool)"));
  std::vector<ServeGraphDev> tab(count);
  for (int i = 0; i < count; i++) {
    const wgm_csr& g = graphs[i];
    int32_t* base = (int32_t*)s->d_graphs;
    e = hipMemcpy(base + off_row[i], g.row_ptr, sizeof(int32_t) * (g.n + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && g.m > 0) e = hipMemcpy(base + off_col[i], g.col, sizeof(int32_t) * g.m, hipMemcpyHostToDevice);
    if (e == hipSuccess && g.wt && g.m > 0)
      e = hipMemcpy(base + off_wt[i], g.wt, sizeof(int32_t) * g.m, hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipMemcpy(graph pool)"));
    int32_t maxdeg = 0;
    for (int32_t x = 0; x < g.n; x++) maxdeg = std::max(maxdeg, g.row_ptr[x + 1] - g.row_ptr[x]);
    tab[i] = ServeGraphDev{base + off_row[i], base + off_col[i], g.wt ? base + off_wt[i] : nullptr, g.n, (int)g.m,
                           maxdeg > 4 ? kServeDeltaHighDegree : kServeDelta, 0};
  }
  e = hipMalloc(&s->d_gtab, sizeof(ServeGraphDev) * count);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipMalloc(graph table)"));
  e = hipMemcpy(s->d_gtab, tab.data(), sizeof(ServeGraphDev) * count, hipMemcpyHostToDevice);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipMemcpy(graph table)"));
  // the scheduler cells and the non-persistent task buffers
  rc = wgm_mega_create(rt, &cfg.np, &s->mega);
  if (rc) return bail(rc);
  // device lines + per-slot arrays + state arena
  const int slots = cfg.max_inflight;
  const size_t line_b = kLine * sizeof(int);
  const size_t lines_bytes = 8 * line_b + (size_t)slots * (sizeof(int) * (2 + kSaved) + sizeof(long long));
  e = hipMalloc(&s->d_lines, lines_bytes);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipMalloc(server lines)"));
  // request state: the planned arena, carved per admitted request
  s->arena_words = plan.arena_words;
  if (s->arena_words > 0) {
    e = hipMalloc(&s->d_arena, (size_t)s->arena_words * 4);
    if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipMalloc(request state)"));
    s->free_ranges[0] = s->arena_words;
  }
  // host memory: reqs, recs (64-byte stride), ring (2 * slots), tail
  const int ring_cap = 2 * slots;
  const size_t h_bytes = (size_t)slots * 32 + (size_t)slots * 64 + (size_t)ring_cap * 4 + line_b;
  e = hipHostMalloc((void**)&s->h_mem, h_bytes, hipHostMallocCoherent | hipHostMallocMapped);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipHostMalloc(server cells)"));
  std::memset(s->h_mem, 0, h_bytes);
  e = hipHostGetDevicePointer((void**)&s->h_mem_dev, s->h_mem, 0);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipHostGetDevicePointer"));
  s->h_reqs = (ServeReq*)s->h_mem;
  ServeDev& v = s->dev;
  v.graphs = (const ServeGraphDev*)s->d_gtab;
  v.reqs = (const ServeReq*)s->h_mem_dev;
  char* recs_h = s->h_mem + (size_t)slots * 32;
  char* recs_d = s->h_mem_dev + (size_t)slots * 32;
  s->h_recs = (ServeRec*)recs_h;
  v.recs = (ServeRec*)recs_d;
  s->h_ring = (int*)(recs_h + (size_t)slots * 64);
  v.ring = (const int*)(recs_d + (size_t)slots * 64);
  s->h_ring_tail = s->h_ring + ring_cap;
  v.ring_tail_host = (const int*)(v.ring + ring_cap);
  v.ring_cap = ring_cap;
  v.slots = slots;
  char* p = (char*)s->d_lines;
  auto line = [&]() { int* r = (int*)p; p += line_b; return r; };
  v.ring_tail = line();
  v.ring_head = line();
  v.stop_dev = line();
  v.quit_dev = line();
  v.resume_head = line();
  v.resume_tail = line();
  p += 2 * line_b;
  v.saved_t0 = (long long*)p;
  p += (size_t)slots * sizeof(long long);
  v.resume = (int*)p;
  p += (size_t)slots * sizeof(int);
  v.saved = (int*)p;
  p += (size_t)slots * kSaved * sizeof(int);
  v.saved_pre = (int*)p;
  v.arena = (uint32_t*)s->d_arena;
  s->inflight.assign(slots, Inflight{});
  for (int i = slots - 1; i >= 0; i--) s->free_slots.push_back(i);
  s->clock_hz = (double)rt->info.wall_clock_hz;
  *out = s;
  return WGM_OK;
}

int wgm_serve_destroy(wgm_serve* s) {
  if (!s) return WGM_OK;
  if (s->started) (void)wgm_serve_stop(s);
  if (s->mega) (void)wgm_mega_destroy(s->mega);
  void* ptrs[] = {s->d_graphs, s->d_gtab, s->d_lines, s->d_arena};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (s->h_mem) (void)hipHostFree(s->h_mem);
  delete s;
  return WGM_OK;
}

int64_t wgm_serve_graph_bytes(const wgm_serve* s) { return s ? s->graph_bytes : 0; }
int64_t wgm_serve_arena_bytes(const wgm_serve* s) { return s ? s->arena_words * 4 : 0; }

int wgm_serve_plan(int32_t count, const wgm_csr* graphs, const wgm_serve_config* cfg_in, int64_t* graph_bytes,
                   int64_t* arena_bytes) {
  if (!graphs || count < 1) return fail(nullptr, WGM_ERR_INVALID, "bad arguments");
  wgm_serve_config cfg = serve_defaults(cfg_in);
  ServePlan plan;
  const char* why = serve_plan(count, graphs, cfg, &plan, nullptr, nullptr, nullptr);
  if (graph_bytes) *graph_bytes = plan.graph_words * 4;
  if (arena_bytes) *arena_bytes = plan.arena_words * 4;
  if (why) return fail(nullptr, plan.graph_words * 4 > cfg.budget_bytes ? WGM_ERR_BUDGET : WGM_ERR_INVALID, "%s", why);
  return WGM_OK;
}
int32_t wgm_serve_outstanding(const wgm_serve* s) { return s ? (int32_t)s->outstanding : 0; }
int64_t wgm_serve_inflight_bytes(const wgm_serve* s) { return s ? s->inflight_bytes : 0; }

int wgm_serve_start(wgm_serve* s, int32_t* participating) {
  if (!s) return fail(nullptr, WGM_ERR_INVALID, "NULL server");
  wgm_runtime* rt = s->rt;
  wgm_mega* m = s->mega;
  if (s->started) return fail(rt, WGM_ERR_STATE, "server already started");
  int grid = 0;
  int rc = wgm_resolve_grid(rt, WGM_K_SERVE, s->cfg.wg_size, s->cfg.num_wgs, &grid);
  if (rc) return rc;
  if (grid < 2 + s->cfg.reserve) return fail(rt, WGM_ERR_INVALID, "grid %d leaves no serving workgroup", grid);
  m->wg_size = s->cfg.wg_size;
  m->sd.query_barrier = 0;
  m->sd.serve = 1;
  m->sd.reserve = s->cfg.reserve;
  hstore(m->h_flag, DEVICE_SCHEDULER_INIT);
  hstore(m->h_persistent, -1);
  hstore(m->h_stop, 0);
  hstore(m->h_participating, 0);
  hstore(s->h_ring_tail, 0);
  for (int i = 0; i < s->dev.slots; i++) {
    unsigned long long* w = (unsigned long long*)&s->h_recs[i];
    __hip_atomic_store(w, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  s->tail = 0;
  rc = wgm_reset_ctl(rt, 0, 0);
  if (rc) return rc;
  rc = mega_enqueue_sched_reset(m);
  if (rc) return rc;
  hipLaunchKernelGGL(k_serve_reset, dim3(16), dim3(256), 0, rt->stream, s->dev);
  hipLaunchKernelGGL(k_serve, dim3(grid), dim3(s->cfg.wg_size), serve_lds_bytes(s->cfg.wg_size), rt->stream, rt->ctl,
                     m->sd, s->dev, m->np);
  HIPCHK(rt, hipGetLastError());
  m->running = true;
  s->started = true;
  if (!hwait(m->h_flag, DEVICE_WAITING, rt->timeout_s)) return fail(rt, WGM_ERR_TIMEOUT, "server never reached DEVICE_WAITING");
  m->participating = hload(m->h_participating);
  s->participating = m->participating;
  if (participating) *participating = s->participating;
  const int k = s->participating - s->cfg.reserve;
  if (k < 1) return fail(rt, WGM_ERR_TOO_MANY_GROUPS, "%d participants leave no serving workgroup", s->participating);
  // the server is the persistent task (send_persistent_task)
  hstore(m->h_task_size, k);
  hstore(m->h_flag, DEVICE_TO_PERSISTENT_TASK);
  if (!hwait(m->h_flag, DEVICE_GOT_GROUPS, rt->timeout_s)) return fail(rt, WGM_ERR_TIMEOUT, "no DEVICE_GOT_GROUPS");
  hstore(m->h_flag, DEVICE_TO_EXECUTE);
  if (!hwait(m->h_flag, DEVICE_WAITING, rt->timeout_s)) return fail(rt, WGM_ERR_TIMEOUT, "no DEVICE_WAITING");
  return WGM_OK;
}

static int serve_submit_at(wgm_serve* s, int32_t graph, int32_t query, int32_t src, uint64_t t_arrival, int64_t* id);

int wgm_serve_submit(wgm_serve* s, int32_t graph, int32_t query, int32_t src, int64_t* id) {
  if (!s) return fail(nullptr, WGM_ERR_INVALID, "NULL server");
  return serve_submit_at(s, graph, query, src, now_ns(), id);
}

static int serve_submit_at(wgm_serve* s, int32_t graph, int32_t query, int32_t src, uint64_t t_arrival, int64_t* id) {
  wgm_runtime* rt = s->rt;
  if (graph < 0 || graph >= s->count) return fail(rt, WGM_ERR_INVALID, "graph index %d out of range", graph);
  if (query != WGM_QUERY_BFS && query != WGM_QUERY_SSSP) return fail(rt, WGM_ERR_INVALID, "unknown query %d", query);
  if (query == WGM_QUERY_SSSP && !s->has_wt[graph]) return fail(rt, WGM_ERR_INVALID, "graph %d has no weights", graph);
  if (src < 0 || src >= s->gn[graph]) return fail(rt, WGM_ERR_INVALID, "bad source %d", src);
  const int64_t bytes = state_words(s->gn[graph], query == WGM_QUERY_SSSP) * 4;
  if (bytes > s->arena_words * 4)
    return fail(rt, WGM_ERR_BUDGET, "request state (%lld bytes) can never fit the budget next to the graphs",
                (long long)bytes);
  Pending p{s->next_id++, graph, query, src, t_arrival};
  s->pending.push_back(p);
  s->outstanding++;
  if (id) *id = p.id;
  if (s->started) serve_progress(s);
  return WGM_OK;
}

int wgm_serve_poll(wgm_serve* s, wgm_serve_done* out, int32_t max, int32_t* count) {
  if (!s) return fail(nullptr, WGM_ERR_INVALID, "NULL server");
  if (s->started) serve_progress(s);
  int c = 0;
  while (c < max && !s->done.empty()) {
    if (out) out[c] = s->done.front();
    s->done.pop_front();
    s->outstanding--;
    c++;
  }
  if (count) *count = c;
  return WGM_OK;
}

int wgm_serve_task(wgm_serve* s, int32_t kind, int32_t groups, wgm_task_times* out) {
  if (!s) return fail(nullptr, WGM_ERR_INVALID, "NULL server");
  if (!s->started) return fail(s->rt, WGM_ERR_STATE, "server not started");
  if (kind < WGM_NP_MATMULT_I32 || kind > WGM_NP_REDUCE_MIN) return fail(s->rt, WGM_ERR_INVALID, "unknown task kind");
  if (groups < 1 || groups > s->participating)
    return fail(s->rt, WGM_ERR_TOO_MANY_GROUPS, "cannot give %d of %d workgroups to a task", groups, s->participating);
  return mega_run_task(s->mega, kind, groups, out);
}

int wgm_serve_stop(wgm_serve* s) {
  if (!s) return fail(nullptr, WGM_ERR_INVALID, "NULL server");
  wgm_runtime* rt = s->rt;
  if (!s->started) return WGM_OK;
  // drain: everything admitted finishes (bounded)
  const uint64_t t0 = now_ns();
  while (!s->pending.empty() || (int)s->free_slots.size() < s->dev.slots) {
    serve_progress(s);
    if ((double)(now_ns() - t0) * 1e-9 > rt->timeout_s) {
      (void)wgm_mega_quit(s->mega);
      s->started = false;
      return fail(rt, WGM_ERR_TIMEOUT, "requests still in flight at stop");
    }
    std::this_thread::yield();
  }
  const int rc = wgm_mega_quit(s->mega);
  s->started = false;
  return rc;
}

// Open-loop load: Poisson arrivals, queries submitted by this thread at
// their arrival times, non-persistent tasks run by a second thread in
// arrival order (the scheduler protocol is synchronous).
int wgm_serve_run_load(wgm_serve* s, const wgm_serve_load* load, wgm_serve_report* rep, wgm_serve_done* records,
                       int32_t* nrecords) {
  if (!s || !load || !rep) return fail(s ? s->rt : nullptr, WGM_ERR_INVALID, "NULL argument");
  wgm_runtime* rt = s->rt;
  if (!s->started) return fail(rt, WGM_ERR_STATE, "server not started");
  if (load->rate_per_s <= 0 || load->requests < 1 || load->np_fraction < 0 || load->np_fraction > 1)
    return fail(rt, WGM_ERR_INVALID, "bad load");
  const int np_groups = load->np_groups > 0 ? load->np_groups : 16;
  if (load->np_fraction > 0 && np_groups > s->participating)
    return fail(rt, WGM_ERR_TOO_MANY_GROUPS, "tasks of %d workgroups exceed %d", np_groups, s->participating);
  struct Arrival {
    uint64_t t;
    int task, kind, graph, query;
  };
  std::vector<Arrival> arr(load->requests);
  uint64_t st = load->seed ? load->seed : 6;
  auto u01 = [&]() { return ((double)(wgm_splitmix64_next(&st) >> 11) + 1.0) * 0x1p-53; };  // (0, 1]
  double t = 0.0;
  int ntask = 0;
  for (int i = 0; i < load->requests; i++) {
    t += -std::log(u01()) / load->rate_per_s * 1e9;
    Arrival a{(uint64_t)t, 0, 0, 0, 0};
    if (u01() <= load->np_fraction) {
      a.task = 1;
      a.kind = (ntask++ & 1) ? WGM_NP_REDUCE_MIN : WGM_NP_MATMUL_F32;
    } else {
      a.graph = (int)(((unsigned __int128)wgm_splitmix64_next(&st) * (unsigned)s->count) >> 64);
      const int r = a.graph % 4;
      a.query = (r == 0 || r == 3 || !s->has_wt[a.graph]) ? WGM_QUERY_BFS : WGM_QUERY_SSSP;
    }
    arr[i] = a;
  }
  const uint64_t base = now_ns() + 1000000;  // first arrival 1 ms from now
  std::vector<double> t_total, t_queue, t_resp, t_exec;
  std::atomic<int> task_errors{0};
  std::atomic<int> task_rc{WGM_OK};
  std::thread tasks([&]() {
    for (const Arrival& a : arr) {
      if (!a.task) continue;
      const uint64_t at = base + a.t;
      while (now_ns() < at) {
      }
      const uint64_t t0 = now_ns();
      wgm_task_times tt{};
      const int rc = mega_run_task(s->mega, a.kind, np_groups, &tt);
      const uint64_t t1 = now_ns();
      if (rc == WGM_ERR_MISMATCH) task_errors++;
      else if (rc) {
        task_rc = rc;
        return;
      }
      t_total.push_back((double)(t1 - at) * 1e-3);
      t_queue.push_back((double)(t0 - at) * 1e-3);
      t_resp.push_back(tt.response_us);
      t_exec.push_back(tt.execution_us);
    }
  });
  std::vector<double> q_lat;
  int64_t scanned = 0;
  int nrec = 0;
  uint64_t last_done = base;
  wgm_serve_done buf[64];
  auto harvest = [&]() {
    int32_t c = 0;
    do {
      wgm_serve_poll(s, buf, 64, &c);
      const uint64_t now = now_ns();
      for (int i = 0; i < c; i++) {
        q_lat.push_back(buf[i].queue_us + buf[i].service_us);
        scanned += buf[i].edges_scanned;
        if (records && nrec < load->requests) records[nrec] = buf[i];
        nrec++;
        last_done = now;
      }
    } while (c == 64);
  };
  int rc = WGM_OK;
  for (const Arrival& a : arr) {
    if (a.task) continue;
    const uint64_t at = base + a.t;
    while (now_ns() < at) harvest();
    rc = serve_submit_at(s, a.graph, a.query, 0, at, nullptr);
    if (rc) break;
  }
  const uint64_t t_wait = now_ns();
  while (rc == WGM_OK && s->outstanding > 0) {
    harvest();
    if ((double)(now_ns() - t_wait) * 1e-9 > rt->timeout_s) rc = fail(rt, WGM_ERR_TIMEOUT, "queries still outstanding");
  }
  tasks.join();
  if (rc == WGM_OK && task_rc != WGM_OK) rc = task_rc;
  const uint64_t end = std::max(last_done, now_ns());
  wgm_serve_report r{};
  r.queries = (int32_t)q_lat.size();
  r.tasks = (int32_t)t_total.size();
  r.task_errors = task_errors;
  r.wall_s = (double)(end - base) * 1e-9;
  r.throughput_rps = r.wall_s > 0 ? (double)(r.queries + r.tasks) / r.wall_s : 0.0;
  r.gteps = r.wall_s > 0 ? (double)scanned / r.wall_s * 1e-9 : 0.0;
  r.query_p50_us = pct(q_lat, 0.5);
  r.query_p99_us = pct(q_lat, 0.99);
  r.query_mean_us = mean(q_lat);
  r.task_p50_us = pct(t_total, 0.5);
  r.task_p99_us = pct(t_total, 0.99);
  r.task_mean_us = mean(t_total);
  r.task_queue_p99_us = pct(t_queue, 0.99);
  r.task_response_p99_us = pct(t_resp, 0.99);
  r.task_exec_p99_us = pct(t_exec, 0.99);
  r.max_inflight_bytes = (double)s->max_inflight_bytes;
  *rep = r;
  if (nrecords) *nrecords = std::min(nrec, load->requests);
  return rc;
}

}  // extern "C"
