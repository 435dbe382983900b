// team_frontier.hpp -- BFS / SSSP over one graph by a team of workgroups
// that share state through agent-scope atomics (the batched launch's teams
// and the standalone frontier apps).
#pragma once

#include "graph_apps.hpp"

namespace wgm {

// Wave-aggregated push of `item` (and its distance tag into qa, when given)
// where `want`, one counter atomic per wave.
template <bool COH>
__device__ __forceinline__ void push_item(bool want, int item, int32_t* q, int* cnt, int scope,
                                          uint32_t* qa = nullptr, uint32_t tag = 0) {
  const unsigned long long mask = __ballot(want);
  if (!mask) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (lane == leader) {
    base = scope == __HIP_MEMORY_SCOPE_WORKGROUP
               ? __hip_atomic_fetch_add(cnt, __popcll(mask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
               : __hip_atomic_fetch_add(cnt, __popcll(mask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  base = __builtin_amdgcn_readlane(base, leader);  // leader is wave-uniform: no LDS round trip
  if (want) {
    const int pos = base + __popcll(mask & ((1ull << lane) - 1ull));
    if (COH) {
      __hip_atomic_store(&q[pos], item, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (qa) __hip_atomic_store(&qa[pos], tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      q[pos] = item;
      if (qa) qa[pos] = tag;
    }
  }
}

__device__ __forceinline__ int wave_min_int(int x) {
  for (int off = 32; off > 0; off >>= 1) x = min(x, __shfl_xor(x, off, 64));
  return x;
}

// Counter barrier of a team of k workgroups (k small). Arrival is a fetch_add
// on a monotonic counter; the arrival that completes epoch e publishes
// generation e. Everything the team shares is accessed with agent-scope
// atomics (frontier_batch<..., COH>), so the barrier only has to drain this
// workgroup's outstanding accesses before arriving: no L2 write-back or
// invalidation, which would also evict every other job's cached graph on the
// XCD.
__device__ __forceinline__ void team_barrier(int* cnt, int* gen, int k, int& epoch) {
  drain_vmem();
  __syncthreads();
  epoch++;
  if (threadIdx.x == 0) {
    const int old = add_relaxed(cnt, 1);
    if (old == epoch * k - 1) {
      st_relaxed(gen, epoch);
    } else {
      while (ld_relaxed(gen) < epoch) {
      }
    }
  }
  __syncthreads();
}

// A BFS level taken bottom-up (direction-optimizing BFS, Beamer et al.,
// SC 2012): every unvisited vertex of the member's range looks for an
// in-neighbour in the frontier (dist == level), four in-edges per round with
// an early exit, and if it finds one takes level + 1 and queues itself. A
// vertex is written only by its own thread, so no atomic on dist is needed;
// the distances are read and written with agent-scope accesses (coherent
// across the XCDs' L2s, as the top-down step's). The set of vertices it
// discovers is the top-down level's: the unvisited vertices with an in-edge
// from the frontier. The level's counters are rotated as frontier_level
// rotates them.
// A level runs bottom-up while the frontier holds more than n / kBottomUpDiv
// vertices, or more than the vertices not yet reached (the tail of a BFS,
// where a few unvisited vertices scan their in-edges instead of the frontier
// its out-edges).
constexpr int kBottomUpDiv = 16;

__device__ __forceinline__ void frontier_bottom_up(const GraphDev& g, const FrontierDev& s, int level, int rank,
                                                   int nteam) {
  const int bd = blockDim.x, tid = threadIdx.x, lane = tid & 63;
  int32_t* out_q = (level & 1) ? s.q0 : s.q1;
  int* out_cnt = &s.counts[((level + 1) % 3) * kLine];
  const uint32_t cur = (uint32_t)level, next = (uint32_t)level + 1u;
  int lo, hi;
  block_range(g.n, rank, nteam, &lo, &hi);
  for (int base = lo; base < hi; base += bd) {
    const int v = base + tid;
    bool found = false;
    if (v < hi && __hip_atomic_load(&s.dist[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == kInf) {
      const int b = g.rrow[v], e = g.rrow[v + 1];
      for (int j = b; j < e && !found; j += 4) {
        int u[4];
#pragma unroll
        for (int q = 0; q < 4; q++) u[q] = j + q < e ? g.rcol[j + q] : -1;
        uint32_t du[4];
#pragma unroll
        for (int q = 0; q < 4; q++)
          du[q] = u[q] >= 0 ? __hip_atomic_load(&s.dist[u[q]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kInf;
#pragma unroll
        for (int q = 0; q < 4; q++) found = found || du[q] == cur;
      }
      if (found) __hip_atomic_store(&s.dist[v], next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const unsigned long long mask = __ballot(found);
    if (mask) {
      const int leader = __ffsll((long long)mask) - 1;
      int pos = 0;
      if (lane == leader) pos = __hip_atomic_fetch_add(out_cnt, __popcll(mask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      pos = __builtin_amdgcn_readlane(pos, leader) + __popcll(mask & ((1ull << lane) - 1ull));
      if (found) __hip_atomic_store(&out_q[pos], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (rank == 0 && tid == 0) {
    st_relaxed(&s.counts[((level + 2) % 3) * kLine], 0);
    if (s.chunk > 0 && s.claims) st_relaxed(&s.claims[((level + 2) % 3) * kLine], 0);
    if (s.heavy > 0) st_relaxed(&s.hcnt[((level + 1) % 3) * kLine], 0);
  }
}

// A graph served by a team of k workgroups (rank 0..k-1). Everything the
// team shares -- queue entries and their distance tags, a frontier vertex's
// distance, counters -- goes through agent-scope atomics, so `bar` (the
// team's barrier) only has to drain, never to write back or invalidate an
// L2. SSSP runs near-far bands of width nf.delta (0: frontier
// Bellman-Ford), with the far counters and band minimum in device lines.
// U: edges per thread per frontier pass (frontier_batch). HEAVY: teams whose
// FrontierDev sets heavy > 0 (the standalone kernels) run the heavy-entry
// pass; the batched kernel's teams (heavy = 0) are built without it.
// Reference: drelax2 (src/lonestar/graph_apps/bfs/device/bfs.cl:214-302,
// src/lonestar/graph_apps/sssp/device/sssp.cl:262-318).
template <bool SSSP, int U = 1, bool HEAVY = false, class Barrier>
__device__ void run_team(const GraphDev& g, const FrontierDev& s, const NearFar& nf, int src, int rank, int k,
                         int* lds, int* iters_out, Barrier&& bar) {
  const bool near_far = SSSP && nf.delta > 0;
  int lo, hi;
  block_range(g.n, rank, k, &lo, &hi);
  // initial state through the coherent path (other members' atomics see it)
  for (int v = lo + (int)threadIdx.x; v < hi; v += blockDim.x) {
    __hip_atomic_store(&s.dist[v], (v == src) ? 0u : kInf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (near_far) __hip_atomic_store(&nf.flag[v], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (rank == 0 && threadIdx.x == 0) {
    __hip_atomic_store(&s.q0[0], src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (SSSP) __hip_atomic_store(&s.qa0[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    st_relaxed(&s.counts[0], 1);
    st_relaxed(&s.counts[kLine], 0);
    st_relaxed(&s.counts[2 * kLine], 0);
    for (int l = 0; l < 3; l++) {
      st_relaxed(&s.claims[l * kLine], 0);
      if (s.heavy > 0) st_relaxed(&s.hcnt[l * kLine], 0);
    }
    if (near_far) {
      st_relaxed(&nf.cnt[0], 0);
      st_relaxed(&nf.cnt[kLine], 0);
      st_relaxed(&nf.cnt[2 * kLine], INT32_MAX);
    }
  }
  bar();
  const int bd = blockDim.x;
  int level = 0, fi = 0;
  int seen = 0;  // BFS: vertices reached so far, this level's frontier included
  uint32_t thr = near_far ? nf.delta : 0xffffffffu;
  while (true) {
    int in_n;
    while ((in_n = ld_relaxed(&s.counts[(level % 3) * kLine])) != 0) {
      // every member read the same counts after the barriers: the same choice
      seen += in_n;
      if (!SSSP && g.rrow && (in_n > g.n / kBottomUpDiv || in_n > g.n - seen))
        frontier_bottom_up(g, s, level, rank, k);
      else
        frontier_level<SSSP, true, SSSP, U>(g, s, level, rank, k, lds, thr, near_far ? nf.far(fi) : nullptr,
                                            &nf.cnt[fi * kLine], nf.flag);
      bar();
      // heavy entries deferred by the pass (read after the barrier: every
      // member sees the same count and takes the same number of barriers)
      if (HEAVY && s.heavy > 0 && ld_relaxed(&s.hcnt[(level % 3) * kLine]) != 0) {
        frontier_heavy<SSSP, SSSP>(g, s, level, lds, thr, near_far ? nf.far(fi) : nullptr, &nf.cnt[fi * kLine],
                                   nf.flag);
        bar();
      }
      level++;
    }
    if (!near_far) break;
    const int fn = ld_relaxed(&nf.cnt[fi * kLine]);
    if (fn == 0) break;
    int mn = INT32_MAX;
    const uint32_t t_old = thr;  // far entries below it were relaxed in the band (dropped)
    for (int i = rank * bd + (int)threadIdx.x; i < fn; i += k * bd) {
      const int v = __hip_atomic_load(&nf.far(fi)[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t d = __hip_atomic_load(&s.dist[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (d >= t_old) mn = min(mn, (int)d);
    }
    mn = wave_min_int(mn);
    if ((threadIdx.x & 63) == 0 && mn != INT32_MAX)
      __hip_atomic_fetch_min(&nf.cnt[2 * kLine], mn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bar();
    thr = (uint32_t)ld_relaxed(&nf.cnt[2 * kLine]) + nf.delta;
    int32_t* lq = (level & 1) ? s.q1 : s.q0;  // this level's (empty) queue
    uint32_t* lqa = (level & 1) ? s.qa1 : s.qa0;
    for (int base = rank * bd; base < fn; base += k * bd) {
      const int i = base + (int)threadIdx.x;
      bool to_near = false, to_far = false;
      int v = 0;
      uint32_t dv = 0;
      if (i < fn) {
        v = __hip_atomic_load(&nf.far(fi)[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        dv = __hip_atomic_load(&s.dist[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (dv < thr) {
          __hip_atomic_store(&nf.flag[v], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          to_near = dv >= t_old;
        } else {
          to_far = true;
        }
      }
      push_item<true>(to_near, v, lq, &s.counts[(level % 3) * kLine], __HIP_MEMORY_SCOPE_AGENT, lqa, dv);
      push_item<true>(to_far, v, nf.far(1 - fi), &nf.cnt[(1 - fi) * kLine], __HIP_MEMORY_SCOPE_AGENT);
    }
    bar();
    // every member has read q[fi] and the minimum; the next split follows at
    // least one more team barrier (the minimum moved to the near queue)
    if (rank == 0 && threadIdx.x == 0) {
      st_relaxed(&nf.cnt[fi * kLine], 0);
      st_relaxed(&nf.cnt[2 * kLine], INT32_MAX);
    }
    fi = 1 - fi;
  }
  if (rank == 0 && threadIdx.x == 0) *iters_out = level;
}

}  // namespace wgm
