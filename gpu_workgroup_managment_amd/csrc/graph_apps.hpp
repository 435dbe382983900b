// graph_apps.hpp -- persistent graph kernels for gfx950 on the wgm runtime.
//
// Frontier BFS / SSSP (Lonestar drelax2: src/lonestar/graph_apps/bfs/device/bfs.cl:104-302,
// src/lonestar/graph_apps/sssp/device/sssp.cl:104-318), topology-driven
// colouring and MIS (Pannotia color_combined / mis_combined:
// src/pannotia/graph_apps/color/device/color.cl:3-97,
// src/pannotia/graph_apps/mis/device/mis.cl:5-125) and pull PageRank.
//
// The per-level / per-round bodies take (rank, nteam): the workgroup's index
// inside the set of workgroups that share the graph, and that set's size. The
// standalone kernels pass the discovered id and count; the batched kernel
// passes a team; the mega-kernel passes the resizable persistent ids.
#pragma once

#include "wgm_device.hpp"

namespace wgm {

constexpr uint32_t kInf = 1000000000u;  // MYINFINITY (src/lonestar/graph_apps/bfs/device/bfs.cl:33)

struct GraphDev {
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* wt;
  const uint32_t* prio;
  int32_t n;
  int32_t m;
  // SSSP: one word per edge, (col << 8) | wt, for graphs whose weights fit 8
  // bits (nullptr: col and wt); one load per relaxed edge instead of two
  const uint32_t* pw;
  // in-edge CSR (sources of each vertex's in-edges); a team's BFS may then
  // take its wide levels bottom-up (nullptr: top-down only)
  const int32_t* rrow;
  const int32_t* rcol;
};

// Frontier state. Level L's frontier is q[L & 1] with its length in
// counts[(L % 3) * kLine]; three counters let one barrier end a level: during
// level L the workgroups read counter L%3, push into (L+1)%3 and rank 0 zeroes
// (L+2)%3, whose last readers finished before the previous barrier. The
// claim counters (dynamic chunking) rotate the same way.
struct FrontierDev {
  uint32_t* dist;
  uint32_t* stamp;  // SSSP: last level v was queued for (dedups pushes)
  int32_t* q0;
  int32_t* q1;
  int* counts;      // 3 lines
  int* claims;      // 3 lines (dynamic chunk claiming), may be null
  int chunk;        // frontier vertices per claim; <= 0: static round-robin chunks
  // SSSP with distance-tagged entries (PAIRS): qa[i] is the distance q[i] was
  // queued with; an entry whose vertex has since dropped lower is stale and
  // skipped (its newer entry carries the lower distance). One atomic per
  // improving edge instead of two; a level queues at most one entry per
  // relaxation, so each queue holds max(n, m) + 1 entries.
  uint32_t* qa0;
  uint32_t* qa1;
  // row start << 3 | out-degree of each entry's vertex (the batched
  // low-degree path, which loads it when it queues the entry)
  uint32_t* qr0;
  uint32_t* qr1;
  // when set: edges relaxed (out-degrees of the live entries processed) are
  // added here, one atomic per frontier batch -- the traversal's work
  unsigned long long* relaxed;
  // teams (COH) with heavy > 0 (the standalone apps, one team of every
  // participant; the batched launch's teams of <= 16 gain nothing from it):
  // an entry of more than `heavy` out-edges is
  // deferred to the heavy list (vertex, distance, claim counter; fewer than
  // m / heavy + 1 entries per level) and its edges are then claimed piece by
  // piece by the whole team (frontier_heavy). hcnt: 3 lines, by level % 3.
  int32_t* hv;
  uint32_t* hdu;
  int* hclaim;
  int* hcnt;
  int heavy;
};

// Heavy-entry threshold of a team with bd-thread workgroups: a workgroup's
// frontier batch spreads up to about bd entries' edges evenly, so an entry
// with several batches' worth of edges would hold back the workgroup that
// claimed it. Never below kHeavyMin, which sizes the heavy lists.
constexpr int kHeavyMin = 1024;
__host__ __device__ constexpr int heavy_threshold(int bd) { return 4 * bd > kHeavyMin ? 4 * bd : kHeavyMin; }

// Dynamic LDS layout for one workgroup of bd threads: end[bd], beg[bd], du[bd], wave[17], claim[1]
__host__ __device__ constexpr int frontier_lds_bytes(int bd) { return (3 * bd + 18) * 4; }

// LDS-typed pointer: loads, stores and atomics through it stay ds_*
// instructions (through a generic pointer they become flat accesses, which
// wait on both the LDS and the memory counters).
typedef __attribute__((address_space(3))) uint32_t lds_u32;

// An empty register barrier: the compiler may not fold a later use of v back
// into the conditional block that produced it, where it would wait for the
// load or atomic before the next slot's is issued.
__device__ __forceinline__ void opaque(uint32_t& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void opaque(int& v) { asm volatile("" : "+v"(v)); }

// Relax edge idx of a source at distance du (SSSP; BFS claims the target
// for level `next`), then the wave pushes what improved: one counter atomic
// per wave for the next level's queue (with distance tags under PAIRS) and
// one for the far queue. Called by every lane of the wave; `active` marks
// lanes with an edge.
// bits (one workgroup owning the graph, LOCAL): one bit per vertex in LDS --
// BFS marks a vertex visited there and the discoverer stores its level (no
// read of dist and no atomic on it); SSSP keeps its far flags there instead
// of in_far.
template <bool SSSP, bool LOCAL, bool COH, bool PAIRS>
__device__ __forceinline__ void relax_push(const GraphDev& g, const FrontierDev& s, bool active, int idx, uint32_t du,
                                           int32_t* out_q, uint32_t* out_qa, int* out_cnt, uint32_t next,
                                           uint32_t thr, int32_t* far_q, int* far_cnt, uint32_t* in_far,
                                           lds_u32* bits = nullptr) {
  constexpr int scope = LOCAL ? __HIP_MEMORY_SCOPE_WORKGROUP : __HIP_MEMORY_SCOPE_AGENT;
  bool want = false, far_want = false;
  int u = 0;
  uint32_t alt = 0;
  if (active) {
    uint32_t w = 0u;
    if (SSSP && g.pw) {
      const uint32_t word = g.pw[idx];
      u = (int)(word >> 8);
      w = word & 255u;
    } else {
      u = g.col[idx];
      if (SSSP) w = (uint32_t)g.wt[idx];
    }
    const uint32_t bit = 1u << (u & 31);
    if (!SSSP && LOCAL && bits) {
      want = (__hip_atomic_fetch_or(&bits[u >> 5], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & bit) == 0u;
      if (want) s.dist[u] = next;
    } else if (!SSSP) {
      if (s.dist[u] == kInf) {
        uint32_t expect = kInf;
        want = __hip_atomic_compare_exchange_strong(&s.dist[u], &expect, next, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                    scope);
      }
    } else {
      alt = du + w;
      if (alt < s.dist[u]) {
        const uint32_t old = __hip_atomic_fetch_min(&s.dist[u], alt, __ATOMIC_RELAXED, scope);
        if (alt < old) {
          if (alt < thr) want = PAIRS || __hip_atomic_fetch_max(&s.stamp[u], next, __ATOMIC_RELAXED, scope) < next;
          else if (LOCAL && bits)
            far_want = (__hip_atomic_fetch_or(&bits[u >> 5], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) &
                        bit) == 0u;
          else far_want = __hip_atomic_exchange(&in_far[u], 1u, __ATOMIC_RELAXED, scope) == 0u;
        }
      }
    }
  }
  const unsigned long long mask = __ballot(want);
  if (mask) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = __hip_atomic_fetch_add(out_cnt, __popcll(mask), __ATOMIC_RELAXED, scope);
    base = __builtin_amdgcn_readlane(base, leader);  // leader is wave-uniform: no LDS round trip
    if (want) {
      const int pos = base + __popcll(mask & ((1ull << lane) - 1ull));
      if (COH) {
        __hip_atomic_store(&out_q[pos], u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (PAIRS) __hip_atomic_store(&out_qa[pos], alt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        out_q[pos] = u;
        if (PAIRS) out_qa[pos] = alt;
      }
    }
  }
  if (SSSP && far_q) {
    const unsigned long long fmask = __ballot(far_want);
    if (fmask) {
      const int lane = threadIdx.x & 63;
      const int leader = __ffsll((long long)fmask) - 1;
      int base = 0;
      if (lane == leader) base = __hip_atomic_fetch_add(far_cnt, __popcll(fmask), __ATOMIC_RELAXED, scope);
      base = __builtin_amdgcn_readlane(base, leader);  // leader is wave-uniform: no LDS round trip
      if (far_want) {
        int32_t* slot = &far_q[base + __popcll(fmask & ((1ull << lane) - 1ull))];
        if (COH) __hip_atomic_store(slot, u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else *slot = u;
      }
    }
  }
}

// relax_push for U edges per lane: edges idx[0..U) of sources at distances
// du[]. Each step's loads or atomics are issued for every slot before any
// result is used (the results pass through opaque() after the issuing loop),
// so a lane waits for one round trip per step instead of U; the wave then
// pushes all U slots' next-level entries with one counter atomic and its far
// entries with one more. Called by every lane of the wave; `active` marks the
// slots with an edge.
template <int U, bool SSSP, bool LOCAL, bool COH, bool PAIRS>
__device__ __forceinline__ void relax_push_n(const GraphDev& g, const FrontierDev& s, const bool* active,
                                             const int* idx, const uint32_t* du, int32_t* out_q, uint32_t* out_qa,
                                             int* out_cnt, uint32_t next, uint32_t thr, int32_t* far_q, int* far_cnt,
                                             uint32_t* in_far, lds_u32* bits = nullptr) {
  constexpr int scope = LOCAL ? __HIP_MEMORY_SCOPE_WORKGROUP : __HIP_MEMORY_SCOPE_AGENT;
  const bool packed = SSSP && g.pw;
  uint32_t wd[U], wt[U];
#pragma unroll
  for (int x = 0; x < U; x++) {
    wd[x] = 0u;
    wt[x] = 0u;
    if (active[x]) {
      if (packed) {
        wd[x] = g.pw[idx[x]];
      } else {
        wd[x] = (uint32_t)g.col[idx[x]];
        if (SSSP) wt[x] = (uint32_t)g.wt[idx[x]];
      }
    }
  }
  int u[U];
  uint32_t alt[U], cur[U];
#pragma unroll
  for (int x = 0; x < U; x++) {
    opaque(wd[x]);
    opaque(wt[x]);
    u[x] = packed ? (int)(wd[x] >> 8) : (int)wd[x];
    alt[x] = SSSP ? du[x] + (packed ? (wd[x] & 255u) : wt[x]) : next;
    // the target's distance filters the atomic (a stale value is never
    // below the current one); BFS with visited bits claims in one LDS step
    cur[x] = 0u;
    if (active[x]) {
      if (!SSSP && LOCAL && bits) {
        const uint32_t bit = 1u << (u[x] & 31);
        cur[x] = (__hip_atomic_fetch_or(&bits[u[x] >> 5], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & bit)
                     ? 0u : kInf;
      } else {
        cur[x] = s.dist[u[x]];
      }
    }
  }
  uint32_t old[U];  // the target's distance before this slot's update (0: no update)
#pragma unroll
  for (int x = 0; x < U; x++) {
    opaque(cur[x]);
    old[x] = 0u;
    if (!SSSP && LOCAL && bits) {
      old[x] = cur[x];
    } else if (!SSSP) {
      if (active[x] && cur[x] == kInf) {
        uint32_t expect = kInf;  // left at kInf by a successful exchange
        __hip_atomic_compare_exchange_strong(&s.dist[u[x]], &expect, next, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             scope);
        old[x] = expect;
      }
    } else if (active[x] && alt[x] < cur[x]) {
      old[x] = __hip_atomic_fetch_min(&s.dist[u[x]], alt[x], __ATOMIC_RELAXED, scope);
    }
  }
  bool near[U], farc[U];
  uint32_t res[U];  // stamp before the update (near), or the far flag's previous state
#pragma unroll
  for (int x = 0; x < U; x++) {
    opaque(old[x]);
    const bool better = SSSP ? alt[x] < old[x] : old[x] == kInf;
    near[x] = better && (!SSSP || alt[x] < thr);
    farc[x] = SSSP && better && alt[x] >= thr;
    res[x] = 0u;
    if (!SSSP) {
      if (LOCAL && bits && near[x]) s.dist[u[x]] = next;
    } else if (near[x]) {
      if (!PAIRS) res[x] = __hip_atomic_fetch_max(&s.stamp[u[x]], next, __ATOMIC_RELAXED, scope);
    } else if (farc[x]) {
      const uint32_t bit = 1u << (u[x] & 31);
      if (LOCAL && bits)
        res[x] = __hip_atomic_fetch_or(&bits[u[x] >> 5], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & bit;
      else
        res[x] = __hip_atomic_exchange(&in_far[u[x]], 1u, __ATOMIC_RELAXED, scope);
    }
  }
  bool want[U], far_want[U];
  unsigned long long mask[U], fmask[U];
  int cnt = 0, fcnt = 0;
#pragma unroll
  for (int x = 0; x < U; x++) {
    opaque(res[x]);
    want[x] = near[x] && (!SSSP || PAIRS || res[x] < next);
    far_want[x] = farc[x] && res[x] == 0u;
    mask[x] = __ballot(want[x]);
    fmask[x] = __ballot(far_want[x]);
    cnt += __popcll(mask[x]);
    fcnt += __popcll(fmask[x]);
  }
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  const int leader = __ffsll((long long)__ballot(1)) - 1;  // the first active lane
  if (cnt) {
    int base = 0;
    if (lane == leader) base = __hip_atomic_fetch_add(out_cnt, cnt, __ATOMIC_RELAXED, scope);
    base = __builtin_amdgcn_readlane(base, leader);  // leader is wave-uniform: no LDS round trip
#pragma unroll
    for (int x = 0; x < U; x++) {
      if (want[x]) {
        const int pos = base + __popcll(mask[x] & below);
        if (COH) {
          __hip_atomic_store(&out_q[pos], u[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (PAIRS) __hip_atomic_store(&out_qa[pos], alt[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          out_q[pos] = u[x];
          if (PAIRS) out_qa[pos] = alt[x];
        }
      }
      base += __popcll(mask[x]);
    }
  }
  if (SSSP && far_q && fcnt) {
    int base = 0;
    if (lane == leader) base = __hip_atomic_fetch_add(far_cnt, fcnt, __ATOMIC_RELAXED, scope);
    base = __builtin_amdgcn_readlane(base, leader);
#pragma unroll
    for (int x = 0; x < U; x++) {
      if (far_want[x]) {
        int32_t* slot = &far_q[base + __popcll(fmask[x] & below)];
        if (COH) __hip_atomic_store(slot, u[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else *slot = u[x];
      }
      base += __popcll(fmask[x]);
    }
  }
}

// Edges each thread takes per pass of frontier_batch<..., U> in the
// standalone SSSP kernel: their loads and atomics are in flight together
// (relax_push_n), so a pass over a batch's edges costs about one chain of
// round trips per kEdgeSlots edges per thread. RMAT scale 20 (ef 16) push
// SSSP, 64 x 1024 workgroups: 3.99 ms at one slot, 3.52 at two, 3.29 at
// three, 3.20 at four. BFS keeps one: its levels are short (ER 2^16: 0.099 ms
// at one slot, 0.110 at four). The batched, serving and merged kernels keep
// one too: their kernels' other paths are scheduled for a register budget
// that the extra slots' state would exceed (the batched kernel's
// single-workgroup grid SSSP ran 4% slower alone, 9% inside the batch).
constexpr int kEdgeSlots = 4;

// Relax the edges of frontier entries in_q[a, b) (b - a <= blockDim.x).
// The workgroup scans the entries' degrees and spreads their edges evenly
// over its threads, U per thread per pass (edges e, e + bd, ...), each
// finding its source entry by binary search in LDS (with U > 1 the slots'
// searches step together), so a high-degree vertex does not serialise one
// thread. BFS claims a vertex
// with one CAS (INF -> level+1) and pushes it exactly once; SSSP lowers dist
// with atomic min and pushes a vertex at most once per level through
// `stamp`. Pushes are wave-aggregated: one counter atomic per wave.
// LOCAL: the push counter is in LDS and only this workgroup touches the
// graph (workgroup-scope atomics). Reference: processnode2 / processedge2 /
// wl_push_1item (src/lonestar/graph_apps/bfs/device/bfs.cl:82-212).
// COH (teams without cache maintenance at their barriers): the state other
// workgroups write -- queue entries, a frontier vertex's own distance -- is
// read and written with agent-scope atomic accesses, which stay coherent
// across the XCDs' L2s without flushing or invalidating them; the read-only
// CSR stays cached. Plain reads of dist[u] only filter candidates (a stale
// value is never below the current one and the atomic decides).
// Near-far (SSSP, far_q != nullptr): a vertex whose distance drops to
// alt >= thr is not queued for the next level but deferred to far_q, at most
// once while its in_far flag is set; the caller processes it in a later
// band. Final distances are the same (they are unique); the work is closer
// to Dijkstra's order.
template <bool SSSP, bool LOCAL, bool COH = false, bool PAIRS = false, int U = 1>
__device__ __forceinline__ void frontier_batch(const GraphDev& g, const FrontierDev& s, const int32_t* in_q, int a,
                                               int b, int32_t* out_q, int* out_cnt, uint32_t next, int* lds,
                                               uint32_t thr = 0xffffffffu, int32_t* far_q = nullptr,
                                               int* far_cnt = nullptr, uint32_t* in_far = nullptr,
                                               const uint32_t* in_qa = nullptr, uint32_t* out_qa = nullptr,
                                               int* heavy_cnt = nullptr, lds_u32* bits = nullptr) {
  const int bd = blockDim.x, tid = threadIdx.x;
  int* s_end = lds;
  int* s_beg = lds + bd;
  uint32_t* s_du = (uint32_t*)(lds + 2 * bd);
  int* s_wave = lds + 3 * bd;
  const int i = a + tid;
  int beg = 0, deg = 0, hv_v = 0;
  uint32_t du = 0;
  bool hvy = false;
  if (i < b) {
    const int v = COH ? __hip_atomic_load(&in_q[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : in_q[i];
    beg = g.row_ptr[v];
    deg = g.row_ptr[v + 1] - beg;
    if (SSSP) du = COH ? __hip_atomic_load(&s.dist[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : s.dist[v];
    if (PAIRS) {
      const uint32_t tag = COH ? __hip_atomic_load(&in_qa[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : in_qa[i];
      if (tag != du) deg = 0;  // stale: v was queued again with a lower distance
    }
    if (COH && heavy_cnt && deg > s.heavy) {
      hvy = true;
      hv_v = v;
    }
  }
  if (COH && heavy_cnt) {  // defer heavy entries to the team-wide pass
    const unsigned long long hm = __ballot(hvy);
    if (hm) {
      const int lane = tid & 63;
      const int leader = __ffsll((long long)hm) - 1;
      int base = 0;
      if (lane == leader) base = __hip_atomic_fetch_add(heavy_cnt, __popcll(hm), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      base = __builtin_amdgcn_readlane(base, leader);  // leader is wave-uniform: no LDS round trip
      if (hvy) {
        const int pos = base + __popcll(hm & ((1ull << lane) - 1ull));
        __hip_atomic_store(&s.hv[pos], hv_v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&s.hdu[pos], du, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&s.hclaim[pos], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s.relaxed) atomicAdd(s.relaxed, (unsigned long long)deg);
        deg = 0;
      }
    }
  }
  int total;
  const int excl = wg_exclusive_scan(deg, s_wave, &total);
  if (s.relaxed && tid == 0 && total > 0) atomicAdd(s.relaxed, (unsigned long long)total);
  s_end[tid] = excl + deg;
  s_beg[tid] = beg - excl;
  if (SSSP) s_du[tid] = du;
  __syncthreads();
  if constexpr (U == 1) {
    for (int eb = 0; eb < total; eb += bd) {
      const int e = eb + tid;
      int idx = 0;
      uint32_t du_e = 0;
      if (e < total) {
        int lo = 0, hi = bd - 1;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (s_end[mid] > e) hi = mid; else lo = mid + 1;
        }
        idx = s_beg[lo] + e;
        du_e = SSSP ? s_du[lo] : 0u;
      }
      relax_push<SSSP, LOCAL, COH, PAIRS>(g, s, e < total, idx, du_e, out_q, out_qa, out_cnt, next, thr, far_q,
                                          far_cnt, in_far, bits);
    }
  } else {
    // bisection steps that narrow bd candidates to one: ceil(log2 bd)
    const int steps = 32 - __clz(bd - 1);
    for (int eb = 0; eb < total; eb += U * bd) {
      int e[U], lo[U], hi[U], idx[U];
      bool act[U];
      uint32_t du_e[U];
#pragma unroll
      for (int x = 0; x < U; x++) {
        e[x] = eb + x * bd + tid;
        act[x] = e[x] < total;
        lo[x] = 0;
        hi[x] = act[x] ? bd - 1 : 0;
      }
      // the first entry whose end is past e: the slots' LDS reads of one step
      // are issued together (a settled slot re-reads its own entry)
      for (int k = 0; k < steps; k++) {
#pragma unroll
        for (int x = 0; x < U; x++) {
          const int mid = (lo[x] + hi[x]) >> 1;
          const bool past = s_end[mid] > e[x];
          const bool open = lo[x] < hi[x];
          hi[x] = open && past ? mid : hi[x];
          lo[x] = open && !past ? mid + 1 : lo[x];
        }
      }
#pragma unroll
      for (int x = 0; x < U; x++) {
        idx[x] = act[x] ? s_beg[lo[x]] + e[x] : 0;
        du_e[x] = SSSP && act[x] ? s_du[lo[x]] : 0u;
      }
      relax_push_n<U, SSSP, LOCAL, COH, PAIRS>(g, s, act, idx, du_e, out_q, out_qa, out_cnt, next, thr, far_q,
                                               far_cnt, in_far, bits);
    }
  }
  __syncthreads();
}

// Near-far state of one traversal: two far queues (n + 1 entries each) and
// their counters, the in_far flags, the band minimum and width.
// A band split drops the far entries whose distance is already below the
// finished band's threshold: any improvement below the threshold queues the
// vertex near, so such a vertex was relaxed at that distance in the band
// (moving it back to near would relax its edges twice), and the band minimum
// is taken over the rest, which all lie at or above that threshold.
struct NearFar {
  int32_t* q[2];
  // select, not index: a runtime index into q[] puts the pair in scratch
  __device__ __forceinline__ int32_t* far(int i) const { return i ? q[1] : q[0]; }
  int* cnt;          // lines: [0] count of q[0], [kLine] count of q[1], [2 kLine] band minimum
  uint32_t* flag;    // in_far[n]
  uint32_t delta;    // band width; 0: plain frontier Bellman-Ford
};

// One level of frontier relaxation over q[level & 1] by a team of nteam
// workgroups. Static mode: chunks of bd entries round-robin over the team.
// Dynamic mode (s.chunk > 0): workgroups claim up to s.chunk entries at a
// time from a per-level counter, which balances uneven chunks. A narrow
// level is claimed in smaller pieces -- about two per member, at least 64
// entries -- so that it still spreads over the team: with s.chunk entries per
// claim, a level of a few thousand entries would keep all but a few members
// idle while those walk every edge. Members may claim different sizes (a
// claim takes [c0, c0 + its size) from the counter), so the claimed ranges
// still tile the level.
template <bool SSSP, bool COH = false, bool PAIRS = false, int U = 1>
__device__ __forceinline__ void frontier_level(const GraphDev& g, const FrontierDev& s, int level,
                                               int rank, int nteam, int* lds, uint32_t thr = 0xffffffffu,
                                               int32_t* far_q = nullptr, int* far_cnt = nullptr,
                                               uint32_t* in_far = nullptr) {
  const int bd = blockDim.x, tid = threadIdx.x;
  const int in_n = ld_relaxed(&s.counts[(level % 3) * kLine]);
  const int32_t* in_q = (level & 1) ? s.q1 : s.q0;
  int32_t* out_q = (level & 1) ? s.q0 : s.q1;
  const uint32_t* in_qa = (level & 1) ? s.qa1 : s.qa0;
  uint32_t* out_qa = (level & 1) ? s.qa0 : s.qa1;
  int* out_cnt = &s.counts[((level + 1) % 3) * kLine];
  const uint32_t next = (uint32_t)level + 1u;
  int* heavy_cnt = COH && s.heavy > 0 ? &s.hcnt[(level % 3) * kLine] : nullptr;
  if (s.chunk > 0 && s.claims) {
    int* s_claim = lds + 3 * bd + 17;
    int* claim = &s.claims[(level % 3) * kLine];
    const int chunk = min(s.chunk, max(64, (in_n / (2 * nteam) + 63) & ~63));
    while (true) {
      if (tid == 0) *s_claim = add_relaxed(claim, chunk);
      __syncthreads();
      const int c0 = *s_claim;
      __syncthreads();
      if (c0 >= in_n) break;
      const int c1 = min(c0 + chunk, in_n);
      for (int a = c0; a < c1; a += bd)
        frontier_batch<SSSP, false, COH, PAIRS, U>(g, s, in_q, a, min(a + bd, c1), out_q, out_cnt, next, lds, thr,
                                                   far_q, far_cnt, in_far, in_qa, out_qa, heavy_cnt);
    }
    if (rank == 0 && tid == 0) st_relaxed(&s.claims[((level + 2) % 3) * kLine], 0);
  } else {
    for (int a = rank * bd; a < in_n; a += nteam * bd)
      frontier_batch<SSSP, false, COH, PAIRS, U>(g, s, in_q, a, min(a + bd, in_n), out_q, out_cnt, next, lds, thr,
                                                 far_q, far_cnt, in_far, in_qa, out_qa, heavy_cnt);
  }
  if (rank == 0 && tid == 0) {
    st_relaxed(&s.counts[((level + 2) % 3) * kLine], 0);
    // the heavy count of level - 2: every member read it (right after the
    // barrier that ended level - 2's relaxation) before the barrier that
    // ended level - 1's; level - 1's own count may still be being read
    if (heavy_cnt) st_relaxed(&s.hcnt[((level + 1) % 3) * kLine], 0);
  }
}

// Second pass of a team's level over the heavy entries frontier_batch
// deferred. Their out-edges are taken as one list, bd entries at a time: the
// workgroup scans the bd entries' degrees (as frontier_batch does), and every
// workgroup of the team claims bd consecutive edges of that list per claim
// (one counter per bd entries, hclaim[chunk], zeroed when the entries were
// appended), so a hub's edges spread over the whole team and a workgroup
// claims about (heavy edges / bd) / team size times per chunk, not once per
// heavy entry. Runs between two team barriers; the pushes go to the same
// next-level queue.
template <bool SSSP, bool PAIRS>
__device__ __forceinline__ void frontier_heavy(const GraphDev& g, const FrontierDev& s, int level, int* lds,
                                               uint32_t thr = 0xffffffffu, int32_t* far_q = nullptr,
                                               int* far_cnt = nullptr, uint32_t* in_far = nullptr) {
  const int bd = blockDim.x, tid = threadIdx.x;
  int* s_end = lds;
  int* s_beg = lds + bd;
  uint32_t* s_du = (uint32_t*)(lds + 2 * bd);
  int* s_wave = lds + 3 * bd;
  int* s_claim = lds + 3 * bd + 17;
  int32_t* out_q = (level & 1) ? s.q0 : s.q1;
  uint32_t* out_qa = (level & 1) ? s.qa0 : s.qa1;
  int* out_cnt = &s.counts[((level + 1) % 3) * kLine];
  const uint32_t next = (uint32_t)level + 1u;
  const int nh = __builtin_amdgcn_readfirstlane(ld_relaxed(&s.hcnt[(level % 3) * kLine]));
  for (int h0 = 0; h0 < nh; h0 += bd) {
    int beg = 0, deg = 0;
    uint32_t du = 0;
    if (h0 + tid < nh) {
      const int v = __hip_atomic_load(&s.hv[h0 + tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      du = __hip_atomic_load(&s.hdu[h0 + tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      beg = g.row_ptr[v];
      deg = g.row_ptr[v + 1] - beg;
    }
    int total;
    const int excl = wg_exclusive_scan(deg, s_wave, &total);
    s_end[tid] = excl + deg;
    s_beg[tid] = beg - excl;
    s_du[tid] = du;
    int* claim = &s.hclaim[h0 / bd];
    while (true) {
      // the scan's stores above, or the last claim's reads of s_claim, come first
      __syncthreads();
      if (tid == 0) *s_claim = __hip_atomic_fetch_add(claim, bd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      const int c0 = __builtin_amdgcn_readfirstlane(*s_claim);
      if (c0 >= total) break;
      const int e = c0 + tid;
      int idx = 0;
      uint32_t du_e = 0;
      if (e < total) {
        int lo = 0, hi = bd - 1;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (s_end[mid] > e) hi = mid; else lo = mid + 1;
        }
        idx = s_beg[lo] + e;
        du_e = s_du[lo];
      }
      relax_push<SSSP, false, true, PAIRS>(g, s, e < total, idx, du_e, out_q, out_qa, out_cnt, next, thr, far_q,
                                           far_cnt, in_far);
    }
    __syncthreads();  // every thread has read s_end before the next chunk's scan rewrites it
  }
}

// ----------------------------------------------------------- colouring
// One barrier per round instead of the reference's two: during round c a
// neighbour is "uncoloured at the start of the round" iff its colour is -1 or
// c, so the max-priority test and the colour write share one pass with the
// same outcome as colour1 -> barrier -> colour2. The degree-1 quirk
// (`start != end - 1`, src/pannotia/graph_apps/color/device/color.cl:46) is kept:
// a vertex with exactly one neighbour ignores it. The stop flag of round c is
// line c % 3 of `flags`.
// The standalone kernels (the worklist rounds below) take vertices
// kTopoGroup to a thread per pass (i, i + bd, i + 2 bd, ...: each load
// coalesced across the wave) and issue the group's work level by level --
// their rows and priorities, then the first kTopoEdges edges of each, then
// those neighbours' states and priorities -- so a thread waits for a few
// memory round trips per group instead of a dependent chain per vertex and
// per edge. Edges past the first kTopoEdges of a vertex are walked one by one
// afterwards. The registers this takes (~90 VGPRs) would spill in the
// mega-kernel, whose resizable tasks keep one vertex at a time (color_round,
// mis_phase_a / _b).
constexpr int kTopoGroup = 4;
constexpr int kTopoEdges = 4;

__device__ __forceinline__ bool color_round(const GraphDev& g, int32_t* color, int round, int lo, int hi) {
  bool any = false;
  for (int v = lo + (int)threadIdx.x; v < hi; v += blockDim.x) {
    if (color[v] != -1) continue;
    const int beg = g.row_ptr[v], end = g.row_ptr[v + 1];
    long long mx = -1;
    if (end - beg != 1) {
      for (int e = beg; e < end; e++) {
        const int u = g.col[e];
        const int cu = color[u];
        if (cu == -1 || cu == round) {
          any = true;
          const long long pu = (long long)g.prio[u];
          mx = pu > mx ? pu : mx;
        }
      }
    }
    if ((long long)g.prio[v] > mx) color[v] = round;
  }
  return any;
}

// ------------------------------------------------------------------ MIS
// State per vertex: -1 undecided, 2r joined in round r, 2r+1 removed in
// round r. Phase A (= mis1 + the join half of mis2): an undecided vertex joins
// if its priority is below every neighbour undecided at the round start
// (state -1 or 2r). Phase B (= the neighbour half of mis2 + mis3): an
// undecided vertex with a neighbour that joined this round is removed.
// Two barriers per round instead of three; the sets are those of
// src/pannotia/graph_apps/mis/device/mis.cl:21-124.
// The standalone kernel's forms (mis_phase_a_wl / _b_wl) group their loads
// as the colour rounds do.
__device__ __forceinline__ bool mis_phase_a(const GraphDev& g, int32_t* st, int32_t* s_out, int round,
                                            int lo, int hi) {
  bool any = false;
  const int joined = 2 * round;
  for (int v = lo + (int)threadIdx.x; v < hi; v += blockDim.x) {
    if (st[v] != -1) continue;
    any = true;
    long long mn = 0x100000000LL;  // above every uint32 priority
    for (int e = g.row_ptr[v]; e < g.row_ptr[v + 1]; e++) {
      const int u = g.col[e];
      const int su = st[u];
      if (su == -1 || su == joined) {
        const long long pu = (long long)g.prio[u];
        mn = pu < mn ? pu : mn;
      }
    }
    if ((long long)g.prio[v] < mn) {
      st[v] = joined;
      s_out[v] = 2;
    }
  }
  return any;
}

__device__ __forceinline__ void mis_phase_b(const GraphDev& g, int32_t* st, int round, int lo, int hi) {
  const int joined = 2 * round;
  for (int v = lo + (int)threadIdx.x; v < hi; v += blockDim.x) {
    if (st[v] != -1) continue;
    for (int e = g.row_ptr[v]; e < g.row_ptr[v + 1]; e++) {
      if (st[g.col[e]] == joined) {
        st[v] = joined + 1;
        break;
      }
    }
  }
}

// ------------------------------------------------------- worklist rounds
// The standalone colour and MIS kernels visit only the vertices still open:
// a workgroup keeps the open vertices of its block in its share ([lo, ...))
// of one of two lists, visits that list in the next round and writes the
// vertices that stay open to the other. A vertex's decision depends only on
// its neighbours' states, read as in the topology-driven rounds, and only its
// own workgroup writes it, so the colours, sets and round counts are the same.
// On the 2048^2 grid colour's 16 rounds visit 13.9 M vertices instead of
// 67.1 M. The first round visits the whole block (no list).
__device__ __forceinline__ void wl_push(bool stay, int v, int32_t* lout, int lo, int* cnt) {
  const unsigned long long mask = __ballot(stay);
  if (mask) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = __hip_atomic_fetch_add(cnt, __popcll(mask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    base = __builtin_amdgcn_readlane(base, leader);
    if (stay) lout[lo + base + __popcll(mask & ((1ull << lane) - 1ull))] = v;
  }
}

// The i-th open vertex of the block: list entry, or lo + i in the first round.
__device__ __forceinline__ int wl_vertex(const int32_t* lin, int lo, int i) { return lin ? lin[lo + i] : lo + i; }

__device__ __forceinline__ bool color_round_wl(const GraphDev& g, int32_t* color, int round, int lo, int cnt,
                                               const int32_t* lin, int32_t* lout, int* ocnt) {
  bool any = false;
  const int bd = blockDim.x;
  for (int base = (int)threadIdx.x; base < cnt; base += kTopoGroup * bd) {
    int v[kTopoGroup], beg[kTopoGroup], deg[kTopoGroup];
    uint32_t pv[kTopoGroup];
    bool open[kTopoGroup];
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++) {
      open[k] = base + k * bd < cnt;
      v[k] = open[k] ? wl_vertex(lin, lo, base + k * bd) : 0;
    }
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++) {
      beg[k] = open[k] ? g.row_ptr[v[k]] : 0;
      const int end = open[k] ? g.row_ptr[v[k] + 1] : 0;
      pv[k] = open[k] ? g.prio[v[k]] : 0u;
      deg[k] = end - beg[k];
      if (deg[k] == 1) deg[k] = 0;  // a vertex with exactly one neighbour ignores it
    }
    int u[kTopoGroup][kTopoEdges];
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++)
#pragma unroll
      for (int j = 0; j < kTopoEdges; j++) u[k][j] = j < deg[k] ? g.col[beg[k] + j] : -1;
    int cu[kTopoGroup][kTopoEdges];
    uint32_t pu[kTopoGroup][kTopoEdges];
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++)
#pragma unroll
      for (int j = 0; j < kTopoEdges; j++) {
        cu[k][j] = u[k][j] >= 0 ? color[u[k][j]] : 0;
        pu[k][j] = u[k][j] >= 0 ? g.prio[u[k][j]] : 0u;
      }
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++) {
      bool coloured = false;
      if (open[k]) {
        long long mx = -1;
#pragma unroll
        for (int j = 0; j < kTopoEdges; j++) {
          if (u[k][j] >= 0 && (cu[k][j] == -1 || cu[k][j] == round)) {
            any = true;
            mx = (long long)pu[k][j] > mx ? (long long)pu[k][j] : mx;
          }
        }
        for (int e = beg[k] + kTopoEdges; e < beg[k] + deg[k]; e++) {
          const int w = g.col[e];
          const int cw = color[w];
          if (cw == -1 || cw == round) {
            any = true;
            const long long pw = (long long)g.prio[w];
            mx = pw > mx ? pw : mx;
          }
        }
        coloured = (long long)pv[k] > mx;
        if (coloured) color[v[k]] = round;
      }
      wl_push(open[k] && !coloured, v[k], lout, lo, ocnt);
    }
  }
  return any;
}

// MIS phase A over the open list: an open vertex joins if its priority is
// below every neighbour undecided at the round start (state -1 or 2r).
__device__ __forceinline__ bool mis_phase_a_wl(const GraphDev& g, int32_t* st, int32_t* s_out, int round, int lo,
                                               int cnt, const int32_t* lin) {
  bool any = false;
  const int joined = 2 * round;
  const int bd = blockDim.x;
  for (int base = (int)threadIdx.x; base < cnt; base += kTopoGroup * bd) {
    int v[kTopoGroup], beg[kTopoGroup], deg[kTopoGroup];
    uint32_t pv[kTopoGroup];
    bool open[kTopoGroup];
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++) {
      open[k] = base + k * bd < cnt;
      v[k] = open[k] ? wl_vertex(lin, lo, base + k * bd) : 0;
    }
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++) {
      beg[k] = open[k] ? g.row_ptr[v[k]] : 0;
      deg[k] = (open[k] ? g.row_ptr[v[k] + 1] : 0) - beg[k];
      pv[k] = open[k] ? g.prio[v[k]] : 0u;
    }
    int u[kTopoGroup][kTopoEdges];
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++)
#pragma unroll
      for (int j = 0; j < kTopoEdges; j++) u[k][j] = j < deg[k] ? g.col[beg[k] + j] : -1;
    int su[kTopoGroup][kTopoEdges];
    uint32_t pu[kTopoGroup][kTopoEdges];
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++)
#pragma unroll
      for (int j = 0; j < kTopoEdges; j++) {
        su[k][j] = u[k][j] >= 0 ? st[u[k][j]] : 0;
        pu[k][j] = u[k][j] >= 0 ? g.prio[u[k][j]] : 0u;
      }
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++) {
      if (!open[k]) continue;
      any = true;
      long long mn = 0x100000000LL;  // above every uint32 priority
#pragma unroll
      for (int j = 0; j < kTopoEdges; j++)
        if (u[k][j] >= 0 && (su[k][j] == -1 || su[k][j] == joined))
          mn = (long long)pu[k][j] < mn ? (long long)pu[k][j] : mn;
      for (int e = beg[k] + kTopoEdges; e < beg[k] + deg[k]; e++) {
        const int w = g.col[e];
        const int sw = st[w];
        if (sw == -1 || sw == joined) {
          const long long pw = (long long)g.prio[w];
          mn = pw < mn ? pw : mn;
        }
      }
      if ((long long)pv[k] < mn) {
        st[v[k]] = joined;
        s_out[v[k]] = 2;
      }
    }
  }
  return any;
}

// MIS phase B over the same list: a vertex that did not join and has a
// neighbour that joined this round is removed; the rest stay open.
__device__ __forceinline__ void mis_phase_b_wl(const GraphDev& g, int32_t* st, int round, int lo, int cnt,
                                               const int32_t* lin, int32_t* lout, int* ocnt) {
  const int joined = 2 * round;
  const int bd = blockDim.x;
  for (int base = (int)threadIdx.x; base < cnt; base += kTopoGroup * bd) {
    int v[kTopoGroup], sv[kTopoGroup], beg[kTopoGroup], deg[kTopoGroup];
    bool open[kTopoGroup];
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++) {
      const bool in = base + k * bd < cnt;
      v[k] = in ? wl_vertex(lin, lo, base + k * bd) : 0;
      sv[k] = in ? st[v[k]] : 0;
      open[k] = in && sv[k] == -1;
    }
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++) {
      beg[k] = open[k] ? g.row_ptr[v[k]] : 0;
      deg[k] = (open[k] ? g.row_ptr[v[k] + 1] : 0) - beg[k];
    }
    int u[kTopoGroup][kTopoEdges];
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++)
#pragma unroll
      for (int j = 0; j < kTopoEdges; j++) u[k][j] = j < deg[k] ? g.col[beg[k] + j] : -1;
    int su[kTopoGroup][kTopoEdges];
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++)
#pragma unroll
      for (int j = 0; j < kTopoEdges; j++) su[k][j] = u[k][j] >= 0 ? st[u[k][j]] : 0;
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++) {
      bool hit = false;
      if (open[k]) {
#pragma unroll
        for (int j = 0; j < kTopoEdges; j++) hit = hit || (u[k][j] >= 0 && su[k][j] == joined);
        for (int e = beg[k] + kTopoEdges; e < beg[k] + deg[k] && !hit; e++) hit = st[g.col[e]] == joined;
        if (hit) st[v[k]] = joined + 1;
      }
      wl_push(open[k] && !hit, v[k], lout, lo, ocnt);
    }
  }
}

// ------------------------------------------------------------ PageRank
// Pull iteration r'(v) = base + d * sum_{u in N(v)} r(u)/deg(u), FP32, the sum
// taken in CSR order by one thread, so the result is bitwise reproducible.
// contrib holds r(u)/deg(u) (0 for deg 0).
__device__ __forceinline__ void pagerank_init(const GraphDev& g, float* contrib, float r0, int lo, int hi) {
  for (int v = lo + (int)threadIdx.x; v < hi; v += blockDim.x) {
    const int deg = g.row_ptr[v + 1] - g.row_ptr[v];
    contrib[v] = deg ? r0 / (float)deg : 0.0f;
  }
}

// GROUPED (the standalone kernel): kTopoGroup vertices per thread per pass,
// their rows, then the first kTopoEdges edges of each, then those
// neighbours' contributions, all in flight together (three round trips per
// group instead of two per edge); each sum still adds its terms one by one in
// edge order, so the result is the same bit for bit.
template <bool GROUPED = false>
__device__ __forceinline__ void pagerank_iter(const GraphDev& g, const float* cin, float* cout, float* rank,
                                              float damping, float base, bool last, int lo, int hi) {
  if (!GROUPED) {
    for (int v = lo + (int)threadIdx.x; v < hi; v += blockDim.x) {
      const int beg = g.row_ptr[v], end = g.row_ptr[v + 1];
      float sum = 0.0f;
      for (int e = beg; e < end; e++) sum = sum + cin[g.col[e]];
      const float r = __builtin_fmaf(damping, sum, base);
      if (last) rank[v] = r;
      else cout[v] = (end > beg) ? r / (float)(end - beg) : 0.0f;
    }
    return;
  }
  const int bd = blockDim.x;
  for (int b0 = lo + (int)threadIdx.x; b0 < hi; b0 += kTopoGroup * bd) {
    int v[kTopoGroup], beg[kTopoGroup], end[kTopoGroup];
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++) {
      v[k] = b0 + k * bd;
      const bool in = v[k] < hi;
      beg[k] = in ? g.row_ptr[v[k]] : 0;
      end[k] = in ? g.row_ptr[v[k] + 1] : 0;
    }
    int u[kTopoGroup][kTopoEdges];
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++)
#pragma unroll
      for (int j = 0; j < kTopoEdges; j++) u[k][j] = beg[k] + j < end[k] ? g.col[beg[k] + j] : -1;
    float cu[kTopoGroup][kTopoEdges];
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++)
#pragma unroll
      for (int j = 0; j < kTopoEdges; j++) cu[k][j] = u[k][j] >= 0 ? cin[u[k][j]] : 0.0f;
#pragma unroll
    for (int k = 0; k < kTopoGroup; k++) {
      if (v[k] >= hi) continue;
      float sum = 0.0f;
#pragma unroll
      for (int j = 0; j < kTopoEdges; j++)
        if (u[k][j] >= 0) sum = sum + cu[k][j];
      for (int e = beg[k] + kTopoEdges; e < end[k]; e++) sum = sum + cin[g.col[e]];
      const float r = __builtin_fmaf(damping, sum, base);
      if (last) rank[v[k]] = r;
      else cout[v[k]] = (end[k] > beg[k]) ? r / (float)(end[k] - beg[k]) : 0.0f;
    }
  }
}

// Contiguous vertex block of workgroup `rank` among `nteam` (keeps grid
// neighbours in one workgroup and so in one XCD's L2).
__device__ __forceinline__ void block_range(int n, int rank, int nteam, int* lo, int* hi) {
  const long long per = ((long long)n + nteam - 1) / nteam;
  long long a = per * rank, b = a + per;
  if (a > n) a = n;
  if (b > n) b = n;
  *lo = (int)a;
  *hi = (int)b;
}

}  // namespace wgm
