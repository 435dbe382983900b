// batched.hip -- many graphs in one persistent launch (device work queue),
// the per-graph relaunch baseline, hipGraph capture and the autotuner.
//
// Not in the reference, which runs one graph per process
// (src/host_skeleton/skel.cpp:250-385). The per-graph body is the same
// frontier relaxation as the standalone BFS / SSSP kernels
// (src/lonestar/graph_apps/bfs/device/bfs.cl:214-302,
// src/lonestar/graph_apps/sssp/device/sssp.cl:262-318).
//
// Work distribution: job j (one graph + query) owns `team_j` consecutive
// tickets; a workgroup that draws ticket t serves job j = the last job whose
// first ticket is <= t, as team rank t - first_j. Tickets are drawn in job
// order (largest graphs first), so at most one team is partially formed at
// any time and its missing members are the next free workgroups: with every
// puller co-resident (discovery), no team waits forever.

#include "apps_internal.hpp"
#include "team_frontier.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

using namespace wgm;

namespace wgm {

// Low-degree graphs (out-degree <= kLowDegree: grids, meshes, road-like)
// traverse thousands of narrow levels, each a chain of memory round trips:
// they take the degree-bounded frontier step, and a team's cross-workgroup
// barrier would add to every level, so one workgroup serves them up to
// kLowDegreeSingle edges. Measured on the 512-query batch: with 2-workgroup
// teams for the 1M-edge grids the batch took 17.4 ms instead of 10.2 ms.
constexpr int kLowDegree = 4;
constexpr long long kLowDegreeSingle = 1ll << 22;
// Frontier entries per thread group in flight together. A narrow level's
// cost is instruction issue, and each unrolled slot adds its scalar checks
// to every level: on the 512-query batch, 4 slots took 8.99 ms per step,
// 2 took 8.75 and 1 took 8.64 (levels past bd / K entries take more passes).
constexpr int kLowSlots = 1;

// The low-degree step keeps the first low_lds_cap(bd) entries of each of its
// two frontier queues (vertex, row info, distance tag) in LDS behind the
// frontier step's own LDS; entries past that spill to the HBM queues at the
// same index. A level then reads its entries from LDS and ends without
// waiting on queue stores to reach L2. 2 * bd entries keep a 1024-thread
// workgroup under 64 KB of LDS, and two such workgroups fit one CU.
__host__ __device__ constexpr int low_lds_cap(int bd) { return 2 * bd; }
// A 1024-thread workgroup is alone on its CU (its register and wave counts
// allow one), so it may also hold the far flags of a low-degree SSSP job
// as an LDS bitmap of kFarBitWords words (graphs up to 2^18 vertices):
// setting a far flag is then an LDS atomic instead of a round trip to L2.
constexpr int kFarBitWords = 8192;
__host__ __device__ constexpr int batch_far_bits_bytes(int bd) { return bd >= 1024 ? 4 * kFarBitWords : 0; }

// Packed rows of low-degree graphs: kLowDegree words per vertex, (target << 8)
// | weight, kEllNone past the degree. A graph gets them when every weight fits
// 8 bits and its vertex ids fit 24 (kEllNone is no valid word then).
constexpr uint32_t kEllNone = 0xffffffffu;
// vertices per thread per round of a reached-edge sum (loads in flight)
constexpr int kReachPer = 8;
constexpr int32_t kEllMaxN = (1 << 24) - 1;
// Distances kept in the L2 (low_ell == 2): a low-degree SSSP query on one
// 1024-thread workgroup updates its distances with plain loads and stores,
// which stay in its XCD's L2, instead of atomics, which are all performed
// beyond it. Each pass first folds its candidates per target in an LDS hash
// table of kCombSlots 64-bit words, (distance << 32) | (target << 11) |
// candidate (slot and thread, so that words are unique and the minimum has
// one owner), open addressing with linear probes. A pass takes kCombU
// entries per 4-lane group, at most kCombU * blockDim.x <= 2048 candidates,
// so a table is at most half full and every probe sequence ends.
// Three tables serve consecutive passes in turn: the winner of a slot clears
// it after the next pass's barrier, when every reader of the slot is past it,
// and the pass after that (the slot's next user) starts after the clear.
// Passes are the cost of a level (each is two memory round trips and two
// barriers): two entries per group halve the passes of the levels wider than
// blockDim.x / 4 entries.
constexpr int kCombBits = 12;
constexpr int kCombSlots = 1 << kCombBits;
constexpr int kCombU = 2;
constexpr unsigned long long kCombEmpty = ~0ull;
// Its own LDS layout: two frontier queues of kCombCap (vertex, tag) entries,
// the far bitmap, the three tables (152 KB at 1024 threads).
constexpr int kCombCap = 1536;
__host__ __device__ constexpr int comb_lds_bytes(int bd) {
  return bd >= 1024 ? 2 * 2 * 4 * kCombCap + 4 * kFarBitWords + 3 * kCombSlots * 8 : 0;
}
__host__ __device__ constexpr int batch_lds_bytes(int bd) {
  return frontier_lds_bytes(bd) + 2 * 3 * 4 * low_lds_cap(bd) + batch_far_bits_bytes(bd) > comb_lds_bytes(bd)
             ? frontier_lds_bytes(bd) + 2 * 3 * 4 * low_lds_cap(bd) + batch_far_bits_bytes(bd)
             : comb_lds_bytes(bd);
}
typedef __attribute__((address_space(3))) unsigned long long lds_u64;

// The per-vertex bitmap: batch_far_bits_bytes(bd) after the low-degree queues.
__device__ __forceinline__ lds_u32* batch_bits(int* lds) {
  return (lds_u32*)(reinterpret_cast<uint32_t*>(lds) +
                    (frontier_lds_bytes((int)blockDim.x) + 2 * 3 * 4 * low_lds_cap((int)blockDim.x)) / 4);
}
// The combine layout's far bitmap and tables.
__device__ __forceinline__ lds_u32* comb_bits(int* lds) { return (lds_u32*)lds + 2 * 2 * kCombCap; }
__device__ __forceinline__ lds_u64* batch_comb(int* lds) { return (lds_u64*)(comb_bits(lds) + kFarBitWords); }

__device__ __forceinline__ unsigned long long comb_word(int u, uint32_t alt, int cand) {
  return ((unsigned long long)alt << 32) | ((unsigned long long)(uint32_t)u << 11) | (uint32_t)cand;
}
// Slot of target u (< 2^21) in a combine table, where its word is folded in
// by a minimum: the first candidate of u claims an empty slot, later ones
// find u there. -1 only if the table were full, which a pass cannot make it.
__device__ __forceinline__ int comb_insert(lds_u64* t, int u, unsigned long long mine) {
  int h = (int)(((uint32_t)u * 2654435761u) >> (32 - kCombBits));
  for (int probe = 0; probe < kCombSlots; probe++) {
    unsigned long long old = kCombEmpty;
    if (__hip_atomic_compare_exchange_strong(&t[h], &old, mine, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP))
      return h;
    if (((uint32_t)old >> 11) == (uint32_t)u) {
      __hip_atomic_fetch_min(&t[h], mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      return h;
    }
    h = (h + 1) & (kCombSlots - 1);
  }
  return -1;
}

// Workgroup barrier that orders LDS only: memory loads stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

struct BatchJob {
  long long row_off;  // into the row_ptr pool
  long long col_off;  // into the col / wt pools
  long long v_off;    // into the per-vertex pools (dist, in_far)
  long long q_off;    // into the queue pools: two queues of qcap entries (q and, for SSSP, distance tags qa)
  long long f_off;    // into the far-queue pool (2 * (n + 1) entries) and the in_far pool (n)
  int n, m, query, team, src, ticket;
  int index;          // caller's graph index
  int maxdeg;         // largest out-degree (picks the degree-bounded frontier step)
  int qcap;           // entries per queue: n + 1 (BFS), max(n, m) + 1 (SSSP, one entry per relaxation)
  uint32_t delta;     // SSSP near-far band width of this query (0: off)
  long long e_off;    // into the packed-row pool (kLowDegree words per vertex), -1: none
  long long p_off;    // into the packed-edge pool (one word per edge, SSSP), -1: none
};

struct BatchDev {
  const BatchJob* jobs;
  int njobs;
  int total_tickets;
  const int32_t* row;
  const int32_t* col;
  const int32_t* wt;
  uint32_t* dist;
  int32_t* q;
  uint32_t* qa;    // distance tags of SSSP queue entries (same offsets as q)
  uint32_t* qr;    // row_info of queue entries (low-degree jobs)
  int* counts;     // 3 lines per job
  int* claims;     // 3 lines per job
  int chunk;
  int* tbar;       // 2 lines per job: arrival counter, generation
  int* ticket;     // one line
  int* iters;      // per job
  int max_team;    // largest team in the job table
  int32_t* far;    // near-far: far queues
  uint32_t* in_far;
  int* fcnt;       // 3 lines per job: far counts and band minimum
  unsigned long long* jt;  // per job: first member's start, last member's end (wall clock)
  unsigned long long* reached;  // per job: out-degrees of the reached vertices, summed after the traversal
  unsigned long long* relaxed;  // per job: edges relaxed from live frontier entries (work)
  unsigned long long* snap;     // control cells copied for the host: clock[0], clock[1], count, short team
  int far_bits;                 // 1: low-degree single-workgroup queries keep per-vertex bits in LDS (far / visited)
  const uint32_t* ell;          // packed rows of the low-degree graphs (BatchJob::e_off)
  const uint32_t* pw;           // packed edges of the other SSSP graphs (BatchJob::p_off)
  int low_ell;                  // 1: low-degree single-workgroup queries read the packed rows
};


// A graph served by one workgroup: level counters live in LDS and the
// workgroup's own barrier orders the levels; all atomics are workgroup scope.
// SSSP with nf.delta > 0 runs near-far bands: levels relax only below the
// band's threshold; when the near frontier empties, the threshold moves to
// the smallest deferred distance + delta and the far queue is split.
// lds_flags: one bit per vertex in the LDS bitmap (batch_bits) -- BFS's
// visited marks, SSSP's far flags -- instead of per-vertex words in HBM.
template <bool SSSP>
__device__ void run_single(const GraphDev& g, const FrontierDev& s, const NearFar& nf, int src, int* lds, int* s_cnt,
                           int* iters_out, bool lds_flags) {
  const bool near_far = SSSP && nf.delta > 0;
  __shared__ int s_far[3];  // far counts of q[0], q[1]; band minimum
  lds_u32* bits = lds_flags && (near_far || !SSSP) ? batch_bits(lds) : nullptr;
  for (int v = threadIdx.x; v < g.n; v += blockDim.x) {
    s.dist[v] = (v == src) ? 0u : kInf;
    if (near_far && !bits) nf.flag[v] = 0u;
  }
  if (bits)
    for (int w = threadIdx.x; w < (g.n + 31) / 32; w += blockDim.x)
      bits[w] = !SSSP && w == (src >> 5) ? 1u << (src & 31) : 0u;
  if (threadIdx.x == 0) {
    s.q0[0] = src;
    if (SSSP) s.qa0[0] = 0u;
    s_cnt[0] = 1;
    s_cnt[1] = 0;
    s_cnt[2] = 0;
    s_far[0] = 0;
    s_far[1] = 0;
    s_far[2] = INT32_MAX;
  }
  __syncthreads();
  const int bd = blockDim.x;
  int level = 0, fi = 0;
  uint32_t thr = near_far ? nf.delta : 0xffffffffu;
  while (true) {
    for (;; level++) {
      const int in_n = s_cnt[level % 3];
      if (in_n == 0) break;
      const int32_t* in_q = (level & 1) ? s.q1 : s.q0;
      int32_t* out_q = (level & 1) ? s.q0 : s.q1;
      const uint32_t* in_qa = (level & 1) ? s.qa1 : s.qa0;
      uint32_t* out_qa = (level & 1) ? s.qa0 : s.qa1;
      for (int a = 0; a < in_n; a += bd)
        frontier_batch<SSSP, true, false, SSSP>(g, s, in_q, a, min(a + bd, in_n), out_q, &s_cnt[(level + 1) % 3],
                                                (uint32_t)level + 1u, lds, thr, near_far ? nf.far(fi) : nullptr,
                                                &s_far[fi], nf.flag, in_qa, out_qa, nullptr, bits);
      if (threadIdx.x == 0) s_cnt[(level + 2) % 3] = 0;
      __syncthreads();
    }
    if (!near_far) break;
    const int fn = s_far[fi];
    if (fn == 0) break;
    const uint32_t t_old = thr;  // far entries below it were relaxed in the band (dropped)
    int mn = INT32_MAX;
    for (int i = threadIdx.x; i < fn; i += bd) {
      const uint32_t d = s.dist[nf.far(fi)[i]];
      if (d >= t_old) mn = min(mn, (int)d);
    }
    mn = wave_min_int(mn);
    if ((threadIdx.x & 63) == 0) atomicMin(&s_far[2], mn);
    __syncthreads();
    thr = (uint32_t)s_far[2] + nf.delta;
    int32_t* lq = (level & 1) ? s.q1 : s.q0;  // this level's (empty) queue
    uint32_t* lqa = (level & 1) ? s.qa1 : s.qa0;
    for (int base = 0; base < fn; base += bd) {
      const int i = base + threadIdx.x;
      bool to_near = false, to_far = false;
      int v = 0;
      uint32_t dv = 0;
      if (i < fn) {
        v = nf.far(fi)[i];
        dv = s.dist[v];
        if (dv < thr) {
          // each deferred vertex is in one far queue once: it moves once
          if (bits) __hip_atomic_fetch_and(&bits[v >> 5], ~(1u << (v & 31)), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
          else nf.flag[v] = 0u;
          to_near = dv >= t_old;
        } else {
          to_far = true;  // stays flagged in the other far queue
        }
      }
      push_item<false>(to_near, v, lq, &s_cnt[level % 3], __HIP_MEMORY_SCOPE_WORKGROUP, lqa, dv);
      push_item<false>(to_far, v, nf.far(1 - fi), &s_far[1 - fi], __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      s_far[fi] = 0;
      s_far[2] = INT32_MAX;
    }
    fi = 1 - fi;
    __syncthreads();
  }
  if (threadIdx.x == 0) *iters_out = level;
}

// Row start and out-degree of v packed into one queue word (row start
// < 2^29, degree <= 7), loaded by the producer alongside its atomic so the
// consumer skips a dependent round trip.
__device__ __forceinline__ uint32_t row_info(const GraphDev& g, int v) {
  const int rs = g.row_ptr[v];
  return ((uint32_t)rs << 3) | (uint32_t)(g.row_ptr[v + 1] - rs);
}


// Frontier queues of the low-degree step: entry i of queue p (0 or 1) in
// LDS below cap, in the HBM queue at index i above.
// LDS-typed (lds_u32), so that the two paths of load/store stay ds_* and
// global_* instructions: through a generic pointer the compiler merges them
// into one flat access, which then waits on both the LDS and the memory
// counters.

struct LowQueues {
  lds_u32* lds;  // [queue][vertex, row info, tag][cap]; F = 2: [queue][vertex, tag][cap] (no row info)
  int cap;
  const FrontierDev* s;
  int F;
  // rows: the entry carries row_info (CSR rows; packed rows need none)
  __device__ __forceinline__ void load(int p, int i, int& v, uint32_t& ri, uint32_t& du, bool tag,
                                       bool rows = true) const {
    ri = 0u;
    if (i < cap) {
      const lds_u32* b = lds + p * F * cap;
      v = (int)b[i];
      if (rows) ri = b[cap + i];
      du = tag ? b[(F - 1) * cap + i] : 0u;
    } else {
      v = (p ? s->q1 : s->q0)[i];
      if (rows) ri = (p ? s->qr1 : s->qr0)[i];
      du = tag ? (p ? s->qa1 : s->qa0)[i] : 0u;
    }
  }
  __device__ __forceinline__ void store(int p, int i, int v, uint32_t ri, uint32_t du, bool tag,
                                        bool rows = true) const {
    if (i < cap) {
      lds_u32* b = lds + p * F * cap;
      b[i] = (uint32_t)v;
      if (rows) b[cap + i] = ri;
      if (tag) b[(F - 1) * cap + i] = du;
    } else {
      (p ? s->q1 : s->q0)[i] = v;
      if (rows) (p ? s->qr1 : s->qr0)[i] = ri;
      if (tag) (p ? s->qa1 : s->qa0)[i] = du;
    }
  }
};

// A low-degree graph (out-degree <= K) on one workgroup: K threads per
// frontier entry, one edge each, and entries that carry what the consumer
// would otherwise fetch in sequence: the vertex, its row (row_info) and, for
// SSSP, the distance it was queued with. With the queues in LDS a pass over
// entries is two dependent HBM round trips -- edge (+ the vertex's current
// distance); atomic (+ the target's row) -- and each thread keeps U entries'
// round trips in flight together, so a level of up to U * bd / K entries is
// one pass and one barrier. An SSSP entry whose vertex has since dropped
// lower is skipped (its newer entry carries the lower distance). Near-far
// bands as in run_single; a far queue of at most bd entries is split in one
// pass that keeps each entry's distance and row in registers across the
// band minimum.
// ELL: the graph is read from packed rows instead of the CSR -- K words per
// vertex, word k = (target << 8) | weight of its k-th edge, kEllNone past its
// degree -- so the K threads of an entry load one 4K-byte row together, an
// entry needs no row_info, and the second round trip is the atomic alone.
// COMB (SSSP on packed rows with the LDS bitmap): no atomic on dist at all.
// A pass loads each candidate target's distance with a plain load and folds
// its candidates per target in an LDS combine table (comb_insert); after an
// LDS-only barrier the one winner of each slot, if below the loaded
// distance, stores it and queues the target. The workgroup is the only
// writer of the graph's distances, so its L2 holds them and the second round
// trip is an L2 hit instead of an atomic beyond the L2. A level of several
// passes orders one pass's stores before the next pass's loads with a full
// barrier.
template <bool SSSP, int K, int U, bool ELL, bool COMB = false>
__device__ void run_low(const GraphDev& g, const FrontierDev& s, const NearFar& nf, int src, int* lds, int* s_cnt,
                        int* iters_out, bool lds_flags, const uint32_t* ell) {
  const bool near_far = SSSP && nf.delta > 0;
  __shared__ int s_far[3];  // far counts of q[0], q[1]; band minimum
  const LowQueues lq = COMB ? LowQueues{(lds_u32*)lds, kCombCap, &s, 2}
                            : LowQueues{(lds_u32*)(reinterpret_cast<uint32_t*>(lds) + frontier_lds_bytes((int)blockDim.x) / 4),
                                        low_lds_cap((int)blockDim.x), &s, 3};
  // one bit per vertex in LDS (lds_flags): SSSP's far flags (else a word per
  // vertex in HBM); BFS's visited marks, which then replace the atomic on
  // dist (else the atomic minimum decides the discovery)
  lds_u32* fbits = COMB ? comb_bits(lds) : batch_bits(lds);
  const bool bits = near_far && lds_flags;
  const bool vis = !SSSP && lds_flags;
  for (int v = threadIdx.x; v < g.n; v += blockDim.x) {
    s.dist[v] = (v == src) ? 0u : kInf;
    if (near_far && !bits) nf.flag[v] = 0u;
  }
  if (bits || vis)
    for (int w = threadIdx.x; w < (g.n + 31) / 32; w += blockDim.x)
      fbits[w] = vis && w == (src >> 5) ? 1u << (src & 31) : 0u;
  lds_u64* comb = batch_comb(lds);
  if (COMB)
    for (int i = threadIdx.x; i < 3 * kCombSlots; i += blockDim.x) comb[i] = kCombEmpty;
  if (threadIdx.x == 0) {
    lq.store(0, 0, src, ELL ? 0u : row_info(g, src), 0u, SSSP, !ELL);
    s_cnt[0] = 1;
    s_cnt[1] = 0;
    s_cnt[2] = 0;
    s_far[0] = 0;
    s_far[1] = 0;
    s_far[2] = INT32_MAX;
  }
  __syncthreads();
  const int bd = blockDim.x, per = bd / K;
  const int slot = (int)threadIdx.x / K, k = (int)threadIdx.x % K;
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  int level = 0, fi = 0;
  uint32_t thr = near_far ? nf.delta : 0xffffffffu;
  uint32_t nrel = 0;  // edges this thread relaxed (live entries' edges)
  int tcur = 0;       // COMB: this pass's table
  // COMB: a queued target's packed row is loaded as it is queued, so that the
  // next level finds it in the cache; the loads complete by the level's
  // barrier and their values are only folded into pfacc (never read)
  uint32_t pfv[U], pfacc = 0;
#pragma unroll
  for (int x = 0; x < U; x++) pfv[x] = 0u;
  int clr[U];         // COMB: slot this thread won in the previous pass (cleared after this pass's barrier)
#pragma unroll
  for (int x = 0; x < U; x++) clr[x] = -1;
  while (true) {
    for (;; level++) {
      const int in_n = __builtin_amdgcn_readfirstlane(s_cnt[level % 3]);
      if (in_n == 0) break;
      // counter (level + 2) % 3 was last read at the start of level - 1
      if (threadIdx.x == 0) s_cnt[(level + 2) % 3] = 0;
      const int p = level & 1;  // in queue; 1 - p is the out queue
      int* out_cnt = &s_cnt[(level + 1) % 3];
      const uint32_t next = (uint32_t)level + 1u;
      for (int a = 0; a < in_n; a += per * U) {
        // slot x holds entries of this wave only while a + x per + its first
        // slot < in_n: a scalar test, so a wave runs no code for empty slots
        // (most levels fill one slot, and only some waves of it) -- issue
        // time, not memory, bounds a narrow level. COMB passes have barriers:
        // every wave runs them, and the per-slot tests below skip the rest.
        const int wslot = __builtin_amdgcn_readfirstlane((int)threadIdx.x & ~63) / K;
        if (COMB && a > 0) {
          __syncthreads();  // the last pass's stores before this pass's loads
#pragma unroll
          for (int x = 0; x < U; x++) {
            pfacc ^= pfv[x];
            pfv[x] = 0u;
          }
        }
        if (!COMB && a + wslot >= in_n) continue;
        bool lv[U];
#pragma unroll
        for (int x = 0; x < U; x++) lv[x] = a + x * per + wslot < in_n;
        int v[U], e[U];
        uint32_t du[U];
        bool ok[U];
#pragma unroll
        for (int x = 0; x < U; x++) {
          const int i = a + x * per + slot;
          uint32_t ri = 0;
          v[x] = 0;
          du[x] = 0;
          if (lv[x] && i < in_n) lq.load(p, i, v[x], ri, du[x], SSSP, !ELL);
          ok[x] = ELL ? (lv[x] && i < in_n) : (int)(ri & 7u) > k;
          e[x] = ELL ? v[x] * K + k : (int)(ri >> 3) + k;
        }
        // round trip 1, every slot in flight: the edge and the entry's distance
        int u[U];
        uint32_t w[U], dv[U];
#pragma unroll
        for (int x = 0; x < U; x++) {
          u[x] = 0;
          w[x] = 0u;
          dv[x] = 0u;
          if (!lv[x]) continue;
          if (ELL) {
            const uint32_t word = ok[x] ? ell[e[x]] : kEllNone;
            dv[x] = SSSP && ok[x] ? s.dist[v[x]] : 0u;
            ok[x] = word != kEllNone;
            u[x] = ok[x] ? (int)(word >> 8) : 0;
            w[x] = word & 255u;
          } else {
            u[x] = ok[x] ? g.col[e[x]] : 0;
            w[x] = SSSP && ok[x] ? (uint32_t)g.wt[e[x]] : 0u;
            dv[x] = SSSP && ok[x] ? s.dist[v[x]] : 0u;
          }
        }
        // round trip 2, every slot in flight: the atomic and the target's row.
        // Only the memory operations sit under each slot's condition; what
        // reads their results comes after every slot has issued.
        uint32_t alt[U], old[U], r0[U], r1[U];
        int hs[U];  // COMB: the target's slot in this pass's table
#pragma unroll
        for (int x = 0; x < U; x++) {
          hs[x] = -1;
          const bool live = ok[x] && (!SSSP || dv[x] == du[x]);
          nrel += live ? 1u : 0u;
          alt[x] = SSSP ? du[x] + w[x] : next;
          // the row loads are unconditional (a dead slot reads row 0, which
          // every graph has): only the atomic needs the slot's condition
          r0[x] = 0u;
          r1[x] = 0u;
          old[x] = 0u;  // not live: never an improvement, never kInf
          if (!lv[x]) continue;
          if (!ELL) {
            int uu = live ? u[x] : 0;
            opaque(uu);  // one unconditional load pair, not one per arm of live
            r0[x] = (uint32_t)g.row_ptr[uu];
            r1[x] = (uint32_t)g.row_ptr[uu + 1];
          }
          if (live) {
            if (COMB) {
              // the target's distance (plain load, in flight across the
              // barrier) and the candidate into the combine table
              old[x] = s.dist[u[x]];
              hs[x] = comb_insert(comb + tcur * kCombSlots, u[x], comb_word(u[x], alt[x], x * bd + (int)threadIdx.x));
            } else if (vis) {
              // BFS with visited bits: the first to set u's bit discovers it
              const uint32_t bit = 1u << (u[x] & 31);
              old[x] = (__hip_atomic_fetch_or(&fbits[u[x] >> 5], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) &
                        bit) ? 0u : kInf;
            } else {
              // BFS: every distance set so far is <= level + 1 = alt, so the
              // minimum changes only kInf and old == kInf marks the discovery
              // (a compare-and-swap without its compare inside the block)
              old[x] = __hip_atomic_fetch_min(&s.dist[u[x]], alt[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
          }
        }
        bool won[U];
#pragma unroll
        for (int x = 0; x < U; x++) won[x] = false;
        if constexpr (COMB) {
          // every candidate is in the table; the slot's word is the minimum
          // over its target's candidates, and the candidate whose word it is
          // wins it (the thread bits break ties)
          lds_barrier();
          lds_u64* tb = comb + tcur * kCombSlots;
          lds_u64* tp = comb + (tcur == 0 ? 2 : tcur - 1) * kCombSlots;
#pragma unroll
          for (int x = 0; x < U; x++) {
            if (clr[x] >= 0) tp[clr[x]] = kCombEmpty;
            clr[x] = -1;
            if (hs[x] >= 0) {
              won[x] = tb[hs[x]] == comb_word(u[x], alt[x], x * bd + (int)threadIdx.x);
              if (won[x]) clr[x] = hs[x];
            }
          }
          tcur = tcur == 2 ? 0 : tcur + 1;
        }
#pragma unroll
        for (int x = 0; x < U; x++) {
          opaque(old[x]);
          opaque(r0[x]);
          opaque(r1[x]);
        }
        bool want[U], far_cand[U];
        uint32_t ru[U];
#pragma unroll
        for (int x = 0; x < U; x++) {
          ru[x] = (r0[x] << 3) | (r1[x] - r0[x]);
          const bool better = COMB ? won[x] && alt[x] < old[x] : SSSP ? alt[x] < old[x] : old[x] == kInf;
          want[x] = better && (!SSSP || alt[x] < thr);
          far_cand[x] = SSSP && better && alt[x] >= thr;
          // the discoverer writes the level: nothing reads dist during a BFS
          if (vis && want[x]) s.dist[u[x]] = alt[x];
          if (COMB && better) s.dist[u[x]] = alt[x];
          if (COMB && want[x]) pfv[x] = ell[(size_t)u[x] * K];
        }
#pragma unroll
        for (int x = 0; x < U; x++) {
          if (!lv[x]) continue;
          // the far flag's exchange is issued first, so that its round trip
          // overlaps the near queue's counter atomic below
          bool far_want = false;
          if (SSSP) {
            if (bits) {
              const uint32_t bit = 1u << (u[x] & 31);
              far_want = far_cand[x] && (__hip_atomic_fetch_or(&fbits[u[x] >> 5], bit, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_WORKGROUP) & bit) == 0u;
            } else {
              far_want = far_cand[x] && __hip_atomic_exchange(&nf.flag[u[x]], 1u, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_WORKGROUP) == 0u;
            }
          }
          const unsigned long long mask = __ballot(want[x]);
          if (mask) {
            const int leader = __ffsll((long long)mask) - 1;
            int base = 0;
            if (lane == leader)
              base = __hip_atomic_fetch_add(out_cnt, __popcll(mask), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            base = __builtin_amdgcn_readlane(base, leader);  // leader is wave-uniform: no LDS round trip
            if (want[x]) lq.store(1 - p, base + __popcll(mask & below), u[x], ru[x], alt[x], SSSP, !ELL);
          }
          if (SSSP) push_item<false>(far_want, u[x], nf.far(fi), &s_far[fi], __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
      __syncthreads();
#pragma unroll
      for (int x = 0; x < U; x++) {
        pfacc ^= pfv[x];
        pfv[x] = 0u;
      }
    }
    if (!near_far) break;
    // uniform: read after the level's barrier (scalar, so the branches
    // around the barriers below stay uniform for the compiler too)
    const int fn = __builtin_amdgcn_readfirstlane(s_far[fi]);
    if (fn == 0) break;
    const int p = level & 1;  // this level's (empty) queue
    // move a far entry below the new threshold to the near queue, else to the other far queue
    const uint32_t t_old = thr;  // far entries below it were relaxed in the band (dropped)
    auto split = [&](bool in, int v, uint32_t dv, uint32_t rv) {
      bool to_near = false, to_far = false;
      if (in) {
        if (dv < thr) {
          // each deferred vertex is in one far queue once: it moves once
          if (bits) __hip_atomic_fetch_and(&fbits[v >> 5], ~(1u << (v & 31)), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
          else nf.flag[v] = 0u;
          to_near = dv >= t_old;
        } else {
          to_far = true;  // stays flagged in the other far queue
        }
      }
      const unsigned long long mask = __ballot(to_near);
      if (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        int nb = 0;
        if (lane == leader)
          nb = __hip_atomic_fetch_add(&s_cnt[level % 3], __popcll(mask), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
        nb = __builtin_amdgcn_readlane(nb, leader);
        if (to_near) lq.store(p, nb + __popcll(mask & below), v, rv, dv, true, !ELL);
      }
      push_item<false>(to_far, v, nf.far(1 - fi), &s_far[1 - fi], __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    if (fn <= bd) {
      const bool in = (int)threadIdx.x < fn;
      int v = 0;
      uint32_t dv = 0, rv = 0;
      if (in) {
        v = nf.far(fi)[threadIdx.x];
        dv = s.dist[v];
        rv = ELL ? 0u : row_info(g, v);
      }
      const int mn = wave_min_int(in && dv >= t_old ? (int)dv : INT32_MAX);
      if (lane == 0) atomicMin(&s_far[2], mn);
      __syncthreads();
      thr = (uint32_t)__builtin_amdgcn_readfirstlane(s_far[2]) + nf.delta;
      split(in, v, dv, rv);
    } else {
      int mn = INT32_MAX;
      for (int i = threadIdx.x; i < fn; i += bd) {
        const uint32_t d = s.dist[nf.far(fi)[i]];
        if (d >= t_old) mn = min(mn, (int)d);
      }
      mn = wave_min_int(mn);
      if (lane == 0) atomicMin(&s_far[2], mn);
      __syncthreads();
      thr = (uint32_t)__builtin_amdgcn_readfirstlane(s_far[2]) + nf.delta;
      for (int base = 0; base < fn; base += bd) {
        const int i = base + threadIdx.x;
        const bool in = i < fn;
        int v = 0;
        uint32_t dv = 0, rv = 0;
        if (in) {
          v = nf.far(fi)[i];
          dv = s.dist[v];
          rv = ELL ? 0u : row_info(g, v);
        }
        split(in, v, dv, rv);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      s_far[fi] = 0;
      s_far[2] = INT32_MAX;
    }
    fi = 1 - fi;
    __syncthreads();
  }
  asm volatile("" ::"v"(pfacc));  // keeps the row loads (their values are unused)
  for (int d = 32; d > 0; d >>= 1) nrel += __shfl_down(nrel, d, 64);
  if ((threadIdx.x & 63) == 0 && nrel && s.relaxed) atomicAdd(s.relaxed, (unsigned long long)nrel);
  if (threadIdx.x == 0) *iters_out = level;
}


// PER_GRAPH: the same body instantiated for the one-launch-per-graph
// baseline (a distinct symbol, so profiles separate the two modes).
template <int PER_GRAPH>
__global__ void __launch_bounds__(1024) k_batched(Ctl c, BatchDev b) {
  extern __shared__ int lds[];
  __shared__ int s_job[2];
  __shared__ int s_cnt[3];
  int n;
  const int id = discovery_protocol(c, &n);
  if (id < 0 || c.im_exit) return;
  // A team larger than the participants could never form: report and leave.
  if (n < b.max_team) {
    if (id == 0 && threadIdx.x == 0) st_relaxed(&c.lines[kLine], n);
    return;
  }
  clock_start(c, id);
  while (true) {
    if (threadIdx.x == 0) {
      const int t = add_relaxed(b.ticket, 1);
      int j = -1, r = 0;
      if (t < b.total_tickets) {
        int lo = 0, hi = b.njobs - 1;
        while (lo < hi) {  // last job with ticket <= t
          const int mid = (lo + hi + 1) >> 1;
          if (b.jobs[mid].ticket <= t) lo = mid; else hi = mid - 1;
        }
        j = lo;
        r = t - b.jobs[j].ticket;
        __hip_atomic_fetch_min(&b.jt[2 * j], (unsigned long long)wall_clock64(), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      }
      s_job[0] = j;
      s_job[1] = r;
    }
    __syncthreads();
    const int j = s_job[0], r = s_job[1];
    __syncthreads();
    if (j < 0) break;
    const BatchJob jb = b.jobs[j];
    GraphDev g{};
    g.row_ptr = b.row + jb.row_off;
    g.col = b.col + jb.col_off;
    g.wt = b.wt ? b.wt + jb.col_off : nullptr;
    g.n = jb.n;
    g.m = jb.m;
    g.pw = jb.p_off >= 0 ? b.pw + jb.p_off : nullptr;
    FrontierDev s{};
    s.dist = b.dist + jb.v_off;
    s.stamp = nullptr;  // SSSP queues carry distance tags instead
    s.q0 = b.q + jb.q_off;
    s.q1 = b.q + jb.q_off + jb.qcap;
    s.qa0 = b.qa + jb.q_off;
    s.qa1 = b.qa + jb.q_off + jb.qcap;
    s.qr0 = b.qr + jb.q_off;
    s.qr1 = b.qr + jb.q_off + jb.qcap;
    s.counts = b.counts + (size_t)j * 3 * kLine;
    s.claims = b.claims + (size_t)j * 3 * kLine;
    s.chunk = b.chunk;
    s.relaxed = b.relaxed + j;
    const bool sssp = jb.query == WGM_QUERY_SSSP;
    NearFar nf{};
    nf.delta = sssp && b.far ? jb.delta : 0u;
    if (nf.delta) {
      nf.q[0] = b.far + jb.f_off;
      nf.q[1] = b.far + jb.f_off + jb.n + 1;
      nf.flag = b.in_far + jb.v_off;
      nf.cnt = b.fcnt + (size_t)j * 3 * kLine;
    }
    if (jb.team <= 1) {
      // per-vertex bits in LDS when the bitmap holds the graph
      const bool lds_flags = b.far_bits && batch_far_bits_bytes((int)blockDim.x) >= 4 * ((jb.n + 31) / 32);
      // degree-bounded step for graphs whose out-degree never exceeds 4
      // (row starts must fit row_info's 29 bits)
      if (jb.maxdeg <= kLowDegree && jb.m < (1 << 29)) {
        if (b.low_ell && jb.e_off >= 0) {
          const uint32_t* ell = b.ell + jb.e_off;
          // the combine tables hold a pass's candidates at most half full
          // and candidate ids fit 11 bits
          if (sssp && b.low_ell == 2 && lds_flags && nf.delta && (int)blockDim.x * kCombU <= kCombSlots / 2)
            run_low<true, kLowDegree, kCombU, true, true>(g, s, nf, jb.src, lds, s_cnt, &b.iters[j], true, ell);
          else if (sssp) run_low<true, kLowDegree, kLowSlots, true>(g, s, nf, jb.src, lds, s_cnt, &b.iters[j], lds_flags, ell);
          else run_low<false, kLowDegree, kLowSlots, true>(g, s, nf, jb.src, lds, s_cnt, &b.iters[j], lds_flags, ell);
        } else {
          if (sssp)
            run_low<true, kLowDegree, kLowSlots, false>(g, s, nf, jb.src, lds, s_cnt, &b.iters[j], lds_flags, nullptr);
          else run_low<false, kLowDegree, kLowSlots, false>(g, s, nf, jb.src, lds, s_cnt, &b.iters[j], lds_flags, nullptr);
        }
      } else {
        if (sssp) run_single<true>(g, s, nf, jb.src, lds, s_cnt, &b.iters[j], lds_flags);
        else run_single<false>(g, s, nf, jb.src, lds, s_cnt, &b.iters[j], lds_flags);
      }
    } else {
      int* tcnt = b.tbar + (size_t)j * 2 * kLine;
      int* tgen = tcnt + kLine;
      int epoch = 0;
      const int team = jb.team;
      auto bar = [&]() { team_barrier(tcnt, tgen, team, epoch); };
      if (sssp) run_team<true>(g, s, nf, jb.src, r, jb.team, lds, &b.iters[j], bar);
      else run_team<false>(g, s, nf, jb.src, r, jb.team, lds, &b.iters[j], bar);
    }
    // The job's reached-edge sum (out-degrees of its reached vertices), by
    // the workgroups that ran it, each over its share of the vertices. A
    // team left its traversal after a barrier that drained every member's
    // atomics, and reads dist coherently (its members' L2s may hold stale
    // lines); one workgroup reads its own writes.
    {
      int lo = 0, hi = jb.n;
      if (jb.team > 1) block_range(jb.n, r, jb.team, &lo, &hi);
      unsigned long long local = 0;
      // kReachPer vertices per thread per round, all loads issued before any
      // use: one memory round trip per kReachPer * blockDim.x vertices (the
      // largest grid's workgroup sweeps 257k vertices at the end of the step)
      for (int base = lo; base < hi; base += kReachPer * (int)blockDim.x) {
        uint32_t d[kReachPer];
        int r0[kReachPer], r1[kReachPer];
#pragma unroll
        for (int q = 0; q < kReachPer; q++) {
          const int v = base + q * (int)blockDim.x + (int)threadIdx.x;
          const bool in = v < hi;
          d[q] = kInf;
          if (in) d[q] = jb.team > 1 ? __hip_atomic_load(&s.dist[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                     : s.dist[v];
          r0[q] = in ? g.row_ptr[v] : 0;
          r1[q] = in ? g.row_ptr[v + 1] : 0;
        }
#pragma unroll
        for (int q = 0; q < kReachPer; q++)
          if (d[q] != kInf) local += (unsigned long long)(r1[q] - r0[q]);
      }
      for (int sh = 32; sh > 0; sh >>= 1) local += __shfl_down(local, sh, 64);
      if ((threadIdx.x & 63) == 0 && local) atomicAdd(&b.reached[j], local);
    }
    // Thread 0 stamps the end before the loop's closing barrier: a thread-0
    // section after it would sit next to the ticket claim across the back
    // edge, and the two merged branches split the job loop into a divergent
    // inner loop around the barriers (lanes of wave 0 then meet them apart).
    if (threadIdx.x == 0)
      __hip_atomic_fetch_max(&b.jt[2 * j + 1], (unsigned long long)wall_clock64(), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
  }
  // the last workgroup to leave stamps the end of the batch
  if (threadIdx.x == 0) {
    fence_release();
    if (add_acq_rel(&c.lines[0], 1) == n - 1) c.clock[1] = wall_clock64();
  }
}

// Per-launch reset of the work queue and team barriers (capture-safe: a
// kernel, not a memset).
__global__ void k_batch_reset(BatchDev b) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int i = t; i < b.njobs * 2; i += gridDim.x * blockDim.x) {
    b.tbar[i * kLine] = 0;
    b.jt[i] = (i & 1) ? 0ull : ~0ull;
    if (i < b.njobs) {
      b.reached[i] = 0ull;
      b.relaxed[i] = 0ull;
    }
  }
  if (t == 0) *b.ticket = 0;
}

// Sum of out-degrees of reached vertices per job, after the traversal:
// job0 + blockIdx.y is the job, blockIdx.x a chunk of kReachChunk of its
// vertices (the blocks past a small job's end leave at once). Each thread
// issues all kReachPer of its loads before summing (one round of memory
// latency per block), then one atomic per block.
constexpr int kReachChunk = 256 * kReachPer;
constexpr int kMaxGridY = 65535;
__global__ void __launch_bounds__(256) k_batch_reached(BatchDev b, int job0) {
  const int j = job0 + (int)blockIdx.y;
  const int n = b.jobs[j].n;
  const int v0 = (int)blockIdx.x * kReachChunk;
  if (v0 >= n) return;
  const uint32_t* dist = b.dist + b.jobs[j].v_off;
  const int32_t* row = b.row + b.jobs[j].row_off;
  uint32_t d[kReachPer];
  int r0[kReachPer], r1[kReachPer];
#pragma unroll
  for (int k = 0; k < kReachPer; k++) {
    const int v = v0 + k * 256 + (int)threadIdx.x;
    const bool in = v < n;
    d[k] = in ? dist[v] : kInf;
    r0[k] = in ? row[v] : 0;
    r1[k] = in ? row[v + 1] : 0;
  }
  unsigned long long local = 0;
#pragma unroll
  for (int k = 0; k < kReachPer; k++)
    if (d[k] != kInf) local += (unsigned long long)(r1[k] - r0[k]);
  for (int s = 32; s > 0; s >>= 1) local += __shfl_down(local, s, 64);
  __shared__ unsigned long long s_sum[4];
  if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
    if (t) atomicAdd(&b.reached[j], t);
  }
}

// The control cells the host reads after a run, next to the per-job results
// so that one copy brings everything back.
__global__ void k_batch_snap(Ctl c, BatchDev b) {
  if (threadIdx.x == 0) {
    b.snap[0] = c.clock[0];
    b.snap[1] = c.clock[1];
    b.snap[2] = (unsigned long long)(long long)*c.count;
    b.snap[3] = (unsigned long long)(long long)c.lines[kLine];
  }
}

}  // namespace wgm

#define fail wgm_fail
#define HIPCHK WGM_HIPCHK

struct wgm_batch {
  wgm_runtime* rt = nullptr;
  int count = 0;
  std::vector<int32_t> n, query, src, maxdeg;
  std::vector<int64_t> m, row_off, col_off, v_off, q_off, e_off, p_off;
  int64_t total_rows = 0, total_edges = 0, total_v = 0, total_q = 0, total_e = 0, total_p = 0;
  uint32_t* ell = nullptr;  // packed rows of the low-degree graphs (e_off, -1: none)
  uint32_t* pw = nullptr;   // packed edges of the other SSSP graphs (p_off, -1: none)
  bool weighted = false;
  int32_t *row = nullptr, *col = nullptr, *wt = nullptr, *q = nullptr, *counts = nullptr, *tbar = nullptr;
  int32_t* claims = nullptr;
  BatchJob* d_jobs_one = nullptr;  // relaunch baseline: the same jobs, tickets from 0
  int chunk = 0;
  int32_t *ticket = nullptr, *iters = nullptr;  // iters points into res
  unsigned long long* jt = nullptr;     // per job: start, end (wall clock); points into res
  std::vector<double> start_us, end_us;  // per caller index, after a run
  std::vector<double> dur_us;            // per caller index: last run's duration (empty before one)
  int short_team = 0;                    // discovered workgroups when a team could not form (0: none)
  bool jobs_by_time = false;             // job table ordered by measured durations
  uint32_t *dist = nullptr, *qa = nullptr, *qr = nullptr;
  std::vector<int32_t> qcap;      // entries per queue, per graph
  int32_t* far = nullptr;         // near-far queues, 2 (n + 1) per job (aligned)
  uint32_t* in_far = nullptr;     // near-far flags, per vertex (v_off)
  int32_t* fcnt = nullptr;        // 3 lines per job
  int64_t total_f = 0;
  std::vector<int64_t> f_off;
  BatchJob* d_jobs = nullptr;
  // per-job results and control snapshot in one block (one copy per run):
  // reached[count], jt[2 count], relaxed[count], snap[4], then iters (int32)[count]
  unsigned long long* res = nullptr;
  size_t res_bytes = 0;
  std::vector<BatchJob> jobs;           // current job table (LPT order)
  wgm_batch_config jobs_cfg{};          // config and grid the device job table was built for
  int jobs_grid = -1;
  std::vector<int64_t> reached;         // per caller index, after a run
  std::vector<int> levels;              // per caller index, after a run
  // captured launch
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  wgm_batch_config graph_cfg{};
  int graph_grid = 0;
  // standalone baseline: heavy lists for the largest graph and their 3 count lines
  int32_t* sa_heavy = nullptr;
  size_t sa_hcap = 0;
};

namespace {

// Team size: ceil(m / team_edges) (sssp_team_edges for SSSP queries), capped;
// see kLowDegreeSingle.
int team_size(const wgm_batch_config& cfg, int64_t m, int32_t maxdeg, bool sssp) {
  int64_t te = cfg.team_edges > 0 ? cfg.team_edges : (1 << 16);
  if (sssp && cfg.sssp_team_edges > 0) te = cfg.sssp_team_edges;
  int64_t k = maxdeg <= kLowDegree ? (m + kLowDegreeSingle - 1) / kLowDegreeSingle : (m + te - 1) / te;
  if (k < 1) k = 1;
  if (k > cfg.max_team) k = cfg.max_team;
  return (int)k;
}

BatchDev dev_view(wgm_batch* b) {
  BatchDev d{};
  d.jobs = b->d_jobs;
  d.njobs = b->count;
  d.total_tickets = 0;
  d.max_team = 1;
  for (const auto& j : b->jobs) {
    d.total_tickets += j.team;
    d.max_team = std::max(d.max_team, j.team);
  }
  d.row = b->row;
  d.col = b->col;
  d.wt = b->weighted ? b->wt : nullptr;
  d.dist = b->dist;
  d.qa = b->qa;
  d.qr = b->qr;
  d.q = b->q;
  d.counts = b->counts;
  d.claims = b->claims;
  d.chunk = b->chunk;
  d.tbar = b->tbar;
  d.ticket = b->ticket;
  d.iters = b->iters;
  d.jt = b->jt;
  d.reached = b->res;
  d.relaxed = b->res + 3 * (size_t)b->count;
  d.snap = b->res + 4 * (size_t)b->count;
  d.far = b->far;
  d.in_far = b->in_far;
  d.fcnt = b->fcnt;
  d.ell = b->ell;
  d.pw = b->pw;
  return d;
}

int normalize(wgm_runtime* rt, const wgm_batch_config* in, wgm_batch_config* out) {
  wgm_batch_config c{};
  if (in) c = *in;
  if (c.wg_size <= 0) c.wg_size = 256;
  if (c.wgs_per_cu <= 0) c.wgs_per_cu = 4;
  if (c.max_team <= 0) c.max_team = 16;
  if (c.team_edges <= 0) c.team_edges = 1 << 16;
  if (c.chunk <= 0) c.chunk = c.wg_size;
  if (c.delta < 0) c.delta = 0;
  if (c.sssp_team_edges < 0) c.sssp_team_edges = 0;
  if (c.delta_high_degree < 0) c.delta_high_degree = 0;
  int rc = wgm_check_wg(rt, c.wg_size);
  if (rc) return rc;
  if (c.wgs_per_cu > 32 || c.max_team > 1024) return fail(rt, WGM_ERR_INVALID, "bad batch config");
  *out = c;
  return WGM_OK;
}

// Job table for a config: LPT order, team sizes, tickets. The order is by
// each query's duration in the last run once one has been timed, by edge
// count before: a long, narrow traversal (a large grid's SSSP: ~1500 levels
// on one workgroup) outlasts a larger but shallow graph's team, and starts
// first only when measured.
int build_jobs(wgm_batch* b, const wgm_batch_config& cfg, int grid) {
  std::vector<int> order(b->count);
  std::iota(order.begin(), order.end(), 0);
  const bool timed = b->dur_us.size() == (size_t)b->count;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return timed ? b->dur_us[a] > b->dur_us[c] : b->m[a] > b->m[c]; });
  b->jobs_by_time = timed;
  b->jobs.assign(b->count, BatchJob{});
  int ticket = 0;
  for (int k = 0; k < b->count; k++) {
    const int i = order[k];
    BatchJob& j = b->jobs[k];
    j.row_off = b->row_off[i];
    j.col_off = b->col_off[i];
    j.v_off = b->v_off[i];
    j.q_off = b->q_off[i];
    j.f_off = b->f_off[i];
    j.n = b->n[i];
    j.m = (int)b->m[i];
    j.query = b->query[i];
    const bool sssp = b->query[i] == WGM_QUERY_SSSP;
    j.team = std::min(team_size(cfg, b->m[i], b->maxdeg[i], sssp), std::max(1, grid));
    j.src = b->src[i];
    j.ticket = ticket;
    j.index = i;
    j.maxdeg = b->maxdeg[i];
    j.qcap = b->qcap[i];
    j.delta = (uint32_t)(b->maxdeg[i] > kLowDegree && cfg.delta_high_degree > 0 ? cfg.delta_high_degree : cfg.delta);
    j.e_off = b->e_off[i];
    j.p_off = b->p_off[i];
    ticket += j.team;
  }
  b->chunk = cfg.chunk;
  b->jobs_cfg = cfg;
  b->jobs_grid = grid;
  WGM_HIPCHK(b->rt, hipMemcpy(b->d_jobs, b->jobs.data(), sizeof(BatchJob) * b->count, hipMemcpyHostToDevice));
  std::vector<BatchJob> one = b->jobs;
  for (auto& j : one) j.ticket = 0;
  WGM_HIPCHK(b->rt, hipMemcpy(b->d_jobs_one, one.data(), sizeof(BatchJob) * b->count, hipMemcpyHostToDevice));
  return WGM_OK;
}

int batch_grid(wgm_runtime* rt, const wgm_batch_config& cfg, int* grid) {
  int disc = 0;
  int rc = wgm_resolve_grid(rt, WGM_K_BATCHED, cfg.wg_size, 0, &disc);
  if (rc) return rc;
  int g = cfg.wgs_per_cu * rt->info.num_cus;
  *grid = std::min(g, disc);
  return WGM_OK;
}

// Enqueue one persistent batched launch (reset kernels + the kernel).
// After the traversal: the reached-degree sums and the control snapshot
// (part of the captured launch, so a run ends with one copy to the host).
int enqueue_collect(wgm_batch* b, hipStream_t st, bool reached) {
  wgm_runtime* rt = b->rt;
  const BatchDev d = dev_view(b);
  const int max_n = *std::max_element(b->n.begin(), b->n.end());
  for (int j0 = 0; reached && j0 < b->count; j0 += kMaxGridY)
    hipLaunchKernelGGL(k_batch_reached, dim3((max_n + kReachChunk - 1) / kReachChunk, std::min(kMaxGridY, b->count - j0)),
                       dim3(256), 0, st, d, j0);
  hipLaunchKernelGGL(k_batch_snap, dim3(1), dim3(64), 0, st, rt->ctl, d);
  WGM_HIPCHK(rt, hipGetLastError());
  return WGM_OK;
}

int enqueue_batched(wgm_batch* b, const wgm_batch_config& cfg, int grid, hipStream_t st) {
  wgm_runtime* rt = b->rt;
  BatchDev d = dev_view(b);
  int rc = wgm_enqueue_reset_ctl(rt, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_batch_reset, dim3(64), dim3(256), 0, st, d);
  d.far_bits = cfg.far_flags_lds ? 1 : 0;
  d.low_ell = cfg.low_ell == 2 ? 2 : cfg.low_ell ? 1 : 0;
  hipLaunchKernelGGL(k_batched<0>, dim3(grid), dim3(cfg.wg_size), batch_lds_bytes(cfg.wg_size), st, rt->ctl, d);
  return enqueue_collect(b, st, false);  // k_batched sums each job's reached edges itself
}

// Ends a run: waits for the stream (event ev1 after the work), copies the
// results block once and fills the per-query results and the stats.
int collect_results(wgm_batch* b, wgm_batch_stats* st) {
  wgm_runtime* rt = b->rt;
  WGM_HIPCHK(rt, hipEventRecord(rt->ev1, rt->stream));
  std::vector<unsigned char> raw(b->res_bytes);
  WGM_HIPCHK(rt, hipMemcpyAsync(raw.data(), b->res, b->res_bytes, hipMemcpyDeviceToHost, rt->stream));
  WGM_HIPCHK(rt, hipStreamSynchronize(rt->stream));
  float ms = 0.f;
  WGM_HIPCHK(rt, hipEventElapsedTime(&ms, rt->ev0, rt->ev1));
  const size_t nj = (size_t)b->count;
  const unsigned long long* per_job = reinterpret_cast<const unsigned long long*>(raw.data());
  const unsigned long long* jt = per_job + nj;
  const unsigned long long* relaxed = per_job + 3 * nj;
  const unsigned long long* snap = per_job + 4 * nj;
  const int32_t* iters = reinterpret_cast<const int32_t*>(per_job + 4 * nj + 4);
  if (st) {
    st->wall_ms = ms;
    const unsigned long long c0 = snap[0], c1 = snap[1];
    st->device_ms = (c1 >= c0 && c0) ? (double)(c1 - c0) * 1e3 / (double)rt->info.wall_clock_hz : 0.0;
    st->participating = (int32_t)(long long)snap[2];
  }
  b->short_team = (int)(long long)snap[3];
  unsigned long long t0 = ~0ull;
  for (size_t k = 0; k < nj; k++) t0 = std::min(t0, jt[2 * k]);
  const double us_per_tick = 1e6 / (double)rt->info.wall_clock_hz;
  b->start_us.assign(b->count, -1.0);
  b->end_us.assign(b->count, -1.0);
  bool timed = true;
  b->reached.assign(b->count, 0);
  int64_t total = 0, relax_total = 0;
  for (int k = 0; k < b->count; k++) {
    b->reached[b->jobs[k].index] = (int64_t)per_job[k];
    b->levels[b->jobs[k].index] = iters[k];
    if (jt[2 * k + 1] >= jt[2 * k] && jt[2 * k] != ~0ull) {
      b->start_us[b->jobs[k].index] = (double)(jt[2 * k] - t0) * us_per_tick;
      b->end_us[b->jobs[k].index] = (double)(jt[2 * k + 1] - t0) * us_per_tick;
    } else {
      timed = false;
    }
    total += (int64_t)per_job[k];
    relax_total += (int64_t)relaxed[k];
  }
  if (timed) {
    b->dur_us.resize(b->count);
    for (int i = 0; i < b->count; i++) b->dur_us[i] = b->end_us[i] - b->start_us[i];
  }
  if (st) {
    st->edges_scanned = total;
    st->edges_relaxed = relax_total;
  }
  return WGM_OK;
}

}  // namespace

static int batch_lds(int wg) { return batch_lds_bytes(wg); }

// Extension hooks used by runtime.hip
bool wgm_extra_kernel_info(int kernel, KernelInfo* out) {
  if (kernel == WGM_K_BATCHED) {
    *out = KernelInfo{(const void*)&k_batched<0>, batch_lds};
    return true;
  }
  return wgm_sched_kernel_info(kernel, out);
}

int wgm_discover_extra(wgm_runtime* rt, int kernel, int grid, int wg_size) {
  if (kernel == WGM_K_BATCHED) {
    hipLaunchKernelGGL(k_batched<0>, dim3(grid), dim3(wg_size), batch_lds_bytes(wg_size), rt->stream, rt->ctl,
                       BatchDev{});
    return WGM_OK;
  }
  return wgm_sched_probe(rt, kernel, grid, wg_size);
}

extern "C" {

int wgm_batch_create(wgm_runtime* rt, int32_t count, const wgm_csr* graphs, const int32_t* queries,
                     const int32_t* sources, wgm_batch** out) {
  if (!rt || !graphs || !queries || !out || count < 1) return fail(rt, WGM_ERR_INVALID, "bad batch arguments");
  wgm_batch* b = new wgm_batch();
  b->rt = rt;
  b->count = count;
  b->levels.assign(count, 0);
  auto bail = [&](int rc) { wgm_batch_destroy(b); return rc; };
  for (int i = 0; i < count; i++) {
    const wgm_csr& g = graphs[i];
    if (g.n < 1 || g.m < 0 || g.m > INT32_MAX || !g.row_ptr || (g.m && !g.col) || g.row_ptr[0] != 0 ||
        g.row_ptr[g.n] != g.m)
      return bail(fail(rt, WGM_ERR_INVALID, "graph %d: bad CSR", i));
    if (queries[i] != WGM_QUERY_BFS && queries[i] != WGM_QUERY_SSSP)
      return bail(fail(rt, WGM_ERR_INVALID, "graph %d: bad query", i));
    if (queries[i] == WGM_QUERY_SSSP && !g.wt) return bail(fail(rt, WGM_ERR_INVALID, "graph %d: SSSP needs weights", i));
    const int s = sources ? sources[i] : 0;
    if (s < 0 || s >= g.n) return bail(fail(rt, WGM_ERR_INVALID, "graph %d: bad source", i));
    if (g.wt) b->weighted = true;
    b->n.push_back(g.n);
    b->m.push_back(g.m);
    int32_t md = 0;
    for (int32_t v = 0; v < g.n; v++) md = std::max(md, g.row_ptr[v + 1] - g.row_ptr[v]);
    b->maxdeg.push_back(md);
    b->query.push_back(queries[i]);
    b->src.push_back(s);
    b->row_off.push_back(b->total_rows);
    b->col_off.push_back(b->total_edges);
    b->v_off.push_back(b->total_v);
    b->q_off.push_back(b->total_q);
    b->f_off.push_back(b->total_f);
    bool w8 = g.n <= kEllMaxN;  // vertex ids fit 24 bits and weights 8
    for (int64_t k = 0; w8 && g.wt && k < g.m; k++) w8 = g.wt[k] >= 0 && g.wt[k] <= 255;
    const bool rows = w8 && md <= kLowDegree;
    const bool edges = w8 && md > kLowDegree && queries[i] == WGM_QUERY_SSSP;
    b->e_off.push_back(rows ? b->total_e : -1);
    if (rows) b->total_e += ((int64_t)kLowDegree * g.n + 63) / 64 * 64;
    b->p_off.push_back(edges ? b->total_p : -1);
    if (edges) b->total_p += ((int64_t)g.m + 63) / 64 * 64;
    // SSSP queues carry distance tags and a level queues at most one entry
    // per relaxation; BFS claims each vertex once
    const int64_t cap = (queries[i] == WGM_QUERY_SSSP ? std::max<int64_t>(g.n, g.m) : g.n) + 1;
    if (cap > INT32_MAX / 2) return bail(fail(rt, WGM_ERR_INVALID, "graph %d: too large for the queue pools", i));
    b->qcap.push_back((int32_t)cap);
    // per-job state on its own 256-byte lines: with team members reading
    // through their XCD's L2, no line may be shared with another job
    b->total_rows += g.n + 1;
    b->total_edges += g.m;
    b->total_v += ((int64_t)g.n + 63) / 64 * 64;
    b->total_q += (2 * cap + 63) / 64 * 64;
    b->total_f += ((2 * ((int64_t)g.n + 1)) + 63) / 64 * 64;
  }
  auto alloc = [&](void** p, size_t bytes) { return hipMalloc(p, bytes ? bytes : 4); };
  hipError_t e = alloc((void**)&b->row, sizeof(int32_t) * b->total_rows);
  if (e == hipSuccess) e = alloc((void**)&b->col, sizeof(int32_t) * b->total_edges);
  if (e == hipSuccess && b->weighted) e = alloc((void**)&b->wt, sizeof(int32_t) * b->total_edges);
  if (e == hipSuccess) e = alloc((void**)&b->dist, sizeof(uint32_t) * b->total_v);
  if (e == hipSuccess) e = alloc((void**)&b->q, sizeof(int32_t) * b->total_q);
  if (e == hipSuccess) e = alloc((void**)&b->qa, sizeof(uint32_t) * b->total_q);
  if (e == hipSuccess) e = alloc((void**)&b->qr, sizeof(uint32_t) * b->total_q);
  if (e == hipSuccess) e = alloc((void**)&b->counts, sizeof(int32_t) * 3 * kLine * count);
  if (e == hipSuccess) e = alloc((void**)&b->claims, sizeof(int32_t) * 3 * kLine * count);
  if (e == hipSuccess) e = alloc((void**)&b->d_jobs_one, sizeof(BatchJob) * count);
  if (e == hipSuccess) e = alloc((void**)&b->tbar, sizeof(int32_t) * 2 * kLine * count);
  if (e == hipSuccess) e = alloc((void**)&b->ticket, sizeof(int32_t) * kLine);
  b->res_bytes = sizeof(unsigned long long) * (4 * (size_t)count + 4) + sizeof(int32_t) * (size_t)count;
  if (e == hipSuccess) e = alloc((void**)&b->res, b->res_bytes);
  if (e == hipSuccess) {
    b->jt = b->res + count;
    b->iters = reinterpret_cast<int32_t*>(b->res + 4 * (size_t)count + 4);
  }
  if (e == hipSuccess) e = alloc((void**)&b->d_jobs, sizeof(BatchJob) * count);
  if (e == hipSuccess) e = alloc((void**)&b->far, sizeof(int32_t) * b->total_f);
  if (e == hipSuccess) e = alloc((void**)&b->in_far, sizeof(uint32_t) * b->total_v);
  if (e == hipSuccess) e = alloc((void**)&b->fcnt, sizeof(int32_t) * 3 * kLine * count);
  if (e == hipSuccess) e = alloc((void**)&b->ell, sizeof(uint32_t) * b->total_e);
  if (e == hipSuccess) e = alloc((void**)&b->pw, sizeof(uint32_t) * b->total_p);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipMalloc(batch pools)"));
  // zeroed state: a queue slot read before it is written holds a vertex id of
  // its own job (0 at first), never garbage
  e = hipMemset(b->dist, 0, sizeof(uint32_t) * b->total_v);
  if (e == hipSuccess) e = hipMemset(b->q, 0, sizeof(int32_t) * b->total_q);
  if (e == hipSuccess) e = hipMemset(b->qa, 0, sizeof(uint32_t) * b->total_q);
  if (e == hipSuccess) e = hipMemset(b->qr, 0, sizeof(uint32_t) * b->total_q);
  if (e == hipSuccess) e = hipMemset(b->far, 0, sizeof(int32_t) * b->total_f);
  if (e == hipSuccess) e = hipMemset(b->in_far, 0, sizeof(uint32_t) * b->total_v);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "hipMemset(batch state)"));
  // one staging copy per pool
  std::vector<int32_t> stage;
  stage.reserve((size_t)std::max(b->total_rows, b->total_edges));
  for (int i = 0; i < count; i++) stage.insert(stage.end(), graphs[i].row_ptr, graphs[i].row_ptr + graphs[i].n + 1);
  e = hipMemcpy(b->row, stage.data(), sizeof(int32_t) * stage.size(), hipMemcpyHostToDevice);
  stage.clear();
  for (int i = 0; i < count && e == hipSuccess; i++) stage.insert(stage.end(), graphs[i].col, graphs[i].col + graphs[i].m);
  if (e == hipSuccess && !stage.empty()) e = hipMemcpy(b->col, stage.data(), sizeof(int32_t) * stage.size(), hipMemcpyHostToDevice);
  if (b->weighted) {
    stage.clear();
    for (int i = 0; i < count; i++) {
      if (graphs[i].wt) stage.insert(stage.end(), graphs[i].wt, graphs[i].wt + graphs[i].m);
      else stage.insert(stage.end(), (size_t)graphs[i].m, 1);
    }
    if (e == hipSuccess && !stage.empty()) e = hipMemcpy(b->wt, stage.data(), sizeof(int32_t) * stage.size(), hipMemcpyHostToDevice);
  }
  if (b->total_e > 0 && e == hipSuccess) {
    std::vector<uint32_t> rows((size_t)b->total_e, kEllNone);
    for (int i = 0; i < count; i++) {
      if (b->e_off[i] < 0) continue;
      const wgm_csr& g = graphs[i];
      uint32_t* r = rows.data() + b->e_off[i];
      for (int32_t v = 0; v < g.n; v++)
        for (int32_t k = g.row_ptr[v]; k < g.row_ptr[v + 1]; k++)
          r[(size_t)v * kLowDegree + (k - g.row_ptr[v])] = ((uint32_t)g.col[k] << 8) | (g.wt ? (uint32_t)g.wt[k] : 0u);
    }
    e = hipMemcpy(b->ell, rows.data(), sizeof(uint32_t) * rows.size(), hipMemcpyHostToDevice);
  }
  if (b->total_p > 0 && e == hipSuccess) {
    std::vector<uint32_t> words((size_t)b->total_p, 0u);
    for (int i = 0; i < count; i++) {
      if (b->p_off[i] < 0) continue;
      const wgm_csr& g = graphs[i];
      uint32_t* w = words.data() + b->p_off[i];
      for (int64_t k = 0; k < g.m; k++) w[k] = ((uint32_t)g.col[k] << 8) | (uint32_t)g.wt[k];
    }
    e = hipMemcpy(b->pw, words.data(), sizeof(uint32_t) * words.size(), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "batch upload"));
  *out = b;
  return WGM_OK;
}

int wgm_batch_destroy(wgm_batch* b) {
  if (!b) return WGM_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  void* ptrs[] = {b->row, b->col, b->wt, b->q, b->counts, b->tbar, b->ticket, b->res, b->dist, b->qa, b->qr,
                  b->d_jobs, b->claims, b->d_jobs_one, b->far, b->in_far, b->fcnt, b->sa_heavy, b->ell, b->pw};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete b;
  return WGM_OK;
}

int64_t wgm_batch_bytes(const wgm_batch* b) {
  if (!b) return 0;
  return 4 * (b->total_rows + b->total_edges * (b->weighted ? 2 : 1) + b->total_e + b->total_p);
}

int wgm_batch_run(wgm_runtime* rt, wgm_batch* b, const wgm_batch_config* cfg_in, wgm_batch_stats* st) {
  if (!rt || !b || b->rt != rt) return fail(rt, WGM_ERR_INVALID, "bad runtime/batch");
  wgm_batch_config cfg;
  int rc = normalize(rt, cfg_in, &cfg);
  if (rc) return rc;
  int grid = 0;
  rc = batch_grid(rt, cfg, &grid);
  if (rc) return rc;
  const bool same = b->exec && std::memcmp(&cfg, &b->graph_cfg, sizeof(cfg)) == 0 && grid == b->graph_grid;
  // The captured launch reads the device job table by pointer: make sure it
  // holds this config's table (another run may have rebuilt it meanwhile).
  // Rebuilt once more after the first timed run, in duration order (the
  // captured launch reads the table by pointer; tickets and team sizes, its
  // by-value arguments, depend only on the config).
  const bool table_ok = std::memcmp(&cfg, &b->jobs_cfg, sizeof(cfg)) == 0 && grid == b->jobs_grid &&
                        (b->jobs_by_time || b->dur_us.empty());
  if (!same || !cfg.use_graph || !table_ok) {
    rc = build_jobs(b, cfg, grid);
    if (rc) return rc;
  }
  rt->ctl.im_exit = 0;
  rt->ctl.handshake = 0;
  if (cfg.use_graph && !same) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    HIPCHK(rt, hipStreamBeginCapture(rt->stream, hipStreamCaptureModeThreadLocal));
    int rc2 = enqueue_batched(b, cfg, grid, rt->stream);
    hipError_t e = hipStreamEndCapture(rt->stream, &b->graph);
    if (rc2) return rc2;
    if (e != hipSuccess) return wgm_hip_fail(rt, e, "hipStreamEndCapture");
    HIPCHK(rt, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_cfg = cfg;
    b->graph_grid = grid;
  }
  HIPCHK(rt, hipEventRecord(rt->ev0, rt->stream));
  if (cfg.use_graph) {
    HIPCHK(rt, hipGraphLaunch(b->exec, rt->stream));
  } else {
    rc = enqueue_batched(b, cfg, grid, rt->stream);
    if (rc) return rc;
  }
  wgm_batch_stats local{};
  local.launches = 1;
  local.grid = grid;
  rc = collect_results(b, &local);
  if (rc) return rc;
  if (b->short_team)
    return fail(rt, WGM_ERR_TOO_MANY_GROUPS, "only %d workgroups discovered for teams of up to %d", b->short_team,
                dev_view(b).max_team);
  local.gteps = local.wall_ms > 0 ? (double)local.edges_scanned / (local.wall_ms * 1e6) : 0.0;
  if (st) *st = local;
  return WGM_OK;
}

// Baseline: the same graphs, one persistent launch per graph, back to back
// on the stream (each launch runs discovery and serves exactly one job with
// a team of the same size the batched launch would use).
int wgm_batch_run_relaunch(wgm_runtime* rt, wgm_batch* b, const wgm_batch_config* cfg_in, wgm_batch_stats* st) {
  if (!rt || !b || b->rt != rt) return fail(rt, WGM_ERR_INVALID, "bad runtime/batch");
  wgm_batch_config cfg;
  int rc = normalize(rt, cfg_in, &cfg);
  if (rc) return rc;
  int grid = 0;
  rc = batch_grid(rt, cfg, &grid);
  if (rc) return rc;
  rc = build_jobs(b, cfg, grid);
  if (rc) return rc;
  rt->ctl.im_exit = 0;
  rt->ctl.handshake = 0;
  BatchDev d = dev_view(b);
  HIPCHK(rt, hipEventRecord(rt->ev0, rt->stream));
  for (int k = 0; k < b->count; k++) {
    BatchDev one = d;
    one.far_bits = cfg.far_flags_lds ? 1 : 0;
    one.low_ell = cfg.low_ell == 2 ? 2 : cfg.low_ell ? 1 : 0;
    one.jobs = b->d_jobs_one + k;  // same job, tickets numbered from 0
    one.njobs = 1;
    one.total_tickets = b->jobs[k].team;
    one.max_team = b->jobs[k].team;
    one.counts = d.counts + (size_t)k * 3 * kLine;
    one.claims = d.claims + (size_t)k * 3 * kLine;
    one.tbar = d.tbar + (size_t)k * 2 * kLine;
    one.iters = d.iters + k;
    one.jt = d.jt + 2 * k;
    one.reached = d.reached + k;
    one.relaxed = d.relaxed + k;
    rc = wgm_enqueue_reset_ctl(rt, rt->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, rt->stream, one);
    hipLaunchKernelGGL(k_batched<1>, dim3(b->jobs[k].team), dim3(cfg.wg_size), batch_lds_bytes(cfg.wg_size),
                       rt->stream, rt->ctl, one);
  }
  HIPCHK(rt, hipGetLastError());
  rc = enqueue_collect(b, rt->stream, false);
  if (rc) return rc;
  wgm_batch_stats local{};
  local.launches = b->count;
  local.grid = grid;
  rc = collect_results(b, &local);
  if (rc) return rc;
  local.device_ms = 0.0;  // per-launch clocks: only the wall time spans the whole baseline
  local.participating = 0;
  local.gteps = local.wall_ms > 0 ? (double)local.edges_scanned / (local.wall_ms * 1e6) : 0.0;
  if (st) *st = local;
  return WGM_OK;
}

// Baseline: one launch per query of the standalone app kernel of wgm_bfs /
// wgm_sssp (every discovered workgroup as one team, its near-far bands and
// heavy-entry pass), back to back on the stream, on the batch's resident
// arrays, results checked like the batched launch's.
int wgm_batch_run_standalone(wgm_runtime* rt, wgm_batch* b, const wgm_launch* lc, wgm_batch_stats* st) {
  if (!rt || !b || b->rt != rt) return fail(rt, WGM_ERR_INVALID, "bad runtime/batch");
  const int wg = lc && lc->wg_size > 0 ? lc->wg_size : 1024;
  int grid_bfs = 0, grid_sssp = 0;
  int rc = wgm_resolve_grid(rt, WGM_K_BFS, wg, lc ? lc->num_wgs : 0, &grid_bfs);
  if (!rc && b->weighted) rc = wgm_resolve_grid(rt, WGM_K_SSSP, wg, lc ? lc->num_wgs : 0, &grid_sssp);
  if (rc) return rc;
  if (b->jobs.empty()) {  // a job table (queue and far-queue offsets) is needed, in any order
    wgm_batch_config cfg;
    int grid = 0;
    rc = normalize(rt, nullptr, &cfg);
    if (!rc) rc = batch_grid(rt, cfg, &grid);
    if (!rc) rc = build_jobs(b, cfg, grid);
    if (rc) return rc;
  }
  const size_t hcap = (size_t)(*std::max_element(b->m.begin(), b->m.end()) / kHeavyMin) + 2;
  if (b->sa_hcap < hcap) {
    if (b->sa_heavy) (void)hipFree(b->sa_heavy);
    b->sa_heavy = nullptr;
    b->sa_hcap = 0;
    HIPCHK(rt, hipMalloc(&b->sa_heavy, (3 * hcap + 3 * kLine) * sizeof(int32_t)));
    b->sa_hcap = hcap;
  }
  // a discovery probe above leaves im_exit set in the control block the
  // kernels receive by value
  rt->ctl.im_exit = 0;
  rt->ctl.handshake = 0;
  const BatchDev d = dev_view(b);
  hipLaunchKernelGGL(k_batch_reset, dim3(64), dim3(256), 0, rt->stream, d);
  HIPCHK(rt, hipEventRecord(rt->ev0, rt->stream));
  for (int k = 0; k < b->count; k++) {
    const BatchJob& j = b->jobs[k];
    const bool sssp = j.query == WGM_QUERY_SSSP;
    GraphDev g{};
    g.row_ptr = b->row + j.row_off;
    g.col = b->col + j.col_off;
    g.wt = sssp ? b->wt + j.col_off : nullptr;
    g.n = j.n;
    g.m = j.m;
    FrontierDev s{};
    s.dist = b->dist + j.v_off;
    s.q0 = b->q + j.q_off;
    s.q1 = s.q0 + j.qcap;
    s.qa0 = sssp ? b->qa + j.q_off : nullptr;
    s.qa1 = sssp ? s.qa0 + j.qcap : nullptr;
    s.counts = b->counts + (size_t)k * 3 * kLine;
    s.claims = b->claims + (size_t)k * 3 * kLine;
    s.hv = b->sa_heavy;
    s.hdu = (uint32_t*)(b->sa_heavy + b->sa_hcap);
    s.hclaim = b->sa_heavy + 2 * b->sa_hcap;
    s.hcnt = b->sa_heavy + 3 * b->sa_hcap;
    NearFar nf{};
    if (sssp) {
      nf.q[0] = b->far + j.f_off;
      nf.q[1] = nf.q[0] + j.n + 1;
      nf.flag = b->in_far + j.v_off;
      nf.cnt = b->fcnt + (size_t)k * 3 * kLine;
      nf.delta = wgm_standalone_delta(j.maxdeg);
    }
    rc = wgm_enqueue_reset_ctl(rt, rt->stream);
    if (!rc) rc = wgm_enqueue_frontier(rt, g, s, nf, sssp, j.src, sssp ? grid_sssp : grid_bfs, wg, d.iters + k);
    if (rc) return rc;
  }
  rc = enqueue_collect(b, rt->stream, true);
  if (rc) return rc;
  wgm_batch_stats local{};
  local.launches = b->count;
  local.grid = grid_bfs;
  rc = collect_results(b, &local);
  if (rc) return rc;
  local.device_ms = 0.0;  // per-launch clocks: only the wall time spans the whole baseline
  local.participating = 0;
  local.gteps = local.wall_ms > 0 ? (double)local.edges_scanned / (local.wall_ms * 1e6) : 0.0;
  if (st) *st = local;
  return WGM_OK;
}

int wgm_batch_job_times(const wgm_batch* b, double* start_us, double* end_us) {
  if (!b || !start_us || !end_us) return fail(b ? b->rt : nullptr, WGM_ERR_INVALID, "bad job-times arguments");
  if (b->start_us.empty()) return fail(b->rt, WGM_ERR_INVALID, "no run yet");
  std::copy(b->start_us.begin(), b->start_us.end(), start_us);
  std::copy(b->end_us.begin(), b->end_us.end(), end_us);
  return WGM_OK;
}

int wgm_batch_result(wgm_batch* b, int32_t index, uint32_t* dist_out, int32_t* iterations, int64_t* edges_scanned) {
  if (!b || index < 0 || index >= b->count) return fail(b ? b->rt : nullptr, WGM_ERR_INVALID, "bad batch index");
  wgm_runtime* rt = b->rt;
  if (dist_out)
    HIPCHK(rt, hipMemcpy(dist_out, b->dist + b->v_off[index], sizeof(uint32_t) * b->n[index], hipMemcpyDeviceToHost));
  if (iterations) {
    *iterations = b->reached.empty() ? -1 : b->levels[index];
  }
  if (edges_scanned) *edges_scanned = b->reached.empty() ? -1 : b->reached[index];
  return WGM_OK;
}

// Grid search over workgroup size, workgroups per CU and team granularity;
// keeps the fastest configuration (each candidate timed `reps` times, best kept).
int wgm_batch_autotune(wgm_runtime* rt, wgm_batch* b, int32_t reps, wgm_batch_config* best_out,
                       wgm_batch_stats* best_stats) {
  if (!rt || !b || !best_out) return fail(rt, WGM_ERR_INVALID, "bad autotune arguments");
  if (reps < 1) reps = 1;
  wgm_batch_config best{};
  wgm_batch_stats bst{};
  double best_ms = 1e30;
  auto trial = [&](const wgm_batch_config& c) {
    for (int r = 0; r < reps; r++) {
      wgm_batch_stats s{};
      const int rc = wgm_batch_run(rt, b, &c, &s);
      if (rc) return rc;
      if (s.wall_ms < best_ms) {
        best_ms = s.wall_ms;
        best = c;
        bst = s;
      }
    }
    return (int)WGM_OK;
  };
  // the other knobs start at values that suit a mixed batch: BFS teams of
  // 2^20 edges per workgroup, SSSP teams of 786432, bands of 512 (low
  // degree) and 12, the LDS bits and packed rows on
  wgm_batch_config start{};
  start.chunk = 256;
  start.team_edges = 1 << 20;
  start.max_team = 16;
  start.delta = 512;
  start.delta_high_degree = 12;
  start.sssp_team_edges = 786432;
  start.far_flags_lds = 1;
  start.low_ell = 2;
  // 1. launch shape: workgroup size x workgroups per CU
  const int wgs[] = {64, 128, 256, 512, 1024};
  const int per_cu[] = {1, 2, 4, 8};
  for (int wg : wgs)
    for (int pc : per_cu) {
      if (wg * pc > 2048) continue;  // CU thread capacity
      wgm_batch_config c = start;
      c.wg_size = wg;
      c.wgs_per_cu = pc;
      const int rc = trial(c);
      if (rc) return rc;
    }
  // 2. then one knob at a time at the best so far, twice over the knobs
  //    that interact (team granularities and bands)
  const int chunks[] = {32, 64, 256, 1024, 4096};
  for (int ch : chunks) {
    wgm_batch_config c = best;
    c.chunk = ch;
    const int rc = trial(c);
    if (rc) return rc;
  }
  for (int pass = 0; pass < 2; pass++) {
    const int team_edges[] = {1 << 14, 1 << 16, 1 << 18, 1 << 20, 1 << 22};
    for (int te : team_edges) {
      wgm_batch_config c = best;
      c.team_edges = te;
      const int rc = trial(c);
      if (rc) return rc;
    }
    const int sssp_te[] = {0, 1 << 16, 1 << 18, 1 << 19, 786432, 1 << 20};
    for (int te : sssp_te) {
      wgm_batch_config c = best;
      c.sssp_team_edges = te;
      const int rc = trial(c);
      if (rc) return rc;
    }
    const int deltas[] = {0, 256, 512, 1024, 2048, 4096};
    for (int dl : deltas) {
      wgm_batch_config c = best;
      c.delta = dl;
      const int rc = trial(c);
      if (rc) return rc;
    }
    const int deltas_hi[] = {0, 4, 8, 12, 16, 32, 128};
    for (int dl : deltas_hi) {
      wgm_batch_config c = best;
      c.delta_high_degree = dl;
      const int rc = trial(c);
      if (rc) return rc;
    }
  }
  // 3. the single-workgroup step's per-vertex LDS bits and packed rows
  for (int flags = 0; flags < 6; flags++) {
    wgm_batch_config c = best;
    c.far_flags_lds = flags & 1;
    c.low_ell = flags >> 1;
    const int rc = trial(c);
    if (rc) return rc;
  }
  *best_out = best;
  if (best_stats) *best_stats = bst;
  return WGM_OK;
}

}  // extern "C"
