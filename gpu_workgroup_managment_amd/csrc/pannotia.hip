// pannotia.hip -- the Pannotia SSSP and BC persistent apps.
//
// Pannotia SSSP (sssp_combined, src/pannotia/graph_apps/sssp/device/sssp.cl:11-92,
// plugin src/pannotia/graph_apps/sssp/sssp.h:40-176): min-plus sweeps over the
// in-edges, y[v] = min(x[v], min_j w_j + x[src_j]), until a sweep changes
// nothing. The reference runs Jacobi sweeps with three barriers each (copy,
// relax, compare). Here one sweep relaxes in place and flags any decrease, with
// one barrier per sweep: values only decrease and every one stays an upper
// bound, so the sweeps reach the same unique fixpoint (the shortest
// distances; BIG_NUM where unreached) in at most as many sweeps. The in-edge
// rows are cut into tasks of at most 16 * G edges, G lanes per task, so a hub's
// row is shared by several lane groups and meets at one atomic min.
//
// Pannotia BC (bc_combined, src/pannotia/graph_apps/bc/device/bc.cl:186-235,
// plugin src/pannotia/graph_apps/bc/bc.h): Brandes' betweenness. The reference
// runs one source at a time over all workgroups with global barriers per BFS
// level (two per level forward, one per level backward) and float atomics for
// the path counts and dependencies, which makes its sums order-dependent.
// Here each workgroup takes a whole source: BFS levels with a workgroup queue,
// path counts pulled over in-edges in in-edge order, dependencies pulled over
// out-edges in CSR order, all in FP64 with no atomics on the sums. The
// dependencies of one batch of sources (one per workgroup) land in per-slot
// rows that, after one barrier, every workgroup adds to bc in source order for
// its share of the vertices. The result does not depend on the grid and
// equals the FP64 oracle bit for bit.

#include "apps_internal.hpp"

#include <algorithm>
#include <vector>

using namespace wgm;

namespace wgm {

constexpr int32_t kBigNum = WGM_BIG_NUM;
// Cap on BC's per-slot state (28 bytes a vertex per slot): the grid shrinks to fit.
constexpr unsigned long long kBcSlotBytes = 8ull << 30;

struct MinplusDev {
  const int4* tasks;  // {row, begin, end, 0}
  int32_t ntasks;
  int32_t max_sweeps;  // n + 1 sweeps settle any graph without a negative cycle
  const int32_t* rcol;
  const int32_t* rwt;
  int32_t* dist;
};

template <int G>
__global__ void __launch_bounds__(1024) k_minplus(Ctl c, MinplusDev p, int* iters_out) {
  Team t;
  if (!standalone_begin(c, t)) return;
  const int lane = (int)threadIdx.x % G, grp = (int)threadIdx.x / G, ngrp = (int)blockDim.x / G;
  // tasks interleaved over every lane group of the grid: the consecutive
  // tasks of a hub's row spread over many workgroups
  const int first = t.id * ngrp + grp, stride = t.n * ngrp;
  int* flags = c.lines;
  int sweep = 1;
  for (;; sweep++) {
    bool any = false;
    for (int i = first; i < p.ntasks; i += stride) {
      const int4 tk = p.tasks[i];
      int best = INT32_MAX;
      // a plain (workgroup-scope) read of dist: a value another workgroup
      // lowers during this sweep may be seen late; the barrier makes it
      // current. Four of the lane's edges per round: their sources and
      // weights, then the four gathers, so the lane waits for two memory
      // round trips per four edges rather than per edge.
      int e = tk.y + lane;
      for (; e + 3 * G < tk.z; e += 4 * G) {
        int sv[4], wv[4], dv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          sv[q] = p.rcol[e + q * G];
          wv[q] = p.rwt[e + q * G];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) dv[q] = __hip_atomic_load(&p.dist[sv[q]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
        for (int q = 0; q < 4; q++) best = wv[q] + dv[q] < best ? wv[q] + dv[q] : best;
      }
      for (; e < tk.z; e += G) {
        const int du = __hip_atomic_load(&p.dist[p.rcol[e]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        const int cand = p.rwt[e] + du;
        best = cand < best ? cand : best;
      }
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) {
        const int other = __shfl_xor(best, o, G);
        best = other < best ? other : best;
      }
      if (lane == 0 && best < __hip_atomic_load(&p.dist[tk.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
        const int old = __hip_atomic_fetch_min(&p.dist[tk.x], best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        any |= best < old;
      }
    }
    if (__syncthreads_or(any) && threadIdx.x == 0) st_relaxed(&flags[(sweep % 3) * kLine], 1);
    if (t.id == 0 && threadIdx.x == 0) st_relaxed(&flags[((sweep + 1) % 3) * kLine], 0);
    xcd_barrier(c, t);
    if (ld_relaxed(&flags[(sweep % 3) * kLine]) == 0) break;
    if (sweep == p.max_sweeps) {  // still changing: a negative cycle; every workgroup stops here
      sweep = -1;
      break;
    }
  }
  clock_end(c, t.id);
  if (t.id == 0 && threadIdx.x == 0) *iters_out = sweep;
  standalone_retire(c);
}

// Unreached vertices start at `big` (BIG_NUM for Pannotia SSSP, MYINFINITY
// for Lonestar SSSP's pull direction): big + 255 stays below INT32_MAX, so a
// candidate through an unreached vertex never lowers anything.
__global__ void k_minplus_init(int32_t* dist, int n, int src, int32_t big) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    dist[v] = v == src ? 0 : big;
}

struct BcDev {
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* rrow;
  const int32_t* rcol;
  int32_t n;
  int32_t nsrc;
  int32_t slots;     // per-slot arrays allocated (>= participants)
  int32_t* d;        // [slots][n] BFS level, -1 unreached
  double* sig;       // [slots][n] shortest-path counts
  double* del;       // [slots][n] dependencies of the slot's source (zero between batches)
  int32_t* q;        // [slots][n] vertices in level order
  int32_t* lvl;      // [slots][n + 2] level starts in q
  double* bc;        // [n]
  unsigned long long* levels;  // total BFS levels over all sources
};

__device__ __forceinline__ int wg_ld(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One source on one workgroup; returns its number of BFS levels.
__device__ int bc_source(const BcDev& b, int s, int32_t* d, double* sig, double* del, int32_t* q, int32_t* lvl,
                         int* s_tail) {
  for (int v = (int)threadIdx.x; v < b.n; v += blockDim.x) d[v] = -1;
  __syncthreads();
  if (threadIdx.x == 0) {
    d[s] = 0;
    sig[s] = 1.0;
    q[0] = s;
    lvl[0] = 0;
    lvl[1] = 1;
    *s_tail = 1;
  }
  __syncthreads();
  int L = 0, lo = 0, hi = 1;
  while (lo < hi) {
    // discover level L + 1 (mega_bfs_kernel_func, bc.cl:17-104)
    for (int i = lo + (int)threadIdx.x; i < hi; i += blockDim.x) {
      const int v = q[i];
      for (int e = b.row_ptr[v]; e < b.row_ptr[v + 1]; e++) {
        const int w = b.col[e];
        if (wg_ld(&d[w]) == -1) {
          int expect = -1;
          if (__hip_atomic_compare_exchange_strong(&d[w], &expect, L + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP))
            q[atomicAdd(s_tail, 1)] = w;
        }
      }
    }
    __syncthreads();
    // the same for every lane: keep the level loop's control uniform
    const int tail = __builtin_amdgcn_readfirstlane(*s_tail);
    // path counts of level L + 1, pulled over in-edges in order
    for (int i = hi + (int)threadIdx.x; i < tail; i += blockDim.x) {
      const int w = q[i];
      double acc = 0.0;
      for (int j = b.rrow[w]; j < b.rrow[w + 1]; j++) {
        const int u = b.rcol[j];
        if (wg_ld(&d[u]) == L) acc += sig[u];
      }
      sig[w] = acc;
    }
    if (threadIdx.x == 0) lvl[L + 2] = tail;
    __syncthreads();
    lo = hi;
    hi = tail;
    L++;
  }
  // dependencies, deepest level first (mega_backtrack_kernel_func, bc.cl:106-183)
  for (int l = L - 1; l >= 0; l--) {
    const int a = lvl[l], z = lvl[l + 1];
    for (int i = a + (int)threadIdx.x; i < z; i += blockDim.x) {
      const int v = q[i];
      const double sv = sig[v];
      double acc = 0.0;
      for (int e = b.row_ptr[v]; e < b.row_ptr[v + 1]; e++) {
        const int w = b.col[e];
        if (wg_ld(&d[w]) == l + 1) acc += sv / sig[w] * (1.0 + del[w]);
      }
      del[v] = v == s ? 0.0 : acc;
    }
    __syncthreads();
  }
  return L;
}

__global__ void __launch_bounds__(1024) k_bc(Ctl c, BcDev b, int* iters_out) {
  __shared__ int s_tail;
  Team t;
  if (!standalone_begin(c, t)) return;
  const int S = t.n < b.slots ? t.n : b.slots;
  const size_t off = (size_t)t.id * (size_t)b.n;
  unsigned long long levels = 0;
  int batches = 0;
  for (int base = 0; base < b.nsrc; base += S, batches++) {
    const int s = base + t.id;
    if (t.id < S && s < b.nsrc)
      levels += (unsigned long long)bc_source(b, s, b.d + off, b.sig + off, b.del + off, b.q + off,
                                              b.lvl + (size_t)t.id * ((size_t)b.n + 2), &s_tail);
    xcd_barrier(c, t);
    // bc[v] += delta of each source of the batch, in source order; the rows
    // are cleared for the next batch on the way
    const int cnt = b.nsrc - base < S ? b.nsrc - base : S;
    for (int v = t.id * (int)blockDim.x + (int)threadIdx.x; v < b.n; v += t.n * (int)blockDim.x) {
      double acc = b.bc[v];
      for (int j = 0; j < cnt; j++) {
        double* row = b.del + (size_t)j * (size_t)b.n;
        acc += row[v];
        row[v] = 0.0;
      }
      b.bc[v] = acc;
    }
    xcd_barrier(c, t);
  }
  if (threadIdx.x == 0 && levels) atomicAdd(b.levels, levels);
  clock_end(c, t.id);
  if (t.id == 0 && threadIdx.x == 0) *iters_out = batches;
  standalone_retire(c);
}

}  // namespace wgm

#define fail wgm_fail
#define HIPCHK WGM_HIPCHK

namespace {

constexpr int kTaskEdgesPerLane = 16;

int minplus_lanes(const wgm_graph* g) {
  const double avg = g->n ? (double)g->m / (double)g->n : 0.0;
  return avg < 8.0 ? 4 : avg < 32.0 ? 16 : 64;
}

// In-edge CSR (sources ascending within a row, the order of
// parseCOO_transpose's stable sort) and the min-plus task list, built once on
// the host from the device copy of the graph and kept resident.
int graph_reverse(wgm_graph* g) {
  if (g->rrow) return WGM_OK;
  wgm_runtime* rt = g->rt;
  const int32_t n = g->n;
  const int64_t m = g->m;
  std::vector<int32_t> rp((size_t)n + 1), col((size_t)m), wt(g->wt ? (size_t)m : 0);
  HIPCHK(rt, hipMemcpy(rp.data(), g->row_ptr, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost));
  if (m) HIPCHK(rt, hipMemcpy(col.data(), g->col, sizeof(int32_t) * (size_t)m, hipMemcpyDeviceToHost));
  if (g->wt && m) HIPCHK(rt, hipMemcpy(wt.data(), g->wt, sizeof(int32_t) * (size_t)m, hipMemcpyDeviceToHost));
  std::vector<int32_t> rrow((size_t)n + 1, 0), rcol((size_t)m), rwt(g->wt ? (size_t)m : 0);
  for (int64_t e = 0; e < m; e++) rrow[(size_t)col[(size_t)e] + 1]++;
  for (int32_t v = 0; v < n; v++) rrow[(size_t)v + 1] += rrow[(size_t)v];
  std::vector<int32_t> pos(rrow.begin(), rrow.end() - 1);
  for (int32_t u = 0; u < n; u++)
    for (int32_t e = rp[(size_t)u]; e < rp[(size_t)u + 1]; e++) {
      const int32_t p = pos[(size_t)col[(size_t)e]]++;
      rcol[(size_t)p] = u;
      if (g->wt) rwt[(size_t)p] = wt[(size_t)e];
    }
  const int chunk = kTaskEdgesPerLane * minplus_lanes(g);
  std::vector<int4> tasks;
  tasks.reserve((size_t)n + (size_t)(m / chunk) + 1);
  for (int32_t v = 0; v < n; v++)
    for (int32_t a = rrow[(size_t)v]; a < rrow[(size_t)v + 1]; a += chunk)
      tasks.push_back(int4{v, a, std::min(a + chunk, rrow[(size_t)v + 1]), 0});
  auto alloc = [&](void** p, size_t bytes) -> int {
    hipError_t e = hipMalloc(p, bytes ? bytes : 4);
    return e == hipSuccess ? WGM_OK : wgm_hip_fail(rt, e, "hipMalloc(in-edge CSR)");
  };
  int rc = alloc((void**)&g->rrow, sizeof(int32_t) * ((size_t)n + 1));
  if (!rc) rc = alloc((void**)&g->rcol, sizeof(int32_t) * (size_t)m);
  if (!rc && g->wt) rc = alloc((void**)&g->rwt, sizeof(int32_t) * (size_t)m);
  if (!rc) rc = alloc((void**)&g->tasks, sizeof(int4) * tasks.size());
  if (rc) return rc;
  HIPCHK(rt, hipMemcpy(g->rrow, rrow.data(), sizeof(int32_t) * ((size_t)n + 1), hipMemcpyHostToDevice));
  if (m) HIPCHK(rt, hipMemcpy(g->rcol, rcol.data(), sizeof(int32_t) * (size_t)m, hipMemcpyHostToDevice));
  if (g->wt && m) HIPCHK(rt, hipMemcpy(g->rwt, rwt.data(), sizeof(int32_t) * (size_t)m, hipMemcpyHostToDevice));
  if (!tasks.empty())
    HIPCHK(rt, hipMemcpy(g->tasks, tasks.data(), sizeof(int4) * tasks.size(), hipMemcpyHostToDevice));
  g->ntasks = (int32_t)tasks.size();
  return WGM_OK;
}

const void* minplus_kernel(int lanes) {
  switch (lanes) {
    case 4: return (const void*)&k_minplus<4>;
    case 16: return (const void*)&k_minplus<16>;
    default: return (const void*)&k_minplus<64>;
  }
}

// Launch of a standalone persistent kernel whose first argument is the
// control block (*ctl, refreshed after the reset), as the apps launch theirs:
// grid from the occupancy of `kernel` (capped at max_grid), optional host
// handshake, bounded wait.
int launch_standalone_app(wgm_runtime* rt, int kernel, const void* fn, const wgm_launch* lc, Ctl* ctl, void** args,
                          int max_grid, wgm_run_stats* st) {
  wgm_launch def{256, 0, 0, 0};
  const wgm_launch& l = lc ? *lc : def;
  const int wg = l.wg_size > 0 ? l.wg_size : 256;
  int grid = 0;
  int rc = wgm_resolve_grid(rt, kernel, wg, l.num_wgs, &grid);
  if (rc) return rc;
  if (max_grid > 0 && grid > max_grid) grid = max_grid;
  rc = wgm_reset_ctl(rt, 0, l.handshake ? 1 : 0);
  if (rc) return rc;
  *ctl = rt->ctl;
  HIPCHK(rt, hipEventRecord(rt->ev0, rt->stream));
  HIPCHK(rt, hipLaunchKernel(fn, dim3(grid), dim3(wg), args, 0, rt->stream));
  double host_ms = 0.0;
  if (l.handshake) {
    rc = wgm_host_handshake(rt, &host_ms);
    if (rc) return rc;
  }
  rc = wgm_finish_launch(rt, st);
  if (rc) return rc;
  st->host_ms = host_ms;
  return WGM_OK;
}

// Grid of the BC launch: the occupancy grid, capped so the per-slot state
// stays within kBcSlotBytes.
int bc_grid(wgm_runtime* rt, const wgm_graph* g, const wgm_launch* lc, int* grid) {
  wgm_launch def{256, 0, 0, 0};
  const wgm_launch& l = lc ? *lc : def;
  int rc = wgm_resolve_grid(rt, WGM_K_BC, l.wg_size > 0 ? l.wg_size : 256, l.num_wgs, grid);
  if (rc) return rc;
  const size_t per_slot = (size_t)g->n * 28 + 8;
  const size_t cap = (size_t)kBcSlotBytes / per_slot;
  if (cap < 1) return fail(rt, WGM_ERR_NOMEM, "one BC slot needs %zu bytes", per_slot);
  if ((size_t)*grid > cap) *grid = (int)cap;
  return WGM_OK;
}

}  // namespace

int wgm_graph_in_edges(wgm_graph* g) { return graph_reverse(g); }

bool wgm_pannotia_kernel_info(int kernel, KernelInfo* out) {
  switch (kernel) {
    case WGM_K_MINPLUS: *out = KernelInfo{(const void*)&k_minplus<16>, nullptr}; return true;
    case WGM_K_BC: *out = KernelInfo{(const void*)&k_bc, nullptr}; return true;
    default: return false;
  }
}

int wgm_pannotia_probe(wgm_runtime* rt, int kernel, int grid, int wg_size) {
  switch (kernel) {
    case WGM_K_MINPLUS:
      hipLaunchKernelGGL(k_minplus<16>, dim3(grid), dim3(wg_size), 0, rt->stream, rt->ctl, MinplusDev{},
                         (int*)nullptr);
      return WGM_OK;
    case WGM_K_BC:
      hipLaunchKernelGGL(k_bc, dim3(grid), dim3(wg_size), 0, rt->stream, rt->ctl, BcDev{}, (int*)nullptr);
      return WGM_OK;
    default:
      return WGM_ERR_INVALID;
  }
}

void wgm_pannotia_release(wgm_graph* g) {
  void* ptrs[] = {g->rrow, g->rcol, g->rwt, g->tasks, g->bc_d, g->bc_sig, g->bc_del, g->bc_q, g->bc_lvl, g->bc_sum};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

extern "C" {

int wgm_sssp_minplus(wgm_runtime* rt, wgm_graph* g, int32_t src, const wgm_launch* lc, int32_t* dist_out,
                     wgm_run_stats* st) {
  if (!rt || !g || g->rt != rt) return fail(rt, WGM_ERR_INVALID, "bad runtime/graph");
  if (src < 0 || src >= g->n) return fail(rt, WGM_ERR_INVALID, "source %d out of range", src);
  if (!g->wt) return fail(rt, WGM_ERR_INVALID, "SSSP needs edge weights");
  int rc = graph_reverse(g);
  if (rc) return rc;
  if (!g->i32a) HIPCHK(rt, hipMalloc((void**)&g->i32a, sizeof(int32_t) * (size_t)g->n));
  hipLaunchKernelGGL(k_minplus_init, dim3(256), dim3(256), 0, rt->stream, g->i32a, g->n, src, kBigNum);
  MinplusDev p{(const int4*)g->tasks, g->ntasks, g->n + 1, g->rcol, g->rwt, g->i32a};
  Ctl ctl{};
  void* args[] = {&ctl, &p, &g->iters};
  // the three lane widths share one resource class; the grid is sized on G = 16
  wgm_run_stats local{};
  rc = launch_standalone_app(rt, WGM_K_MINPLUS, minplus_kernel(minplus_lanes(g)), lc, &ctl, args, 0, &local);
  if (rc) return rc;
  int iters = 0;
  HIPCHK(rt, hipMemcpy(&iters, g->iters, sizeof(int), hipMemcpyDeviceToHost));
  if (iters < 0) return fail(rt, WGM_ERR_INVALID, "distances still fall after %d sweeps: negative cycle", g->n + 1);
  local.iterations = iters;
  local.edges_scanned = (int64_t)iters * g->m;  // every sweep reads every in-edge
  if (dist_out) HIPCHK(rt, hipMemcpy(dist_out, g->i32a, sizeof(int32_t) * (size_t)g->n, hipMemcpyDeviceToHost));
  if (st) *st = local;
  return WGM_OK;
}

}  // extern "C"

// Lonestar SSSP in the pull direction: the min-plus sweeps above on the
// graph's own distance array with MYINFINITY for unreached vertices, so the
// result is wgm_sssp's uint32 distances.
int wgm_sssp_pull(wgm_runtime* rt, wgm_graph* g, int32_t src, const wgm_launch* lc, wgm_run_stats* st) {
  int rc = graph_reverse(g);
  if (rc) return rc;
  if (!g->dist) HIPCHK(rt, hipMalloc((void**)&g->dist, sizeof(uint32_t) * (size_t)g->n));
  hipLaunchKernelGGL(k_minplus_init, dim3(256), dim3(256), 0, rt->stream, (int32_t*)g->dist, g->n, src,
                     (int32_t)kInf);
  MinplusDev p{(const int4*)g->tasks, g->ntasks, g->n + 1, g->rcol, g->rwt, (int32_t*)g->dist};
  Ctl ctl{};
  void* args[] = {&ctl, &p, &g->iters};
  rc = launch_standalone_app(rt, WGM_K_MINPLUS, minplus_kernel(minplus_lanes(g)), lc, &ctl, args, 0, st);
  if (rc) return rc;
  int iters = 0;
  HIPCHK(rt, hipMemcpy(&iters, g->iters, sizeof(int), hipMemcpyDeviceToHost));
  if (iters < 0) return fail(rt, WGM_ERR_INVALID, "distances still fall after %d sweeps", g->n + 1);
  st->iterations = iters;
  return WGM_OK;
}

extern "C" {

int wgm_bc(wgm_runtime* rt, wgm_graph* g, int32_t num_sources, const wgm_launch* lc, double* bc_out,
           wgm_run_stats* st) {
  if (!rt || !g || g->rt != rt) return fail(rt, WGM_ERR_INVALID, "bad runtime/graph");
  if (num_sources > g->n) return fail(rt, WGM_ERR_INVALID, "%d sources for %d vertices", num_sources, g->n);
  const int32_t nsrc = num_sources > 0 ? num_sources : g->n;
  int rc = graph_reverse(g);
  if (rc) return rc;
  int grid = 0;
  rc = bc_grid(rt, g, lc, &grid);
  if (rc) return rc;
  if (g->bc_slots < grid) {
    void* old[] = {g->bc_d, g->bc_sig, g->bc_del, g->bc_q, g->bc_lvl};
    for (void* p : old)
      if (p) (void)hipFree(p);
    g->bc_d = g->bc_q = g->bc_lvl = nullptr;
    g->bc_sig = g->bc_del = nullptr;
    g->bc_slots = 0;
    const size_t sn = (size_t)grid * (size_t)g->n;
    hipError_t e = hipMalloc((void**)&g->bc_d, sizeof(int32_t) * sn);
    if (e == hipSuccess) e = hipMalloc((void**)&g->bc_sig, sizeof(double) * sn);
    if (e == hipSuccess) e = hipMalloc((void**)&g->bc_del, sizeof(double) * sn);
    if (e == hipSuccess) e = hipMalloc((void**)&g->bc_q, sizeof(int32_t) * sn);
    if (e == hipSuccess) e = hipMalloc((void**)&g->bc_lvl, sizeof(int32_t) * (size_t)grid * ((size_t)g->n + 2));
    if (e != hipSuccess) return wgm_hip_fail(rt, e, "hipMalloc(BC slots)");
    // dependency rows start at zero; each batch's reduction clears them again
    HIPCHK(rt, hipMemsetAsync(g->bc_del, 0, sizeof(double) * sn, rt->stream));
    g->bc_slots = grid;
  }
  if (!g->bc_sum) HIPCHK(rt, hipMalloc((void**)&g->bc_sum, sizeof(double) * ((size_t)g->n + 1)));
  unsigned long long* levels = (unsigned long long*)(g->bc_sum + g->n);
  HIPCHK(rt, hipMemsetAsync(g->bc_sum, 0, sizeof(double) * ((size_t)g->n + 1), rt->stream));
  BcDev b{g->row_ptr, g->col, g->rrow, g->rcol, g->n, nsrc, g->bc_slots, g->bc_d, g->bc_sig, g->bc_del,
          g->bc_q, g->bc_lvl, g->bc_sum, levels};
  Ctl ctl{};
  void* args[] = {&ctl, &b, &g->iters};
  wgm_run_stats local{};
  rc = launch_standalone_app(rt, WGM_K_BC, (const void*)&k_bc, lc, &ctl, args, grid, &local);
  if (rc) return rc;
  unsigned long long lv = 0;
  HIPCHK(rt, hipMemcpy(&lv, levels, sizeof(lv), hipMemcpyDeviceToHost));
  local.iterations = (int32_t)std::min<unsigned long long>(lv, INT32_MAX);  // BFS levels over all sources
  // TEPS convention for BC: each source traverses every edge once forward
  local.edges_scanned = (int64_t)nsrc * g->m;
  if (bc_out) HIPCHK(rt, hipMemcpy(bc_out, g->bc_sum, sizeof(double) * (size_t)g->n, hipMemcpyDeviceToHost));
  if (st) *st = local;
  return WGM_OK;
}

}  // extern "C"
