// apps.hip -- device-resident graphs and the standalone persistent apps.
//
// Plays the role of the reference's persistent-app plugins plus run_persistent:
// init_persistent_app_for_occupancy uploads the graph, reset_persistent_task
// re-initialises state, the driver launches the standalone kernel, waits on the
// handshake and check_persistent_task reads the result back
// (src/lonestar/graph_apps/bfs/bfs.h:209-370, src/pannotia/graph_apps/color/color.h:41-195,
// src/pannotia/graph_apps/mis/mis.h:40-202, src/host_skeleton/skel.cpp:250-385).
// Unlike the reference, a reset never re-uploads the graph: graph arrays stay
// resident and only the per-run state is re-initialised on the device.

#include "apps_internal.hpp"
#include "team_frontier.hpp"

#include <algorithm>
#include <cstring>

using namespace wgm;

namespace wgm {

// ====================================================== standalone kernels
// Shape of each (standalone.cl, e.g. src/lonestar/graph_apps/bfs/device/standalone.cl:247-320):
// discovery; host handshake; barrier; the application loop with global
// barriers (here the XCD-aware two-level barrier); retire.

// Band widths of the standalone SSSP's near-far bands for weights in
// [1, 255]: wide on graphs of out-degree <= 4 (long, narrow traversals: fewer
// levels), narrower on the others (short, wide ones: less rework). Measured on
// RMAT scale 20 (ef 16), push direction, 1024 x 64 workgroups, four edges per
// thread per pass and the heavy entries claimed as one list: 3.32 ms at 12
// (78 levels), 3.09 at 16, 2.83 at 32, 2.76 at 48, 2.66 at 96, 2.65 at 128
// (18 levels), 2.70 at 192 and 256, 2.67 at 512. With one edge per pass and
// the heavy entries claimed one by one, 16 had been best (6.32 ms at 256
// workgroups; 7.21 at 64, 8.24 at 1024): wide levels then cost more.
constexpr uint32_t kStandaloneDelta = 1024;
constexpr uint32_t kStandaloneDeltaHighDegree = 128;

// BFS / SSSP over the whole grid as one team (run_team): state shared through
// agent-scope atomics, so the per-level two-level barrier skips its L2
// write-back and invalidation (which would refetch the graph every level);
// SSSP queue entries carry distance tags and run in near-far bands.
template <bool SSSP>
__global__ void __launch_bounds__(1024) k_frontier(Ctl c, GraphDev g, FrontierDev s, NearFar nf, int src,
                                                   int* iters_out) {
  extern __shared__ int lds[];
  Team t;
  if (!standalone_begin(c, t)) return;
  run_team<SSSP, SSSP ? kEdgeSlots : 1, true>(g, s, nf, src, t.id, t.n, lds, iters_out,
                                              [&]() { xcd_barrier<false>(c, t); });
  clock_end(c, t.id);
  standalone_retire(c);
}

// wl: two lists of n entries; a workgroup uses [lo, hi) of each for the open
// vertices of its block (worklist rounds). The open count of round r is
// s_wl[r % 3]: reset two rounds after its last read.
__global__ void __launch_bounds__(1024) k_color(Ctl c, GraphDev g, int32_t* color, int32_t* wl, int* iters_out) {
  __shared__ int s_wl[3];
  Team t;
  if (!standalone_begin(c, t)) return;
  int lo, hi;
  block_range(g.n, t.id, t.n, &lo, &hi);
  if (threadIdx.x == 0) s_wl[0] = s_wl[1] = s_wl[2] = 0;
  __syncthreads();
  int* flags = c.lines;
  int round = 1, cnt = hi - lo;
  const int32_t* lin = nullptr;
  for (;; round++) {
    if (threadIdx.x == 0) s_wl[(round + 1) % 3] = 0;
    int32_t* lout = (round & 1) ? wl : wl + g.n;
    const bool any = color_round_wl(g, color, round, lo, cnt, lin, lout, &s_wl[round % 3]);
    if (__syncthreads_or(any) && threadIdx.x == 0) st_relaxed(&flags[(round % 3) * kLine], 1);
    if (t.id == 0 && threadIdx.x == 0) st_relaxed(&flags[((round + 1) % 3) * kLine], 0);
    xcd_barrier(c, t);
    if (ld_relaxed(&flags[(round % 3) * kLine]) == 0) break;
    cnt = s_wl[round % 3];
    lin = lout;
  }
  clock_end(c, t.id);
  if (t.id == 0 && threadIdx.x == 0) *iters_out = round;
  standalone_retire(c);
}

__global__ void __launch_bounds__(1024) k_mis(Ctl c, GraphDev g, int32_t* st, int32_t* s_out, int32_t* wl,
                                             int* iters_out) {
  __shared__ int s_wl[3];
  Team t;
  if (!standalone_begin(c, t)) return;
  int lo, hi;
  block_range(g.n, t.id, t.n, &lo, &hi);
  if (threadIdx.x == 0) s_wl[0] = s_wl[1] = s_wl[2] = 0;
  __syncthreads();
  int* flags = c.lines;
  int round = 1, cnt = hi - lo;
  const int32_t* lin = nullptr;
  for (;; round++) {
    if (threadIdx.x == 0) s_wl[(round + 1) % 3] = 0;
    int32_t* lout = (round & 1) ? wl : wl + g.n;
    const bool any = mis_phase_a_wl(g, st, s_out, round, lo, cnt, lin);
    if (__syncthreads_or(any) && threadIdx.x == 0) st_relaxed(&flags[(round % 3) * kLine], 1);
    if (t.id == 0 && threadIdx.x == 0) st_relaxed(&flags[((round + 1) % 3) * kLine], 0);
    xcd_barrier(c, t);
    if (ld_relaxed(&flags[(round % 3) * kLine]) == 0) break;
    mis_phase_b_wl(g, st, round, lo, cnt, lin, lout, &s_wl[round % 3]);
    xcd_barrier(c, t);
    cnt = s_wl[round % 3];
    lin = lout;
  }
  clock_end(c, t.id);
  if (t.id == 0 && threadIdx.x == 0) *iters_out = round;
  standalone_retire(c);
}

__global__ void __launch_bounds__(1024, 8) k_pagerank(Ctl c, GraphDev g, float* c0, float* c1, float* rank,
                                                   float damping, float base, float r0, int iters,
                                                   int* iters_out) {
  Team t;
  if (!standalone_begin(c, t)) return;
  int lo, hi;
  block_range(g.n, t.id, t.n, &lo, &hi);
  pagerank_init(g, c0, r0, lo, hi);
  xcd_barrier(c, t);
  for (int it = 0; it < iters; it++) {
    const float* cin = (it & 1) ? c1 : c0;
    float* cout = (it & 1) ? c0 : c1;
    pagerank_iter<true>(g, cin, cout, rank, damping, base, it == iters - 1, lo, hi);
    xcd_barrier(c, t);
  }
  clock_end(c, t.id);
  if (t.id == 0 && threadIdx.x == 0) *iters_out = iters;
  standalone_retire(c);
}

// Sum of out-degrees of reached vertices (dist != INF): the TEPS numerator.
__global__ void k_reached_edges(GraphDev g, const uint32_t* dist, unsigned long long* out) {
  unsigned long long local = 0;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += gridDim.x * blockDim.x)
    if (dist[v] != kInf) local += (unsigned long long)(g.row_ptr[v + 1] - g.row_ptr[v]);
  for (int d = 32; d > 0; d >>= 1) local += __shfl_down(local, d, 64);
  if ((threadIdx.x & 63) == 0 && local) atomicAdd(out, local);
}



__global__ void k_fill_i32(int32_t* p, int n, int32_t v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace wgm


#define fail wgm_fail
#define HIPCHK WGM_HIPCHK

namespace {

template <typename T>
int dev_alloc(wgm_runtime* rt, T** p, size_t count) {
  if (*p) return WGM_OK;
  hipError_t e = hipMalloc((void**)p, (count ? count : 1) * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return wgm_hip_fail(rt, e, "hipMalloc(graph state)");
  }
  return WGM_OK;
}

int launch_prologue(wgm_runtime* rt, int kernel, const wgm_launch* lc, int* grid, int* wg) {
  wgm_launch def{256, 0, 0, 0};
  const wgm_launch& l = lc ? *lc : def;
  *wg = l.wg_size > 0 ? l.wg_size : 256;
  int rc = wgm_resolve_grid(rt, kernel, *wg, l.num_wgs, grid);
  if (rc) return rc;
  rc = wgm_reset_ctl(rt, 0, l.handshake ? 1 : 0);
  if (rc) return rc;
  HIPCHK(rt, hipEventRecord(rt->ev0, rt->stream));
  return WGM_OK;
}

int launch_epilogue(wgm_runtime* rt, const wgm_launch* lc, wgm_graph* g, wgm_run_stats* st) {
  HIPCHK(rt, hipGetLastError());
  int rc = WGM_OK;
  double host_ms = 0.0;
  if (lc && lc->handshake) rc = wgm_host_handshake(rt, &host_ms);
  wgm_run_stats local{};
  int rc2 = wgm_finish_launch(rt, &local);
  if (rc == WGM_OK) rc = rc2;
  if (rc) return rc;
  int iters = 0;
  HIPCHK(rt, hipMemcpy(&iters, g->iters, sizeof(int), hipMemcpyDeviceToHost));
  local.iterations = iters;
  local.host_ms = host_ms;
  if (st) *st = local;
  return WGM_OK;
}

}  // namespace

GraphDev wgm_graph_view(const wgm_graph* g) {
  GraphDev v{};
  v.row_ptr = g->row_ptr;
  v.col = g->col;
  v.wt = g->wt;
  v.prio = g->prio;
  v.n = g->n;
  v.m = (int32_t)g->m;
  return v;
}

int wgm_graph_frontier_state(wgm_graph* g, bool sssp, FrontierDev* s) {
  wgm_runtime* rt = g->rt;
  int rc = dev_alloc(rt, &g->dist, g->n);
  if (!rc) rc = dev_alloc(rt, &g->q0, g->n + 1);
  if (!rc) rc = dev_alloc(rt, &g->q1, g->n + 1);
  if (!rc) rc = dev_alloc(rt, &g->counts, 3 * kLine);
  if (!rc && sssp) rc = dev_alloc(rt, &g->stamp, g->n);
  if (rc) return rc;
  s->dist = g->dist;
  s->stamp = sssp ? g->stamp : nullptr;
  s->q0 = g->q0;
  s->q1 = g->q1;
  s->counts = g->counts;
  return WGM_OK;
}

int wgm_graph_app_state(wgm_graph* g, int app) {
  wgm_runtime* rt = g->rt;
  int rc = WGM_OK;
  if (app == 1 || app == 2) rc = dev_alloc(rt, &g->i32a, g->n);
  if (!rc && app == 2) rc = dev_alloc(rt, &g->i32b, g->n);
  if (!rc && app == 3) rc = dev_alloc(rt, &g->f0, g->n);
  if (!rc && app == 3) rc = dev_alloc(rt, &g->f1, g->n);
  if (!rc && app == 3) rc = dev_alloc(rt, &g->f2, g->n);
  return rc;
}

// State of the standalone team traversal: queues of max(n, m) + 1 entries
// with distance tags for SSSP (a level queues at most one entry per
// relaxation), n + 1 for BFS; claim counters; near-far queues and flags.
static int team_frontier_state(wgm_graph* g, bool sssp, FrontierDev* s, NearFar* nf) {
  wgm_runtime* rt = g->rt;
  const int64_t cap = (sssp ? std::max<int64_t>(g->n, g->m) : (int64_t)g->n) + 1;
  if (g->tcap < cap) {  // a BFS-sized pair of queues cannot serve SSSP
    (void)hipFree(g->tq);
    (void)hipFree(g->tqa);
    g->tq = nullptr;
    g->tqa = nullptr;
    g->tcap = 0;
  }
  int rc = dev_alloc(rt, &g->dist, g->n);
  if (!rc) rc = dev_alloc(rt, &g->tq, 2 * (size_t)cap);
  if (!rc && sssp) rc = dev_alloc(rt, &g->tqa, 2 * (size_t)std::max<int64_t>(cap, g->tcap));
  if (!rc) rc = dev_alloc(rt, &g->counts, 3 * kLine);
  if (!rc) rc = dev_alloc(rt, &g->claims, 3 * kLine);
  if (!rc && sssp) rc = dev_alloc(rt, &g->far, 2 * ((size_t)g->n + 1));
  if (!rc && sssp) rc = dev_alloc(rt, &g->in_far, (size_t)g->n);
  if (!rc && sssp) rc = dev_alloc(rt, &g->fcnt, 3 * kLine);
  const size_t hcap = (size_t)(g->m / kHeavyMin) + 2;
  if (!rc) rc = dev_alloc(rt, &g->hv, hcap);
  if (!rc) rc = dev_alloc(rt, &g->hdu, hcap);
  if (!rc) rc = dev_alloc(rt, &g->hclaim, hcap);
  if (!rc) rc = dev_alloc(rt, &g->hcnt, 3 * kLine);
  if (rc) return rc;
  if (g->tcap < cap) g->tcap = cap;
  *s = FrontierDev{};
  s->dist = g->dist;
  s->q0 = g->tq;
  s->q1 = g->tq + g->tcap;
  s->qa0 = sssp ? g->tqa : nullptr;
  s->qa1 = sssp ? g->tqa + g->tcap : nullptr;
  s->counts = g->counts;
  s->claims = g->claims;
  s->hv = g->hv;
  s->hdu = g->hdu;
  s->hclaim = g->hclaim;
  s->hcnt = g->hcnt;
  *nf = NearFar{};
  if (sssp) {
    nf->q[0] = g->far;
    nf->q[1] = g->far + g->n + 1;
    nf->flag = g->in_far;
    nf->cnt = g->fcnt;
    nf->delta = wgm_standalone_delta(g->maxdeg);
  }
  return WGM_OK;
}

int wgm_graph_reached_edges(wgm_graph* g, int64_t* out) {
  wgm_runtime* rt = g->rt;
  unsigned long long* d = nullptr;
  HIPCHK(rt, hipMalloc(&d, sizeof(unsigned long long)));
  (void)hipMemsetAsync(d, 0, sizeof(unsigned long long), rt->stream);
  hipLaunchKernelGGL(k_reached_edges, dim3(256), dim3(256), 0, rt->stream, wgm_graph_view(g), g->dist, d);
  unsigned long long h = 0;
  hipError_t e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, rt->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(rt->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return wgm_hip_fail(rt, e, "reached edges");
  *out = (int64_t)h;
  return WGM_OK;
}

bool wgm_app_kernel_info(int kernel, KernelInfo* out) {
  switch (kernel) {
    case WGM_K_BFS: *out = KernelInfo{(const void*)&k_frontier<false>, wgm_frontier_lds}; return true;
    case WGM_K_SSSP: *out = KernelInfo{(const void*)&k_frontier<true>, wgm_frontier_lds}; return true;
    case WGM_K_COLOR: *out = KernelInfo{(const void*)&k_color, nullptr}; return true;
    case WGM_K_MIS: *out = KernelInfo{(const void*)&k_mis, nullptr}; return true;
    case WGM_K_PAGERANK: *out = KernelInfo{(const void*)&k_pagerank, nullptr}; return true;
    default: return wgm_pannotia_kernel_info(kernel, out);
  }
}

// Discovery probe with the app kernel itself (im_exit set: every workgroup
// returns right after discovery and never reads its other arguments).
int wgm_app_probe(wgm_runtime* rt, int kernel, int grid, int wg_size) {
  const size_t lds = wgm_kernel_lds(kernel, wg_size);
  switch (kernel) {
    case WGM_K_BFS:
      hipLaunchKernelGGL(k_frontier<false>, dim3(grid), dim3(wg_size), lds, rt->stream, rt->ctl, GraphDev{},
                         FrontierDev{}, NearFar{}, 0, (int*)nullptr);
      return WGM_OK;
    case WGM_K_SSSP:
      hipLaunchKernelGGL(k_frontier<true>, dim3(grid), dim3(wg_size), lds, rt->stream, rt->ctl, GraphDev{},
                         FrontierDev{}, NearFar{}, 0, (int*)nullptr);
      return WGM_OK;
    case WGM_K_COLOR:
      hipLaunchKernelGGL(k_color, dim3(grid), dim3(wg_size), 0, rt->stream, rt->ctl, GraphDev{}, (int32_t*)nullptr,
                         (int32_t*)nullptr, (int*)nullptr);
      return WGM_OK;
    case WGM_K_MIS:
      hipLaunchKernelGGL(k_mis, dim3(grid), dim3(wg_size), 0, rt->stream, rt->ctl, GraphDev{}, (int32_t*)nullptr,
                         (int32_t*)nullptr, (int32_t*)nullptr, (int*)nullptr);
      return WGM_OK;
    case WGM_K_PAGERANK:
      hipLaunchKernelGGL(k_pagerank, dim3(grid), dim3(wg_size), 0, rt->stream, rt->ctl, GraphDev{}, (float*)nullptr,
                         (float*)nullptr, (float*)nullptr, 0.f, 0.f, 0.f, 0, (int*)nullptr);
      return WGM_OK;
    default:
      return wgm_pannotia_probe(rt, kernel, grid, wg_size);
  }
}

extern "C" {

int wgm_graph_create(wgm_runtime* rt, const wgm_csr* h, const uint32_t* prio, wgm_graph** out) {
  if (!rt || !h || !out) return fail(rt, WGM_ERR_INVALID, "NULL argument");
  if (h->n < 1 || h->m < 0 || h->m > INT32_MAX || !h->row_ptr || (h->m > 0 && !h->col))
    return fail(rt, WGM_ERR_INVALID, "bad CSR shape n=%d m=%lld", h->n, (long long)h->m);
  if (h->row_ptr[0] != 0 || h->row_ptr[h->n] != h->m)
    return fail(rt, WGM_ERR_INVALID, "row_ptr must start at 0 and end at m");
  int32_t maxdeg = 0;
  for (int32_t v = 0; v < h->n; v++) {
    if (h->row_ptr[v + 1] < h->row_ptr[v]) return fail(rt, WGM_ERR_INVALID, "row_ptr not monotone at %d", v);
    maxdeg = std::max(maxdeg, h->row_ptr[v + 1] - h->row_ptr[v]);
  }
  for (int64_t e = 0; e < h->m; e++)
    if (h->col[e] < 0 || h->col[e] >= h->n) return fail(rt, WGM_ERR_INVALID, "col[%lld] out of range", (long long)e);
  wgm_graph* g = new wgm_graph();
  g->rt = rt;
  g->n = h->n;
  g->m = h->m;
  g->maxdeg = maxdeg;
  auto bail = [&](int rc) { wgm_graph_destroy(g); return rc; };
  int rc = dev_alloc(rt, &g->row_ptr, (size_t)h->n + 1);
  if (!rc) rc = dev_alloc(rt, &g->col, (size_t)h->m);
  if (!rc && h->wt) rc = dev_alloc(rt, &g->wt, (size_t)h->m);
  if (!rc && prio) rc = dev_alloc(rt, &g->prio, (size_t)h->n);
  if (!rc) rc = dev_alloc(rt, &g->iters, 1);
  if (rc) return bail(rc);
  hipError_t e = hipMemcpy(g->row_ptr, h->row_ptr, sizeof(int32_t) * ((size_t)h->n + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess && h->m) e = hipMemcpy(g->col, h->col, sizeof(int32_t) * (size_t)h->m, hipMemcpyHostToDevice);
  if (e == hipSuccess && h->wt && h->m) e = hipMemcpy(g->wt, h->wt, sizeof(int32_t) * (size_t)h->m, hipMemcpyHostToDevice);
  if (e == hipSuccess && prio) e = hipMemcpy(g->prio, prio, sizeof(uint32_t) * (size_t)h->n, hipMemcpyHostToDevice);
  if (e != hipSuccess) return bail(wgm_hip_fail(rt, e, "graph upload"));
  *out = g;
  return WGM_OK;
}

int wgm_graph_destroy(wgm_graph* g) {
  if (!g) return WGM_OK;
  wgm_pannotia_release(g);
  void* ptrs[] = {g->row_ptr, g->col, g->wt, g->prio, g->dist, g->stamp, g->q0, g->q1, g->counts,
                  g->i32a, g->i32b, g->f0, g->f1, g->f2, g->iters, g->tq, g->tqa, g->claims, g->far,
                  g->in_far, g->fcnt, g->hv, g->hdu, g->hclaim, g->hcnt, g->wl};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete g;
  return WGM_OK;
}

int64_t wgm_graph_bytes(const wgm_graph* g) {
  if (!g) return 0;
  int64_t b = 4 * ((int64_t)g->n + 1) + 4 * g->m;
  if (g->wt) b += 4 * g->m;
  if (g->prio) b += 4 * (int64_t)g->n;
  if (g->rrow) b += 4 * ((int64_t)g->n + 1) + (g->rwt ? 8 : 4) * g->m + 16 * (int64_t)g->ntasks;
  return b;
}

}  // extern "C"

// The standalone frontier kernel on any state: prologue (grid, control
// reset, start event), launch on `grid` workgroups of `wg` threads, then the
// finish (HIP event time, device clock, participants); *iters_dev is the
// device cell the kernel writes its level count to. Allocates nothing.
uint32_t wgm_standalone_delta(int maxdeg) { return maxdeg > 4 ? kStandaloneDeltaHighDegree : kStandaloneDelta; }

// Enqueue the standalone frontier kernel (no wait): chunked claims of one
// workgroup's width, heavy entries above heavy_threshold(wg).
int wgm_enqueue_frontier(wgm_runtime* rt, const GraphDev& g, FrontierDev s, const NearFar& nf, bool sssp, int src,
                         int grid, int wg, int* iters_dev) {
  s.chunk = wg;  // frontier entries a workgroup claims at a time
  s.heavy = heavy_threshold(wg);
  const size_t lds = frontier_lds_bytes(wg);
  if (sssp)
    hipLaunchKernelGGL(k_frontier<true>, dim3(grid), dim3(wg), lds, rt->stream, rt->ctl, g, s, nf, src, iters_dev);
  else
    hipLaunchKernelGGL(k_frontier<false>, dim3(grid), dim3(wg), lds, rt->stream, rt->ctl, g, s, nf, src, iters_dev);
  HIPCHK(rt, hipGetLastError());
  return WGM_OK;
}

int wgm_launch_frontier(wgm_runtime* rt, const GraphDev& g, FrontierDev s, const NearFar& nf, bool sssp, int src,
                        int grid, int wg, int handshake, int* iters_dev, wgm_run_stats* st) {
  int rc = wgm_reset_ctl(rt, 0, handshake ? 1 : 0);
  if (rc) return rc;
  HIPCHK(rt, hipEventRecord(rt->ev0, rt->stream));
  rc = wgm_enqueue_frontier(rt, g, s, nf, sssp, src, grid, wg, iters_dev);
  if (rc) return rc;
  double host_ms = 0.0;
  if (handshake) rc = wgm_host_handshake(rt, &host_ms);
  wgm_run_stats local{};
  const int rc2 = wgm_finish_launch(rt, &local);
  if (rc == WGM_OK) rc = rc2;
  if (rc) return rc;
  int iters = 0;
  HIPCHK(rt, hipMemcpy(&iters, iters_dev, sizeof(int), hipMemcpyDeviceToHost));
  local.iterations = iters;
  local.host_ms = host_ms;
  if (st) *st = local;
  return WGM_OK;
}

// Sum of out-degrees of the vertices with dist != INF into *out (a device
// cell the caller has cleared), on the runtime's stream.
int wgm_enqueue_reached_edges(wgm_runtime* rt, const GraphDev& g, const uint32_t* dist, unsigned long long* out) {
  hipLaunchKernelGGL(k_reached_edges, dim3(256), dim3(256), 0, rt->stream, g, dist, out);
  HIPCHK(rt, hipGetLastError());
  return WGM_OK;
}

extern "C" {

static int run_frontier(wgm_runtime* rt, wgm_graph* g, int32_t src, const wgm_launch* lc, uint32_t* dist_out,
                        wgm_run_stats* st, bool sssp) {
  if (!rt || !g || g->rt != rt) return fail(rt, WGM_ERR_INVALID, "bad runtime/graph");
  if (src < 0 || src >= g->n) return fail(rt, WGM_ERR_INVALID, "source %d out of range", src);
  if (sssp && !g->wt) return fail(rt, WGM_ERR_INVALID, "SSSP needs edge weights");
  const int mode = lc ? lc->mode : WGM_SSSP_AUTO;
  if (mode < WGM_SSSP_AUTO || mode > WGM_SSSP_PULL) return fail(rt, WGM_ERR_INVALID, "unknown SSSP mode %d", mode);
  if (sssp && (mode == WGM_SSSP_PULL || (mode == WGM_SSSP_AUTO && g->maxdeg > 4))) {
    wgm_run_stats local{};
    int rc = wgm_sssp_pull(rt, g, src, lc, &local);
    if (!rc) rc = wgm_graph_reached_edges(g, &local.edges_scanned);
    if (rc) return rc;
    if (dist_out) HIPCHK(rt, hipMemcpy(dist_out, g->dist, sizeof(uint32_t) * (size_t)g->n, hipMemcpyDeviceToHost));
    if (st) *st = local;
    return WGM_OK;
  }
  FrontierDev s{};
  NearFar nf{};
  int rc = team_frontier_state(g, sssp, &s, &nf);
  if (rc) return rc;
  const int kernel = sssp ? WGM_K_SSSP : WGM_K_BFS;
  const int wg = lc && lc->wg_size > 0 ? lc->wg_size : 256;
  int grid = 0;
  rc = wgm_resolve_grid(rt, kernel, wg, lc ? lc->num_wgs : 0, &grid);
  if (rc) return rc;
  // BFS takes its wide levels bottom-up over the in-edges
  GraphDev view = wgm_graph_view(g);
  if (!sssp) {
    rc = wgm_graph_in_edges(g);
    if (rc) return rc;
    view.rrow = g->rrow;
    view.rcol = g->rcol;
  }
  wgm_run_stats local{};
  rc = wgm_launch_frontier(rt, view, s, nf, sssp, src, grid, wg, lc ? lc->handshake : 0, g->iters, &local);
  if (rc) return rc;
  rc = wgm_graph_reached_edges(g, &local.edges_scanned);
  if (rc) return rc;
  if (dist_out) HIPCHK(rt, hipMemcpy(dist_out, g->dist, sizeof(uint32_t) * (size_t)g->n, hipMemcpyDeviceToHost));
  if (st) *st = local;
  return WGM_OK;
}

int wgm_bfs(wgm_runtime* rt, wgm_graph* g, int32_t src, const wgm_launch* lc, uint32_t* dist_out,
            wgm_run_stats* st) {
  return run_frontier(rt, g, src, lc, dist_out, st, false);
}

int wgm_sssp(wgm_runtime* rt, wgm_graph* g, int32_t src, const wgm_launch* lc, uint32_t* dist_out,
             wgm_run_stats* st) {
  return run_frontier(rt, g, src, lc, dist_out, st, true);
}

int wgm_color(wgm_runtime* rt, wgm_graph* g, const wgm_launch* lc, int32_t* color_out, wgm_run_stats* st) {
  if (!rt || !g || g->rt != rt) return fail(rt, WGM_ERR_INVALID, "bad runtime/graph");
  if (!g->prio) return fail(rt, WGM_ERR_INVALID, "colouring needs vertex priorities");
  int rc = dev_alloc(rt, &g->i32a, g->n);
  if (!rc) rc = dev_alloc(rt, &g->wl, 2 * (size_t)g->n);
  if (rc) return rc;
  hipLaunchKernelGGL(k_fill_i32, dim3(256), dim3(256), 0, rt->stream, g->i32a, g->n, -1);
  int grid = 0, wg = 0;
  rc = launch_prologue(rt, WGM_K_COLOR, lc, &grid, &wg);
  if (rc) return rc;
  hipLaunchKernelGGL(k_color, dim3(grid), dim3(wg), 0, rt->stream, rt->ctl, wgm_graph_view(g), g->i32a, g->wl,
                     g->iters);
  rc = launch_epilogue(rt, lc, g, st);
  if (rc) return rc;
  if (color_out) HIPCHK(rt, hipMemcpy(color_out, g->i32a, sizeof(int32_t) * (size_t)g->n, hipMemcpyDeviceToHost));
  return WGM_OK;
}

int wgm_mis(wgm_runtime* rt, wgm_graph* g, const wgm_launch* lc, int32_t* s_out, wgm_run_stats* st) {
  if (!rt || !g || g->rt != rt) return fail(rt, WGM_ERR_INVALID, "bad runtime/graph");
  if (!g->prio) return fail(rt, WGM_ERR_INVALID, "MIS needs vertex priorities");
  int rc = dev_alloc(rt, &g->i32a, g->n);
  if (!rc) rc = dev_alloc(rt, &g->i32b, g->n);
  if (!rc) rc = dev_alloc(rt, &g->wl, 2 * (size_t)g->n);
  if (rc) return rc;
  hipLaunchKernelGGL(k_fill_i32, dim3(256), dim3(256), 0, rt->stream, g->i32a, g->n, -1);
  hipLaunchKernelGGL(k_fill_i32, dim3(256), dim3(256), 0, rt->stream, g->i32b, g->n, 0);
  int grid = 0, wg = 0;
  rc = launch_prologue(rt, WGM_K_MIS, lc, &grid, &wg);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mis, dim3(grid), dim3(wg), 0, rt->stream, rt->ctl, wgm_graph_view(g), g->i32a, g->i32b, g->wl,
                     g->iters);
  rc = launch_epilogue(rt, lc, g, st);
  if (rc) return rc;
  if (s_out) HIPCHK(rt, hipMemcpy(s_out, g->i32b, sizeof(int32_t) * (size_t)g->n, hipMemcpyDeviceToHost));
  return WGM_OK;
}

int wgm_pagerank(wgm_runtime* rt, wgm_graph* g, float damping, int32_t iterations, const wgm_launch* lc,
                 float* rank_out, wgm_run_stats* st) {
  if (!rt || !g || g->rt != rt) return fail(rt, WGM_ERR_INVALID, "bad runtime/graph");
  if (iterations < 1 || !(damping >= 0.0f && damping <= 1.0f))
    return fail(rt, WGM_ERR_INVALID, "need iterations >= 1 and damping in [0, 1]");
  int rc = dev_alloc(rt, &g->f0, g->n);
  if (!rc) rc = dev_alloc(rt, &g->f1, g->n);
  if (!rc) rc = dev_alloc(rt, &g->f2, g->n);
  if (rc) return rc;
  const float nf = (float)g->n;
  const float base = (1.0f - damping) / nf;
  const float r0 = 1.0f / nf;
  int grid = 0, wg = 0;
  rc = launch_prologue(rt, WGM_K_PAGERANK, lc, &grid, &wg);
  if (rc) return rc;
  hipLaunchKernelGGL(k_pagerank, dim3(grid), dim3(wg), 0, rt->stream, rt->ctl, wgm_graph_view(g), g->f0, g->f1, g->f2,
                     damping, base, r0, iterations, g->iters);
  rc = launch_epilogue(rt, lc, g, st);
  if (rc) return rc;
  if (rank_out) HIPCHK(rt, hipMemcpy(rank_out, g->f2, sizeof(float) * (size_t)g->n, hipMemcpyDeviceToHost));
  return WGM_OK;
}

}  // extern "C"
