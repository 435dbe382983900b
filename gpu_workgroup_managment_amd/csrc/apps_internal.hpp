// apps_internal.hpp -- device-resident graph object shared by the .hip units.
#pragma once

#include "runtime_internal.hpp"

struct wgm_graph {
  wgm_runtime* rt = nullptr;
  int32_t n = 0;
  int64_t m = 0;
  int32_t maxdeg = 0;  // largest out-degree
  int32_t* row_ptr = nullptr;
  int32_t* col = nullptr;
  int32_t* wt = nullptr;
  uint32_t* prio = nullptr;
  // per-run state, allocated on first use and kept resident
  uint32_t* dist = nullptr;
  uint32_t* stamp = nullptr;
  int32_t* q0 = nullptr;
  int32_t* q1 = nullptr;
  int* counts = nullptr;
  int32_t* i32a = nullptr;
  int32_t* i32b = nullptr;
  float* f0 = nullptr;
  float* f1 = nullptr;
  float* f2 = nullptr;
  int* iters = nullptr;
  int32_t* wl = nullptr;  // colour / MIS worklists: two lists of n entries
  // standalone BFS / SSSP (one team of every participant): two queues of
  // tcap entries with distance tags (SSSP), claim counters, near-far queues
  // and flags
  int32_t* tq = nullptr;
  uint32_t* tqa = nullptr;
  int64_t tcap = 0;
  int* claims = nullptr;
  int32_t* far = nullptr;
  uint32_t* in_far = nullptr;
  int* fcnt = nullptr;
  int32_t* hv = nullptr;  // heavy lists: m / kHeavyMin + 2 entries
  uint32_t* hdu = nullptr;
  int32_t* hclaim = nullptr;
  int* hcnt = nullptr;    // 3 lines
  // in-edge CSR and min-plus tasks (Pannotia SSSP / BC), built on first use
  int32_t* rrow = nullptr;
  int32_t* rcol = nullptr;
  int32_t* rwt = nullptr;
  void* tasks = nullptr;  // int4 {row, begin, end, 0}
  int32_t ntasks = 0;
  // BC per-slot state and the result
  int32_t bc_slots = 0;
  int32_t* bc_d = nullptr;
  double* bc_sig = nullptr;
  double* bc_del = nullptr;
  int32_t* bc_q = nullptr;
  int32_t* bc_lvl = nullptr;
  double* bc_sum = nullptr;
};

wgm::GraphDev wgm_graph_view(const wgm_graph* g);
// defined by the Pannotia unit (pannotia.hip)
bool wgm_pannotia_kernel_info(int kernel, KernelInfo* out);
int wgm_pannotia_probe(wgm_runtime* rt, int kernel, int grid, int wg_size);
void wgm_pannotia_release(wgm_graph* g);
int wgm_graph_frontier_state(wgm_graph* g, bool sssp, wgm::FrontierDev* s);
// per-run arrays of a topology-driven app: 1 colour (i32a), 2 MIS (i32a,
// i32b), 3 PageRank (f0, f1, f2); 0 nothing
int wgm_graph_app_state(wgm_graph* g, int app);
int wgm_graph_reached_edges(wgm_graph* g, int64_t* out);
int wgm_launch_frontier(wgm_runtime* rt, const wgm::GraphDev& g, wgm::FrontierDev s, const wgm::NearFar& nf, bool sssp,
                        int src, int grid, int wg, int handshake, int* iters_dev, wgm_run_stats* st);
int wgm_enqueue_frontier(wgm_runtime* rt, const wgm::GraphDev& g, wgm::FrontierDev s, const wgm::NearFar& nf, bool sssp,
                         int src, int grid, int wg, int* iters_dev);
uint32_t wgm_standalone_delta(int maxdeg);
int wgm_sssp_pull(wgm_runtime* rt, wgm_graph* g, int32_t src, const wgm_launch* lc, wgm_run_stats* st);
// the graph's in-edge CSR (g->rrow, g->rcol), built once and kept resident
int wgm_graph_in_edges(wgm_graph* g);
int wgm_enqueue_reached_edges(wgm_runtime* rt, const wgm::GraphDev& g, const uint32_t* dist, unsigned long long* out);
