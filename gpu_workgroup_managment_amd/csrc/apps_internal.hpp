// apps_internal.hpp -- device-resident graph object shared by the .hip units.
#pragma once

#include "runtime_internal.hpp"

struct wgm_graph {
  wgm_runtime* rt = nullptr;
  int32_t n = 0;
  int64_t m = 0;
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
};

wgm::GraphDev wgm_graph_view(const wgm_graph* g);
int wgm_graph_frontier_state(wgm_graph* g, bool sssp, wgm::FrontierDev* s);
int wgm_graph_reached_edges(wgm_graph* g, int64_t* out);
