// Level-synchronous BFS from vertex 0 written against the scheduler stubs:
// the persistent loop of Lonestar's drelax2
// (src/lonestar/graph_apps/bfs/device/bfs.cl:214-302, worklists as in
// wl_pop_id / wl_push, bfs.cl:62-100) with one resizing barrier per level.
//
// Frontier entries are claimed in chunks of one workgroup's width from a
// per-level counter, so a workgroup forked in at a barrier, or killed at one,
// changes nothing in the result: a vertex's level is claimed once, by
// compare-and-swap. counts[] and claims[] rotate over three slots by
// level % 3: level L reads counts[L % 3] and pushes into counts[(L+1) % 3];
// workgroup 0 clears counts[(L+2) % 3] and claims[(L+1) % 3], whose last
// users finished before the previous barrier. `level` is the restorable
// variable a forked workgroup receives.
__global__ void bfs_level(const int* row_ptr, const int* col, int n, unsigned* dist, int* qa, int* qb, int* counts,
                          int* claims, int* levels_out) {
  int level = 0;
  __shared__ int s_c[1];
  for (int v = get_global_id(0); v < n; v += get_global_size(0)) dist[v] = v == 0 ? 0u : 1000000000u;
  if (get_global_id(0) == 0) {
    qa[0] = 0;
    counts[0] = 1;
    counts[1] = 0;
    counts[2] = 0;
    claims[0] = 0;
    claims[1] = 0;
    claims[2] = 0;
  }
  global_barrier();
  while (true) {
    int in_n = __hip_atomic_load(&counts[level % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (in_n == 0) {
      if (get_group_id(0) == 0 && get_local_id(0) == 0) *levels_out = level;
      break;
    }
    const int* in_q = (level & 1) ? qb : qa;
    int* out_q = (level & 1) ? qa : qb;
    while (true) {
      if (get_local_id(0) == 0) s_c[0] = atomicAdd(&claims[level % 3], get_local_size(0));
      __syncthreads();
      int base = s_c[0];
      __syncthreads();
      if (base >= in_n) break;
      int i = base + get_local_id(0);
      if (i < in_n) {
        int v = in_q[i];
        for (int e = row_ptr[v]; e < row_ptr[v + 1]; e++) {
          int u = col[e];
          if (atomicCAS(&dist[u], 1000000000u, (unsigned)(level + 1)) == 1000000000u) {
            int pos = atomicAdd(&counts[(level + 1) % 3], 1);
            out_q[pos] = u;
          }
        }
      }
      // a barrier between the thread-0 sections of consecutive claims keeps
      // the claim loop one loop (no path threaded into the next claim)
      __syncthreads();
    }
    if (get_group_id(0) == 0 && get_local_id(0) == 0) {
      __hip_atomic_store(&counts[(level + 2) % 3], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&claims[(level + 1) % 3], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    level++;
    resizing_global_barrier();
  }
}
