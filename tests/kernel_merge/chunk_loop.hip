// Persistent work-queue kernel written against the scheduler stubs, in the
// shape of the octree example's loop (kernel_merge/examples/octree/
// octree_km_input.cl: workgroup 0 offers forks, the others offer kills, a
// shared counter hands out work, the loop ends with a break).
//
// Every chunk c of rounds * ceil(n / chunk) is claimed once: out[i] = f(in[i])
// over its elements and hits[c] += 1. `scale` is set by the initialisation,
// which forked workgroups skip: they receive it in the restoration context.
__global__ void chunk_loop(const unsigned* in, unsigned* out, int* hits, int* next, int n, int chunk, int rounds,
                           unsigned mult) {
  unsigned scale = 0;
  int claimed = 0;
  __shared__ int s_c[1];
  if (get_global_id(0) == 0) {
    *next = 0;
  }
  scale = mult;
  global_barrier();
  while (true) {
    int g = get_group_id(0);
    if (g == 0) {
      offer_fork();
    }
    if (get_group_id(0) > 0) {
      offer_kill();
    }
    if (get_local_id(0) == 0) s_c[0] = atomicAdd(next, 1);
    __syncthreads();
    int c = s_c[0];
    __syncthreads();
    int chunks = (n + chunk - 1) / chunk;
    if (c >= rounds * chunks) break;
    if (get_local_id(0) == 0) atomicAdd(&hits[c], 1);
    int lo = (c % chunks) * chunk;
    int hi = lo + chunk < n ? lo + chunk : n;
    for (int i = lo + get_local_id(0); i < hi; i += get_local_size(0)) {
      unsigned x = in[i];
      for (int k = 0; k < 16; k++) x = x * 1664525u + 1013904223u;
      out[i] = x * scale + 1u;
    }
    claimed++;
    // Keep a barrier between the thread-0 sections of consecutive
    // iterations: without one, the compiler may thread the end of one
    // iteration into the next one's claim and split wave 0 across a loop
    // of barriers that only its lane 0 can leave.
    __syncthreads();
  }
}
