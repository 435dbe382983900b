// Octree construction from shared task pools, written against the
// scheduler stubs: an irregular persistent app in the shape of the
// reference's final_octree test (src/test_suite/final_octree/device/octree.cl:150-320).
// Workgroup 0 offers forks and the others offer kills at the top of the task
// loop; a workgroup pops a box from the pools (its own first), splits its
// particles into the eight octants through LDS counters, pushes every octant
// holding more than maxchilds particles as a new task and finalises the
// others as leaves (count into the tree, particles in place, a leaf record).
// The loop ends when every particle sits in a leaf.
//
// Boxes are [c - h, c + h) per axis; octant bit 0 / 1 / 2 is x / y / z on the
// upper side (x >= cx), and a child's centre is c +- h / 2. A box whose half
// width falls below min_h is a leaf whatever its count (coincident points).

struct OctTask {
  float cx, cy, cz, h;
  int flip;  // particles of the box are in newparticles (1) or particles (0)
  int beg, end, treepos;
};

// Pool lock, push and pop: thread 0 only.
__device__ inline void oct_lock(int* lk) {
  int e = 0;
  while (!__hip_atomic_compare_exchange_strong(lk, &e, 1, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT))
    e = 0;
}

__device__ inline void oct_unlock(int* lk) { __hip_atomic_store(lk, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT); }

__device__ inline bool oct_pop(OctTask* t, OctTask* pools, int* locks, int* heads, int pool_size, int id) {
  bool got = false;
  oct_lock(&locks[id]);
  if (heads[id] > 0) {
    heads[id] -= 1;
    *t = pools[pool_size * id + heads[id]];
    got = true;
  }
  oct_unlock(&locks[id]);
  return got;
}

__device__ inline bool oct_push(const OctTask* t, OctTask* pools, int* locks, int* heads, int pool_size, int id) {
  bool pushed = false;
  oct_lock(&locks[id]);
  if (heads[id] < pool_size) {
    pools[pool_size * id + heads[id]] = *t;
    heads[id] += 1;
    pushed = true;
  }
  oct_unlock(&locks[id]);
  return pushed;
}

__device__ inline int oct_box(float4 p, const OctTask& t) {
  return (p.x < t.cx ? 0 : 1) + (p.y < t.cy ? 0 : 2) + (p.z < t.cz ? 0 : 4);
}

__device__ inline OctTask oct_child(const OctTask& t, int i) {
  OctTask c;
  const float q = t.h * 0.5f;
  c.cx = t.cx + ((i & 1) ? q : -q);
  c.cy = t.cy + ((i & 2) ? q : -q);
  c.cz = t.cz + ((i & 4) ? q : -q);
  c.h = q;
  c.flip = t.flip ? 0 : 1;
  c.beg = 0;
  c.end = 0;
  c.treepos = 0;
  return c;
}

// status[0]: particles finalised; [1]: tree size; [2]: leaves; [3]: overflow flag
__global__ void octree(float4* particles, float4* newparticles, unsigned* tree, int tree_cap, int n, int maxchilds,
                       float root_h, float min_h, OctTask* pools, int* locks, int* heads, int num_pools, int pool_size,
                       int* status, int4* leaf_range, float4* leaf_box) {
  __shared__ unsigned count[8];
  __shared__ unsigned sum[8];
  __shared__ OctTask t[1];
  __shared__ int got[1];
  __shared__ int over[1];
  if (get_global_id(0) == 0) {
    for (int i = 0; i < num_pools; i++) {
      heads[i] = 0;
      locks[i] = 0;
    }
    status[0] = 0;
    status[1] = 8;
    status[2] = 0;
    status[3] = 0;
    OctTask root;
    root.cx = 0.0f;
    root.cy = 0.0f;
    root.cz = 0.0f;
    root.h = root_h;
    root.flip = 0;
    root.beg = 0;
    root.end = n;
    root.treepos = 0;
    oct_push(&root, pools, locks, heads, pool_size, 0);
  }
  global_barrier();
  while (true) {
    if (get_group_id(0) == 0) {
      offer_fork();
    }
    if (get_group_id(0) > 0) {
      offer_kill();
    }
    int pool_id = get_group_id(0) % num_pools;
    if (get_local_id(0) == 0) {
      got[0] = 0;
      for (int i = 0; i < num_pools && !got[0]; i++)
        got[0] = oct_pop(&t[0], pools, locks, heads, pool_size, (pool_id + i) % num_pools) ? 1 : 0;
      over[0] = !got[0] && __hip_atomic_load(&status[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= n;
      if (!got[0] && __hip_atomic_load(&status[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) over[0] = 1;
    }
    __syncthreads();
    int have = got[0];
    int stop = over[0];
    __syncthreads();
    if (!have) {
      if (stop) break;
      continue;
    }
    const float4* from = t[0].flip ? newparticles : particles;
    float4* to = t[0].flip ? particles : newparticles;
    for (int i = get_local_id(0); i < 8; i += get_local_size(0)) count[i] = 0;
    __syncthreads();
    for (int i = t[0].beg + get_local_id(0); i < t[0].end; i += get_local_size(0))
      atomicAdd(&count[oct_box(from[i], t[0])], 1u);
    __syncthreads();
    if (get_local_id(0) == 0) {
      sum[0] = count[0];
      for (int x = 1; x < 8; x++) sum[x] = sum[x - 1] + count[x];
    }
    __syncthreads();
    for (int i = t[0].beg + get_local_id(0); i < t[0].end; i += get_local_size(0)) {
      const float4 p = from[i];
      const unsigned pos = atomicSub(&sum[oct_box(p, t[0])], 1u) - 1u;
      to[t[0].beg + pos] = p;
    }
    __syncthreads();
    for (int i = 0; i < 8; i++) {
      OctTask c = oct_child(t[0], i);
      c.beg = t[0].beg + (int)sum[i];
      c.end = c.beg + (int)count[i];
      if ((int)count[i] > maxchilds && c.h >= min_h) {
        // every thread's scatter into `to` is released before the push
        __threadfence();
        __syncthreads();
        if (get_local_id(0) == 0) {
          const unsigned pos = atomicAdd((unsigned*)&status[1], 8u);
          if (pos + 8u > (unsigned)tree_cap) {
            __hip_atomic_store(&status[3], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          } else {
            tree[t[0].treepos + i] = pos;
            c.treepos = (int)pos;
            bool pushed = false;
            for (int j = 0; j < num_pools && !pushed; j++)
              pushed = oct_push(&c, pools, locks, heads, pool_size, (pool_id + j) % num_pools);
            if (!pushed) __hip_atomic_store(&status[3], 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
        __syncthreads();
      } else {
        // a leaf: its particles go to `particles` (already there when flip)
        if (!t[0].flip) {
          for (int j = c.beg + get_local_id(0); j < c.end; j += get_local_size(0)) particles[j] = newparticles[j];
        }
        __threadfence();
        __syncthreads();
        if (get_local_id(0) == 0) {
          tree[t[0].treepos + i] = 0x80000000u | count[i];
          const int k = atomicAdd(&status[2], 1);
          leaf_range[k] = make_int4(c.beg, (int)count[i], t[0].treepos + i, 0);
          leaf_box[k] = make_float4(c.cx, c.cy, c.cz, c.h);
          __hip_atomic_fetch_add(&status[0], (int)count[i], __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
      }
    }
  }
}
