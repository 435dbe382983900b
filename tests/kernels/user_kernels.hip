// user_kernels.hip -- a caller's own persistent kernels, written against
// include/wgm_device_api.h and launched through libwgm (test fixture).
//
// k_user_rounds: a standalone kernel using the stub names (get_group_id,
// get_num_groups, get_local_id, global_barrier): every round each workgroup
// writes its slot, and after the barrier every workgroup checks every slot.
//
// k_user_mega: the pair of kernels of kernel_merge's example
// (kernel_merge/examples/simple_barrier_with_reduction/{simple_barrier,reduction}.cl)
// merged by hand in the form of its expected output (expected/merged.cl):
// the persistent loop with a resizing barrier and a restoration context
// {target, i}, and the MY_reduce minimum as the non-persistent task.
#include <climits>

#include "wgm_device_api.h"

__global__ void k_user_rounds(wgm_kernel_args wa, int* slots, int* out, int rounds) {
  DISCOVERY_PROTOCOL(wa);
  const int g = get_group_id(0), n = get_num_groups(0);
  __shared__ int s_bad;
  for (int r = 1; r <= rounds; r++) {
    if (get_local_id(0) == 0) {
      slots[g] = r * 4096 + g;
      s_bad = 0;
    }
    global_barrier();
    int bad = 0;
    for (int i = get_local_id(0); i < n; i += get_local_size(0)) bad += slots[i] != r * 4096 + i;
    if (bad) atomicAdd(&s_bad, bad);
    __syncthreads();
    if (get_local_id(0) == 0 && s_bad) atomicAdd(&out[0], s_bad);
    global_barrier();
  }
  if (g == 0 && get_local_id(0) == 0) out[1] = n;
  for (int i = get_global_id(0); i < 4 * n; i += get_global_size(0)) atomicAdd(&out[2], 1);
  WGM_STANDALONE_END(wa);
}

struct Restoration_ctx {
  int target;
  int i;
};

__device__ void simple_barrier(const wgm_mega_args& ma, const Kernel_ctx* k, int* iters_out, int target_iters) {
  Restoration_ctx rc = wgm_restoration_ctx<Restoration_ctx>(ma);
  int i = 0;
  if (rc.target != 0) i = rc.i;
  while (true) {
    if (rc.target == 0) {
      i++;
      const Restoration_ctx to_fork{1, i};
      if (global_barrier_resize(ma, k, to_fork)) return;
    }
    rc.target = 0;
    if (i == target_iters) {
      if (k_get_group_id(k) == 0 && threadIdx.x == 0) *iters_out = i;
      return;
    }
  }
}

__device__ void my_reduce(int length, const int* buffer, int* result, int* calls, int* scratch, const Kernel_ctx* k) {
  const int li = get_local_id(0), ls = get_local_size(0);
  const int stride = k_get_global_size(k);
  for (int base = k_get_group_id(k) * ls; base < length; base += stride) {
    const int gi = base + li;
    scratch[li] = gi < length ? buffer[gi] : INT_MAX;
    __syncthreads();
    for (int offset = 1; offset < ls; offset <<= 1) {
      const int mask = (offset << 1) - 1;
      if ((li & mask) == 0 && li + offset < ls) scratch[li] = min(scratch[li], scratch[li + offset]);
      __syncthreads();
    }
    if (li == 0) atomicMin(result, scratch[0]);
    __syncthreads();
  }
  if (li == 0) atomicAdd(calls, 1);
}

__global__ void k_user_mega(wgm_mega_args ma, int length, const int* buffer, int* result, int* calls, int* iters_out,
                            int target_iters) {
  __shared__ int scratch[1024];
  WGM_MEGA_BODY(ma, my_reduce(length, buffer, result, calls, scratch, np_ctx),
                simple_barrier(ma, p_ctx, iters_out, target_iters));
}

// ---- host side of the fixture: kernel handles and plain device buffers
extern "C" {
const void* userk_rounds() { return (const void*)&k_user_rounds; }
const void* userk_mega() { return (const void*)&k_user_mega; }
int userk_alloc(long long bytes, void** out) { return hipMalloc(out, (size_t)bytes) == hipSuccess ? 0 : -1; }
int userk_free(void* p) { return hipFree(p) == hipSuccess ? 0 : -1; }
int userk_h2d(void* dst, const void* src, long long bytes) {
  return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}
int userk_d2h(void* dst, const void* src, long long bytes) {
  return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}
}
