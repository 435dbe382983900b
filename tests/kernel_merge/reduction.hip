// Non-persistent kernel written against the scheduler stubs: the minimum of
// `buffer` into *result (kernel_merge/examples/simple_barrier_with_reduction/
// reduction.cl), with a per-workgroup call count for the tests.
__global__ void MY_reduce(int length, const int* buffer, int* result, int* calls) {
  __shared__ int scratch[1024];
  int local_index = get_local_id(0);
  int local_size = get_local_size(0);
  int stride = get_global_size(0);
  for (int base = get_group_id(0) * local_size; base < length; base += stride) {
    int global_index = base + local_index;
    scratch[local_index] = global_index < length ? buffer[global_index] : INT_MAX;
    __syncthreads();
    for (int offset = 1; offset < local_size; offset <<= 1) {
      int mask = (offset << 1) - 1;
      if ((local_index & mask) == 0 && local_index + offset < local_size) {
        int other = scratch[local_index + offset];
        int mine = scratch[local_index];
        scratch[local_index] = mine < other ? mine : other;
      }
      __syncthreads();
    }
    if (local_index == 0) atomicMin(result, scratch[0]);
    __syncthreads();
  }
  if (local_index == 0) atomicAdd(calls, 1);
}
