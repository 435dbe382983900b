// Persistent kernel written against the scheduler stubs: a loop of resizing
// barriers (kernel_merge/examples/simple_barrier_with_reduction/
// simple_barrier.cl); workgroup 0 records the iterations it reached.
__global__ void simple_barrier(int* iters_out, int target_iters) {
  int i = 0;
  while (true) {
    i++;
    resizing_global_barrier();
    if (i == target_iters) {
      if (get_group_id(0) == 0 && get_local_id(0) == 0) *iters_out = i;
      break;
    }
  }
}
