// The reference's non-persistent task written against the scheduler stubs
// (MY_matmult, src/non_persistent_kernels/matmult/device/matmult.cl:26-67):
// C = A B over int32 with wrap-around, element by element in a grid-stride
// loop; the workgroup that finishes last hashes C's diagonal with djb2
// (matmult.cl:60-66) into *hash, which lives in coherent host memory so the
// host can check it while the mega-kernel still holds the GPU. That
// workgroup also re-arms the counter for the next task.
__global__ void MY_matmult(const int* A, int a_rows, int a_cols, const int* B, int b_cols, int* C, int* counter,
                           int* hash) {
  __shared__ int s_last[1];
  for (int i = get_global_id(0); i < a_rows * b_cols; i += get_global_size(0)) {
    int r = i / b_cols;
    int c = i % b_cols;
    unsigned acc = 0;
    for (int k = 0; k < a_cols; k++) acc += (unsigned)A[r * a_cols + k] * (unsigned)B[k * b_cols + c];
    C[i] = (int)acc;
  }
  __syncthreads();
  if (get_local_id(0) == 0) {
    __threadfence();
    s_last[0] = atomicAdd(counter, 1) == get_num_groups(0) - 1;
  }
  __syncthreads();
  if (s_last[0] && get_local_id(0) == 0) {
    __threadfence();
    unsigned h = 5381;
    for (int i = 0; i < a_rows && i < b_cols; i++)
      h = h * 33u + (unsigned)__hip_atomic_load(&C[i * b_cols + i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(hash, (int)h, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
