// Plugins for the host skeleton (wgm_run_persistent / wgm_run_merged), as a
// caller would write them:
//  - the Lonestar BFS persistent app (src/lonestar/graph_apps/bfs/bfs.h:192-370):
//    reads a Galois .gr graph, runs bfs_level (kernel_merge's standalone or
//    merged form) and checks the levels against a "%d:%d" solution file;
//  - the matmult non-persistent task (src/non_persistent_kernels/matmult/matmult.h:26-200):
//    middle-square inputs from seed 1234, int32 product, djb2 hash of the
//    diagonal checked against the host's product.
// km_bfs_app / km_matmult_task fill the vtables for a caller (the tests).
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "wgm.h"

extern "C" const void* km_standalone_bfs();

namespace {

template <typename T>
int dalloc(T** p, size_t count) {
  return hipMalloc((void**)p, (count ? count : 1) * sizeof(T)) == hipSuccess ? WGM_OK : WGM_ERR_NOMEM;
}

int push_ptr(wgm_args* a, const void* p) { return wgm_args_push(a, &p, sizeof(p)); }
int push_int(wgm_args* a, int v) { return wgm_args_push(a, &v, sizeof(v)); }

// ------------------------------------------------------------ BFS app
struct BfsApp {
  std::string graph_file, solution_file;
  wgm_csr g{};
  int32_t *row = nullptr, *col = nullptr, *qa = nullptr, *qb = nullptr, *counts = nullptr, *claims = nullptr,
          *levels = nullptr;
  uint32_t* dist = nullptr;
  int checks = 0;
};

const char* bfs_name(void*) { return "lonestar_bfs"; }
const void* bfs_kernel(void*) { return km_standalone_bfs(); }

// read the graph file and upload it with the per-run state
int bfs_init_occupancy(void* self, wgm_runtime*) {
  BfsApp* a = (BfsApp*)self;
  int rc = wgm_read_graph(a->graph_file.c_str(), WGM_FMT_GR, 0, &a->g);
  if (rc) return rc;
  const size_t n = (size_t)a->g.n, m = (size_t)a->g.m;
  rc = dalloc(&a->row, n + 1);
  if (!rc) rc = dalloc(&a->col, m);
  if (!rc) rc = dalloc(&a->dist, n);
  if (!rc) rc = dalloc(&a->qa, n + 1);
  if (!rc) rc = dalloc(&a->qb, n + 1);
  if (!rc) rc = dalloc(&a->counts, 3);
  if (!rc) rc = dalloc(&a->claims, 3);
  if (!rc) rc = dalloc(&a->levels, 1);
  if (rc) return rc;
  if (hipMemcpy(a->row, a->g.row_ptr, 4 * (n + 1), hipMemcpyHostToDevice) != hipSuccess) return WGM_ERR_HIP;
  if (m && hipMemcpy(a->col, a->g.col, 4 * m, hipMemcpyHostToDevice) != hipSuccess) return WGM_ERR_HIP;
  return WGM_OK;
}

// the state does not depend on the number of workgroups
int bfs_init_real(void*, wgm_runtime*, int32_t) { return WGM_OK; }

int bfs_args(void* self, wgm_args* args) {
  BfsApp* a = (BfsApp*)self;
  int rc = push_ptr(args, a->row);
  if (!rc) rc = push_ptr(args, a->col);
  if (!rc) rc = push_int(args, a->g.n);
  if (!rc) rc = push_ptr(args, a->dist);
  if (!rc) rc = push_ptr(args, a->qa);
  if (!rc) rc = push_ptr(args, a->qb);
  if (!rc) rc = push_ptr(args, a->counts);
  if (!rc) rc = push_ptr(args, a->claims);
  if (!rc) rc = push_ptr(args, a->levels);
  return rc;
}

// the kernel's initialisation resets levels and worklists; poison the
// output so that a run that does not write it fails the check
int bfs_reset(void* self, wgm_runtime*) {
  BfsApp* a = (BfsApp*)self;
  return hipMemset(a->dist, 0xff, 4 * (size_t)a->g.n) == hipSuccess ? WGM_OK : WGM_ERR_HIP;
}

// diff against the solution file (diff_solution_file_int)
int bfs_check(void* self, wgm_runtime*) {
  BfsApp* a = (BfsApp*)self;
  std::vector<int32_t> d((size_t)a->g.n);
  if (hipMemcpy(d.data(), a->dist, 4 * d.size(), hipMemcpyDeviceToHost) != hipSuccess) return WGM_ERR_HIP;
  a->checks++;
  int32_t bad = -1;
  return wgm_check_solution(a->solution_file.c_str(), WGM_SOL_COLON, d.data(), a->g.n, &bad);
}

int bfs_clean(void* self, wgm_runtime*) {
  BfsApp* a = (BfsApp*)self;
  void* ptrs[] = {a->row, a->col, a->dist, a->qa, a->qb, a->counts, a->claims, a->levels};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  a->row = a->col = a->qa = a->qb = a->counts = a->claims = a->levels = nullptr;
  a->dist = nullptr;
  wgm_csr_free(&a->g);
  a->g = wgm_csr{};
  return WGM_OK;
}

// ------------------------------------------------------- matmult task
struct MatmultTask {
  int dim = 256;
  int32_t expect = 0;
  int32_t *A = nullptr, *B = nullptr, *C = nullptr, *counter = nullptr;
  int* h_hash = nullptr;  // coherent host memory: written by the device, read by check
  int* d_hash = nullptr;
  int checks = 0;
};

const char* mm_name(void*) { return "matmult"; }
const void* mm_kernel(void*) { return nullptr; }

// middle-square digit fill, A then B from one stream (matmult.h:38-50), and
// the expected djb2 hash of the host product's diagonal (matmult.h:144-149)
int mm_init(void* self, wgm_runtime*) {
  MatmultTask* t = (MatmultTask*)self;
  const size_t nn = (size_t)t->dim * t->dim;
  std::vector<int32_t> a(nn), b(nn);
  int32_t seed = 1234;
  for (auto* m : {&a, &b})
    for (size_t i = 0; i < nn; i++) {
      seed = (int32_t)((uint32_t)seed * (uint32_t)seed);
      seed = (seed / 1000) % 1000000;
      (*m)[i] = seed % 10;
    }
  uint32_t h = 5381;
  for (int i = 0; i < t->dim; i++) {
    uint32_t acc = 0;
    for (int k = 0; k < t->dim; k++) acc += (uint32_t)a[(size_t)i * t->dim + k] * (uint32_t)b[(size_t)k * t->dim + i];
    h = h * 33u + acc;
  }
  t->expect = (int32_t)h;
  int rc = dalloc(&t->A, nn);
  if (!rc) rc = dalloc(&t->B, nn);
  if (!rc) rc = dalloc(&t->C, nn);
  if (!rc) rc = dalloc(&t->counter, 1);
  if (rc) return rc;
  if (hipMemcpy(t->A, a.data(), 4 * nn, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(t->B, b.data(), 4 * nn, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(t->counter, 0, 4) != hipSuccess)
    return WGM_ERR_HIP;
  if (hipHostMalloc((void**)&t->h_hash, 64, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess)
    return WGM_ERR_NOMEM;
  if (hipHostGetDevicePointer((void**)&t->d_hash, t->h_hash, 0) != hipSuccess) return WGM_ERR_HIP;
  return WGM_OK;
}

int mm_args(void* self, wgm_args* args) {
  MatmultTask* t = (MatmultTask*)self;
  int rc = push_ptr(args, t->A);
  if (!rc) rc = push_int(args, t->dim);
  if (!rc) rc = push_int(args, t->dim);
  if (!rc) rc = push_ptr(args, t->B);
  if (!rc) rc = push_int(args, t->dim);
  if (!rc) rc = push_ptr(args, t->C);
  if (!rc) rc = push_ptr(args, t->counter);
  if (!rc) rc = push_ptr(args, t->d_hash);
  return rc;
}

// host memory only: the mega-kernel holds the GPU while tasks run
int mm_reset(void* self) {
  MatmultTask* t = (MatmultTask*)self;
  __hip_atomic_store(t->h_hash, INT_MIN, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  return WGM_OK;
}

int mm_check(void* self) {
  MatmultTask* t = (MatmultTask*)self;
  t->checks++;
  const int got = __hip_atomic_load(t->h_hash, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
  return got == t->expect ? WGM_OK : WGM_ERR_MISMATCH;
}

int mm_clean(void* self, wgm_runtime*) {
  MatmultTask* t = (MatmultTask*)self;
  void* ptrs[] = {t->A, t->B, t->C, t->counter};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (t->h_hash) (void)hipHostFree(t->h_hash);
  t->A = t->B = t->C = t->counter = nullptr;
  t->h_hash = t->d_hash = nullptr;
  return WGM_OK;
}

}  // namespace

extern "C" {

// A BFS app on `graph_file` (.gr) checked against `solution_file`.
int km_bfs_app(const char* graph_file, const char* solution_file, wgm_persistent_app* out) {
  if (!graph_file || !solution_file || !out) return WGM_ERR_INVALID;
  BfsApp* a = new BfsApp();
  a->graph_file = graph_file;
  a->solution_file = solution_file;
  *out = wgm_persistent_app{a,           bfs_name, bfs_kernel, bfs_init_occupancy, bfs_init_real, bfs_args,
                            bfs_args,    bfs_reset, bfs_check, bfs_clean};
  return WGM_OK;
}

int km_bfs_checks(const wgm_persistent_app* app) { return ((const BfsApp*)app->self)->checks; }
void km_bfs_free(wgm_persistent_app* app) {
  delete (BfsApp*)app->self;
  app->self = nullptr;
}

// The matmult task on dim x dim matrices.
int km_matmult_task(int32_t dim, wgm_np_task* out) {
  if (dim < 1 || !out) return WGM_ERR_INVALID;
  MatmultTask* t = new MatmultTask();
  t->dim = dim;
  *out = wgm_np_task{t, mm_name, mm_kernel, mm_init, mm_args, mm_reset, mm_check, mm_clean};
  return WGM_OK;
}

int km_matmult_checks(const wgm_np_task* np) { return ((const MatmultTask*)np->self)->checks; }
int32_t km_matmult_expected(const wgm_np_task* np) { return ((const MatmultTask*)np->self)->expect; }
void km_matmult_free(wgm_np_task* np) {
  delete (MatmultTask*)np->self;
  np->self = nullptr;
}

}  // extern "C"
