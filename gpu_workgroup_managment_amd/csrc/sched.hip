// sched.hip -- resizing scheduler and mega-kernels (filled in by the
// scheduler milestone).
#include "apps_internal.hpp"

using namespace wgm;

bool wgm_sched_kernel_info(int kernel, KernelInfo* out) {
  (void)kernel;
  (void)out;
  return false;
}

int wgm_sched_probe(wgm_runtime* rt, int kernel, int grid, int wg_size) {
  (void)grid;
  (void)wg_size;
  return wgm_fail(rt, WGM_ERR_INVALID, "no discovery probe for kernel %d", kernel);
}

void wgm_runtime_release_extras(wgm_runtime* rt) { (void)rt; }
