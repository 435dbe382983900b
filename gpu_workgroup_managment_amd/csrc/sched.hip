// sched.hip -- resizing scheduler, mega-kernels and batched launches
// (filled in by the scheduler milestone).
#include "apps_internal.hpp"

using namespace wgm;

int wgm_discover_extra(wgm_runtime* rt, int kernel, int grid, int wg_size) {
  (void)grid;
  (void)wg_size;
  return wgm_fail(rt, WGM_ERR_INVALID, "no discovery probe for kernel %d", kernel);
}

bool wgm_extra_kernel_info(int kernel, KernelInfo* out) {
  (void)kernel;
  (void)out;
  return false;
}

void wgm_runtime_release_extras(wgm_runtime* rt) { (void)rt; }
