// skeleton.cpp -- the host skeleton: the reference's run_persistent and
// run_merged (src/host_skeleton/skel.cpp:250-385, 405-631) over the two
// plugin vtables of wgm.h, on top of the runtime's C ABI. Host code only.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

#include "wgm.h"

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// One ordered argument list, destroyed with its scope.
struct Args {
  wgm_args* a = nullptr;
  Args() { (void)wgm_args_create(&a); }
  ~Args() { (void)wgm_args_destroy(a); }
};

// get_workgroups_for_non_persistent (skel.cpp:387-401)
int np_groups_for(int ratio, int occupancy) {
  int k = ratio == -1 ? occupancy - 2 : ratio == -2 ? 1 : occupancy / std::max(1, ratio);
  return std::max(1, k);
}

wgm_skel_config defaults(const wgm_skel_config* in) {
  wgm_skel_config c = in ? *in : wgm_skel_config{};
  if (c.threads_per_wg <= 0) c.threads_per_wg = 128;
  if (c.iterations <= 0) c.iterations = 1;
  if (c.non_persistent_wgs == 0) c.non_persistent_wgs = -1;
  if (c.non_persistent_frequency_ms < 0) c.non_persistent_frequency_ms = 0;
  return c;
}

// The reference's persistent-time summary (skel.cpp:360-372).
void summarize(const std::vector<double>& t, wgm_skel_report* r) {
  if (t.empty()) return;
  double sum = 0, mn = t[0], mx = t[0];
  for (double x : t) {
    sum += x;
    mn = std::min(mn, x);
    mx = std::max(mx, x);
  }
  const double mean = sum / (double)t.size();
  double var = 0;
  for (double x : t) var += (x - mean) * (x - mean);
  r->total_ms = sum;
  r->mean_ms = mean;
  r->min_ms = mn;
  r->max_ms = mx;
  r->std_ms = std::sqrt(var / (double)t.size());
}

bool complete(const wgm_persistent_app* a) {
  return a && a->name && a->kernel && a->init_for_occupancy && a->init_for_real && a->set_args_for_occupancy &&
         a->set_args_for_real && a->reset && a->check && a->clean;
}

bool complete(const wgm_np_task* t) {
  return t && t->name && t->init && t->set_args && t->reset && t->check && t->clean;
}

}  // namespace

extern "C" {

int wgm_run_persistent(wgm_runtime* rt, const wgm_persistent_app* app, const wgm_skel_config* cfg_in,
                       wgm_skel_report* out) {
  if (!rt || !complete(app)) return WGM_ERR_INVALID;
  const wgm_skel_config cfg = defaults(cfg_in);
  void* self = app->self;
  const void* kernel = app->kernel(self);
  if (!kernel) return WGM_ERR_INVALID;
  wgm_skel_report rep{};
  int rc = app->init_for_occupancy(self, rt);
  if (rc) return rc;
  auto fail = [&](int code) {
    (void)app->clean(self, rt);
    return code;
  };
  int occupancy = 0;
  {
    Args occ;
    rc = app->set_args_for_occupancy(self, occ.a);
    if (!rc) rc = wgm_discover_occupancy_custom(rt, kernel, cfg.threads_per_wg, cfg.lds_bytes, occ.a, 0, &occupancy);
    if (rc) return fail(rc);
  }
  const int num_wgs = cfg.num_wgs > 0 ? std::min(cfg.num_wgs, occupancy) : occupancy;
  rc = app->init_for_real(self, rt, num_wgs);
  if (rc) return fail(rc);
  Args real;
  rc = app->set_args_for_real(self, real.a);
  if (rc) return fail(rc);
  std::vector<double> times;
  for (int i = 0; i < cfg.iterations; i++) {
    rc = app->reset(self, rt);
    if (rc) return fail(rc);
    wgm_run_stats st{};
    rc = wgm_launch_standalone(rt, kernel, cfg.threads_per_wg, num_wgs, cfg.lds_bytes, real.a, 1, &st);
    if (rc) return fail(rc);
    times.push_back(st.host_ms);
    rc = app->check(self, rt);
    if (rc == WGM_ERR_MISMATCH) rep.persistent_errors++;
    else if (rc) return fail(rc);
  }
  rc = app->clean(self, rt);
  rep.occupancy = occupancy;
  rep.num_wgs = num_wgs;
  rep.iterations = cfg.iterations;
  summarize(times, &rep);
  if (out) *out = rep;
  return rc;
}

int wgm_run_merged(wgm_runtime* rt, const wgm_persistent_app* app, const wgm_np_task* np, const void* merged_kernel,
                   const wgm_skel_config* cfg_in, wgm_skel_report* out) {
  if (!rt || !complete(app) || !complete(np) || !merged_kernel) return WGM_ERR_INVALID;
  const wgm_skel_config cfg = defaults(cfg_in);
  void* self = app->self;
  wgm_skel_report rep{};
  int rc = app->init_for_occupancy(self, rt);
  if (rc) return rc;
  rc = np->init(np->self, rt);
  if (rc) {
    (void)app->clean(self, rt);
    return rc;
  }
  wgm_mega* m = nullptr;
  auto fail = [&](int code) {
    if (m) (void)wgm_mega_destroy(m);  // quits a resident kernel first
    (void)app->clean(self, rt);
    (void)np->clean(np->self, rt);
    return code;
  };
  // arguments: the non-persistent task's first, then the persistent app's
  int occupancy = 0;
  {
    Args occ;
    rc = np->set_args(np->self, occ.a);
    if (!rc) rc = app->set_args_for_occupancy(self, occ.a);
    if (!rc) rc = wgm_discover_occupancy_custom(rt, merged_kernel, cfg.threads_per_wg, cfg.lds_bytes, occ.a, 1, &occupancy);
    if (rc) return fail(rc);
  }
  const int num_wgs = cfg.num_wgs > 0 ? std::min(cfg.num_wgs, occupancy) : occupancy;
  if (num_wgs < 3) return fail(WGM_ERR_INVALID);  // scheduler + persistent master + one task workgroup
  rc = app->init_for_real(self, rt, num_wgs);
  if (rc) return fail(rc);
  Args real;
  rc = np->set_args(np->self, real.a);
  if (!rc) rc = app->set_args_for_real(self, real.a);
  if (rc) return fail(rc);
  // the mega-kernel's scheduler cells (its own task buffers go unused)
  const wgm_np_config npc{WGM_NP_REDUCE_MIN, 32, 1024, 0};
  rc = wgm_mega_create(rt, &npc, &m);
  if (rc) return fail(rc);
  rc = wgm_mega_set_query_barrier(m, cfg.use_query_barrier);
  if (rc) return fail(rc);
  const int np_wgs = np_groups_for(cfg.non_persistent_wgs, num_wgs);
  std::vector<double> times, resp, exec;
  for (int it = 0; it < cfg.iterations; it++) {
    rc = app->reset(self, rt);
    if (!rc) rc = np->reset(np->self);
    if (rc) return fail(rc);
    int participating = 0;
    rc = wgm_mega_launch_custom(m, merged_kernel, cfg.threads_per_wg, num_wgs, cfg.lds_bytes, real.a, &participating);
    if (!rc) rc = wgm_mega_send_persistent_custom(m, participating);
    if (rc) return fail(rc);
    if (!cfg.skip_tasks) {
      const int k = std::min(np_wgs, participating - 1);  // the persistent master is never given up
      double begin = now_ms();
      int sent = 0;
      while (wgm_mega_persistent_running(m) && (cfg.max_tasks <= 0 || sent < cfg.max_tasks)) {
        if (now_ms() - begin >= cfg.non_persistent_frequency_ms) {
          rep.redlines++;
        } else {
          while (now_ms() - begin < cfg.non_persistent_frequency_ms) {
          }
        }
        begin = now_ms();
        rc = np->reset(np->self);
        if (rc) return fail(rc);
        wgm_task_times tt{};
        rc = wgm_mega_send_task_custom(m, k, &tt);
        if (rc) return fail(rc);
        resp.push_back(tt.response_us);
        exec.push_back(tt.execution_us);
        rc = np->check(np->self);
        if (rc == WGM_ERR_MISMATCH) rep.np_errors++;
        else if (rc) return fail(rc);
        sent++;
      }
      rep.tasks += sent;
    }
    double ms = 0;
    rc = wgm_mega_wait_persistent(m, 120.0, &ms);
    if (!rc) rc = wgm_mega_quit(m);
    if (rc) return fail(rc);
    times.push_back(ms);
    rc = app->check(self, rt);
    if (rc == WGM_ERR_MISMATCH) rep.persistent_errors++;
    else if (rc) return fail(rc);
  }
  (void)wgm_mega_destroy(m);
  m = nullptr;
  rc = app->clean(self, rt);
  const int rc2 = np->clean(np->self, rt);
  rep.occupancy = occupancy;
  rep.num_wgs = num_wgs;
  rep.np_wgs = np_wgs;
  rep.iterations = cfg.iterations;
  summarize(times, &rep);
  auto pct = [](std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    return v[(size_t)std::floor(q * (double)(v.size() - 1))];
  };
  rep.response_p50_us = pct(resp, 0.5);
  rep.response_p99_us = pct(resp, 0.99);
  rep.exec_p50_us = pct(exec, 0.5);
  rep.exec_p99_us = pct(exec, 0.99);
  if (out) *out = rep;
  return rc ? rc : rc2;
}

}  // extern "C"
