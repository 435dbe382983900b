// noalloc.hip -- standalone BFS on caller-owned device buffers that
// allocates nothing: the entry point for callers that cannot call hipMalloc
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// it). No counterpart in the reference.
//
// The traversal's working memory -- level, claim and heavy-list counters,
// the level count and reached-edge sum, two frontier queues and the heavy
// lists -- lives in one file-scope device array per GPU, sized for graphs of
// up to WGM_NOALLOC_MAX_N vertices and WGM_NOALLOC_MAX_M edges. Its device
// address is looked up once per device when a runtime is created (the host
// symbol is only a stand-in for it) and passed to the kernels as a pointer.
// Every call on one GPU reuses the array: the kernels initialise what they
// read, and calls that overlap on one GPU race.

#include "apps_internal.hpp"
#include "team_frontier.hpp"

#include <mutex>

using namespace wgm;

namespace {

constexpr size_t kLineBytes = kLine * sizeof(int);
constexpr size_t kHeavyCap = (size_t)(WGM_NOALLOC_MAX_M / kHeavyMin) + 2;  // heavy-list entries
// layout: 11 counter lines (level counts 3, claims 3, heavy-list counts 3,
// the level count 1, the reached-edge sum 1), two queues of n + 1 entries,
// three heavy lists of kHeavyCap entries
constexpr size_t kCounterLines = 11;
constexpr size_t kQueueOff = kCounterLines * kLineBytes;
constexpr size_t kHeavyOff = kQueueOff + ((2 * ((size_t)WGM_NOALLOC_MAX_N + 1) * 4 + 255) / 256) * 256;
constexpr size_t kScratchBytes = kHeavyOff + ((3 * kHeavyCap * 4 + 255) / 256) * 256;
static_assert(kScratchBytes == WGM_NOALLOC_SCRATCH_BYTES, "scratch size and its declared maximum disagree");

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace wgm {
__global__ void k_noalloc_clear(unsigned long long* cell) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *cell = 0ull;
}
}  // namespace wgm

// Called by wgm_create with `device` current: the device address of this
// GPU's copy of the array.
int wgm_noalloc_init(wgm_runtime* rt, int device) {
  if (device < 0 || device >= kMaxDevices) return wgm_fail(rt, WGM_ERR_INVALID, "device %d beyond the scratch table", device);
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[device]) return WGM_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return wgm_hip_fail(rt, e, "hipGetSymbolAddress(scratch)");
  g_scratch[device] = (unsigned char*)p;
  return WGM_OK;
}

extern "C" {

int wgm_bfs_noalloc(wgm_runtime* rt, const int32_t* row_ptr, const int32_t* col, int32_t n, int64_t m, int32_t src,
                    const wgm_launch* lc, uint32_t* dist, wgm_run_stats* st) {
  if (!rt || !row_ptr || !dist || (m > 0 && !col)) return wgm_fail(rt, WGM_ERR_INVALID, "NULL argument");
  if (n < 1 || n > WGM_NOALLOC_MAX_N || m < 0 || m > WGM_NOALLOC_MAX_M)
    return wgm_fail(rt, WGM_ERR_INVALID, "graph of %d vertices and %lld edges exceeds the scratch limits (%d, %lld)", n,
                    (long long)m, WGM_NOALLOC_MAX_N, (long long)WGM_NOALLOC_MAX_M);
  if (src < 0 || src >= n) return wgm_fail(rt, WGM_ERR_INVALID, "source %d out of range", src);
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    base = rt->device < kMaxDevices ? g_scratch[rt->device] : nullptr;
  }
  if (!base) return wgm_fail(rt, WGM_ERR_STATE, "scratch of device %d not set up", rt->device);
  const int wg = lc && lc->wg_size > 0 ? lc->wg_size : 256;
  int rc = wgm_check_wg(rt, wg);
  if (rc) return rc;
  // grid: the caller's, else an occupancy this runtime discovered earlier for
  // the BFS kernel at this size, else the HIP occupancy bound (discovery then
  // drops the workgroups that are not co-resident); no cache is filled here
  int grid = lc && lc->num_wgs > 0 ? lc->num_wgs : 0;
  if (grid <= 0) {
    auto it = rt->disc_cache.find(WGM_K_BFS * 17 + wg / 64);
    if (it != rt->disc_cache.end()) {
      grid = it->second;
    } else {
      int per_cu = 0;
      WGM_HIPCHK(rt, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wgm_kernel_info(WGM_K_BFS).fn, wg,
                                                                  wgm_kernel_lds(WGM_K_BFS, wg)));
      grid = per_cu * rt->info.num_cus;
    }
  }
  if (grid < 1 || grid > rt->max_grid) return wgm_fail(rt, WGM_ERR_INVALID, "grid %d out of range", grid);
  int* lines = (int*)base;
  FrontierDev s{};
  s.dist = dist;
  s.counts = lines;
  s.claims = lines + 3 * kLine;
  s.hcnt = lines + 6 * kLine;
  int* iters = lines + 9 * kLine;
  unsigned long long* reached = (unsigned long long*)(lines + 10 * kLine);
  s.q0 = (int32_t*)(base + kQueueOff);
  s.q1 = s.q0 + n + 1;
  s.hv = (int32_t*)(base + kHeavyOff);
  s.hdu = (uint32_t*)(s.hv + kHeavyCap);
  s.hclaim = (int32_t*)(s.hdu + kHeavyCap);
  GraphDev g{};
  g.row_ptr = row_ptr;
  g.col = col;
  g.n = n;
  g.m = (int32_t)m;
  wgm_run_stats local{};
  rc = wgm_launch_frontier(rt, g, s, NearFar{}, false, src, grid, wg, lc ? lc->handshake : 0, iters, &local);
  if (rc) return rc;
  hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(64), 0, rt->stream, reached);
  WGM_HIPCHK(rt, hipGetLastError());
  rc = wgm_enqueue_reached_edges(rt, g, dist, reached);
  if (rc) return rc;
  unsigned long long edges = 0;
  WGM_HIPCHK(rt, hipMemcpyAsync(&edges, reached, sizeof(edges), hipMemcpyDeviceToHost, rt->stream));
  WGM_HIPCHK(rt, hipStreamSynchronize(rt->stream));
  local.edges_scanned = (int64_t)edges;
  if (st) *st = local;
  return WGM_OK;
}

}  // extern "C"
