/*
 * wgm.h -- C ABI of the MI355X cooperative workgroup runtime (libwgm.so).
 *
 * Every entry point returns an int status (WGM_OK or a negative WGM_ERR_*)
 * and never calls exit(). The reference instead calls check_ocl() -> exit(1)
 * on any OpenCL error (src/scheduler_rt/rt_host/ggc_ocl.h:11-18) and exits when
 * too many workgroups are requested (src/scheduler_rt/rt_host/cl_communicator.h:125-130).
 *
 * Threading: one wgm_runtime per GPU; calls on one runtime are serialised by
 * the caller (the reference's CL_Communicator is not re-entrant either,
 * src/scheduler_rt/rt_host/cl_communicator.h:19-465).
 */
#ifndef WGM_H
#define WGM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define WGM_OK 0
#define WGM_ERR_INVALID (-1)          /* bad argument or shape */
#define WGM_ERR_NOMEM (-2)            /* host or device allocation failed */
#define WGM_ERR_HIP (-3)              /* a HIP call failed; see wgm_last_error */
#define WGM_ERR_NODEV (-4)            /* no usable GPU */
#define WGM_ERR_TOO_MANY_GROUPS (-5)  /* more workgroups asked for than discovered */
#define WGM_ERR_MISMATCH (-6)         /* a result self-check failed */
#define WGM_ERR_TIMEOUT (-7)          /* a bounded host wait expired */
#define WGM_ERR_STATE (-8)            /* call not valid in the current protocol state */
#define WGM_ERR_BUDGET (-9)           /* admission refused: HBM budget exceeded */
#define WGM_ERR_IO (-10)              /* a file could not be opened, read or written */
#define WGM_ERR_FORMAT (-11)          /* a graph or solution file is malformed; see wgm_format_error */

/* ------------------------------------------------------- graphs (host CSR) */
/* 0-based CSR: row_ptr[n+1], col[m] ascending per row, optional wt[m]. */
typedef struct wgm_csr {
  int32_t n;
  int64_t m;
  int32_t* row_ptr;
  int32_t* col;
  int32_t* wt;
} wgm_csr;

#define WGM_GRAPH_RMAT 0
#define WGM_GRAPH_GRID 1
#define WGM_QUERY_BFS 0
#define WGM_QUERY_SSSP 1

typedef struct wgm_batch_item {
  int32_t index;
  int32_t kind;   /* WGM_GRAPH_* */
  int32_t query;  /* WGM_QUERY_* */
  int32_t n;      /* vertices after truncation / squaring */
  int32_t scale;  /* RMAT scale (0 for grids) */
  int32_t rows, cols;
  uint64_t graph_seed;
  uint64_t weight_seed;
} wgm_batch_item;

/* Synthetic inputs. No counterpart in the reference, which reads Galois .gr
 * / DIMACS / METIS files (src/lonestar/graph_apps/bfs/bfs.h:95-169,
 * src/pannotia/graph_parser/parse.cpp:58-281).
 * splitmix64 (Vigna): state += 0x9e3779b97f4a7c15, then the 64-bit mixer. */
uint64_t wgm_splitmix64_next(uint64_t* state);
/* G(n, M): the first `pairs` distinct non-loop pairs, symmetrised. */
int wgm_gen_er(int32_t n, int64_t pairs, uint64_t seed, wgm_csr* out);
/* RMAT, edge_factor * 2^scale draws, self-loops and duplicates removed,
 * edges with an endpoint >= truncate_n dropped (truncate_n <= 0: keep all). */
int wgm_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                 uint64_t seed, int32_t truncate_n, wgm_csr* out);
/* rows x cols 4-neighbour grid, id = r * cols + c, undirected. */
int wgm_gen_grid(int32_t rows, int32_t cols, wgm_csr* out);
/* wt[e] uniform in [lo, hi] in CSR order. Allocates g->wt. */
int wgm_gen_weights(wgm_csr* g, int32_t lo, int32_t hi, uint64_t seed);
/* uint32 priorities, no two adjacent vertices equal. */
int wgm_gen_priorities(const wgm_csr* g, uint64_t seed, uint32_t* prio);
void wgm_csr_free(wgm_csr* g);
/* free() for single arrays the library allocated (e.g. wgm_gen_weights' wt). */
void wgm_free(void* p);
/* ------------------------------------------------------------ file formats
 * The reference's datasets and solution files, read into (and written from)
 * the 0-based CSR above. Readers allocate out's arrays (wgm_csr_free); out->wt
 * is NULL when the file carries no weights. Every reader builds a complete
 * row_ptr and keeps each row's edges in file order.
 *   WGM_FMT_GR      Galois .gr v1 binary, 4-byte or no edge data
 *                   (readFromGR, src/lonestar/graph_apps/bfs/bfs.h:95-169)
 *   WGM_FMT_DIMACS  "p sp n m" / "a u v w", 1-based
 *                   (parseCOO, src/pannotia/graph_parser/parse.cpp:177-281)
 *   WGM_FMT_METIS   "n m [fmt]" then one adjacency line per vertex, fmt 0 or 1
 *                   (parseMetis, parse.cpp:58-175)
 *   WGM_FMT_MM      MatrixMarket coordinate, self-loops dropped
 *                   (parseMM, parse.cpp:467-580)
 *   WGM_FMT_EDGES   "n m" then "src dst [w]" lines, 0-based
 *                   (readFromEdges, src/lonestar/graph_apps/sssp/sssp.h:49-94)
 * Flags: WGM_READ_UNDIRECTED adds the reverse of every arc (the reference's
 * directed = false; for METIS it means each edge is listed at both ends and
 * m counts edges once); WGM_READ_TRANSPOSE gives the transpose
 * (parseCOO_transpose, parse.cpp:692-797). */
#define WGM_FMT_GR 0
#define WGM_FMT_DIMACS 1
#define WGM_FMT_METIS 2
#define WGM_FMT_MM 3
#define WGM_FMT_EDGES 4
#define WGM_READ_UNDIRECTED 1
#define WGM_READ_TRANSPOSE 2
int wgm_read_graph(const char* path, int32_t format, int32_t flags, wgm_csr* out);
/* Writers for .gr, DIMACS (every CSR edge as an arc) and METIS (flags
 * WGM_READ_UNDIRECTED: m written as half the CSR edges). */
int wgm_write_graph(const char* path, int32_t format, int32_t flags, const wgm_csr* g);
/* Message of the last failed format / solution call on this thread. */
const char* wgm_format_error(void);

/* Solution files, one vertex per line:
 *   WGM_SOL_COLON        "%d:%d" index from 0 (Lonestar, bfs.h:336-365)
 *   WGM_SOL_COLON_SPACE  "%d: %d" index from 1 (Pannotia color / SSSP,
 *                        src/pannotia/graph_apps/color/color.h:149-190)
 *   WGM_SOL_PLAIN        "%d" (Pannotia MIS, mis.h:170-195)
 * wgm_check_solution is diff_solution_file_int: the first n values must
 * match; WGM_ERR_MISMATCH (or WGM_ERR_FORMAT for a short or malformed file)
 * with *first_bad = the vertex. The _f32 pair is BC's "%f" file compared
 * within an absolute tolerance (bc.h:176-203 uses 8). */
#define WGM_SOL_COLON 0
#define WGM_SOL_COLON_SPACE 1
#define WGM_SOL_PLAIN 2
int wgm_write_solution(const char* path, int32_t style, const int32_t* values, int32_t n);
int wgm_check_solution(const char* path, int32_t style, const int32_t* values, int32_t n, int32_t* first_bad);
int wgm_write_solution_f32(const char* path, const float* values, int32_t n);
int wgm_check_solution_f32(const char* path, const float* values, int32_t n, float tol, int32_t* first_bad);

/* Batched workload description (count items) and parallel generation. */
int wgm_batch_spec(int32_t count, uint64_t seed, int32_t lo_exp, int32_t hi_exp,
                   wgm_batch_item* items);
int wgm_batch_generate(int32_t count, const wgm_batch_item* items, wgm_csr* graphs,
                       int32_t threads);

/* --------------------------------------------------------------- runtime */
typedef struct wgm_runtime wgm_runtime;

typedef struct wgm_device_info {
  char name[64];
  char arch[32];
  int32_t num_cus;
  int32_t max_wg_size;
  int32_t wavefront_size;
  int64_t wall_clock_hz;   /* rate of the device constant clock */
  int64_t total_mem;
  int32_t lds_per_cu;
} wgm_device_info;

/* Kernels whose occupancy can be queried / discovered. */
#define WGM_K_DISCOVERY 0
#define WGM_K_BARRIER_BENCH 1
#define WGM_K_BFS 2
#define WGM_K_SSSP 3
#define WGM_K_COLOR 4
#define WGM_K_MIS 5
#define WGM_K_PAGERANK 6
#define WGM_K_MEGA 7
#define WGM_K_BATCHED 8
#define WGM_K_SERVE 9
#define WGM_K_MINPLUS 10
#define WGM_K_BC 11
#define WGM_K_COUNT 12

/* Runtime on one GPU: stream, control block (discovery poll, barrier flag
 * lines, XCD lines) and coherent host cells. Replaces the CL_Execution set-up
 * of src/host_skeleton/skel.cpp:633-700 and the barrier / discovery buffers
 * and their resets (src/host_skeleton/skel.cpp:127-155, 278-295). */
int wgm_create(int32_t device, wgm_runtime** out);
int wgm_destroy(wgm_runtime* rt);
const char* wgm_last_error(const wgm_runtime* rt);
const char* wgm_strerror(int32_t status);
int wgm_get_device_info(wgm_runtime* rt, wgm_device_info* out);

/* hipOccupancyMaxActiveBlocksPerMultiprocessor for `kernel` at wg_size,
 * and the CU count: the vendor bound that discovery is checked against
 * (the reference derates its discovered count instead, amd_check,
 * src/host_skeleton/skel.cpp:242-247). */
int wgm_occupancy_query_hip(wgm_runtime* rt, int32_t kernel, int32_t wg_size,
                            int32_t* blocks_per_cu, int32_t* num_cus);

/* Occupancy discovery probe: launch `probe_groups` workgroups of `kernel`'s
 * resource shape with im_exit set; *count = workgroups that joined the poll.
 * probe_groups <= 0 picks 2x the HIP occupancy bound.
 * Replaces get_occupancy_d_ctx and discovery_get_occupancy
 * (src/host_skeleton/skel.cpp:157-165, 223-240). */
int wgm_discover_occupancy(wgm_runtime* rt, int32_t kernel, int32_t wg_size,
                           int32_t probe_groups, int32_t* count);

/* Barrier round trip at a given occupancy: `iterations` x 10 global
 * barriers in one launch. Replaces the os_freq test (kernel
 * src/test_suite/os_freq/device/os_freq.cl:8-46, host
 * src/test_suite/os_freq/host/os_freq.cpp:262-329). Every wait of the
 * FLAGS / SENSE / XCD variants has a budget of 2^22 spins (the environment
 * variable WGM_BARRIER_SPIN_BUDGET, read at each call, replaces it): a barrier
 * that does not complete ends the launch and returns WGM_ERR_TIMEOUT. The
 * COOP variant's grid sync is HIP's and has no such bound. */
#define WGM_BARRIER_FLAGS 0   /* flat master/slave flag barrier, one line per WG */
#define WGM_BARRIER_SENSE 1   /* counter + sense reversal */
#define WGM_BARRIER_XCD 2     /* two-level: members -> XCD leader, leaders <-> leaders */
#define WGM_BARRIER_COOP 3    /* reference point: hipLaunchCooperativeKernel + cooperative_groups grid sync
                                 (the grid is at most the discovered occupancy) */
typedef struct wgm_barrier_result {
  int32_t participating;
  int32_t barriers;
  int32_t epoch_errors;  /* WGs that saw a wrong epoch after a barrier */
  double kernel_ms;      /* HIP event time of the launch */
  double device_ms;      /* device clock between first and last barrier */
  double ns_per_barrier;
} wgm_barrier_result;
int wgm_barrier_bench(wgm_runtime* rt, int32_t variant, int32_t wg_size, int32_t num_wgs,
                      int32_t iterations, wgm_barrier_result* out);

/* Plain device copy of `bytes` (read + write), `iterations` times: achieved
 * GB/s counts both directions. Calibrates HBM traffic counters (no
 * counterpart in the reference). */
int wgm_copy_bench(wgm_runtime* rt, int64_t bytes, int32_t iterations, double* ms_per_copy, double* gbps);

/* ------------------------------------------------ device-resident graphs */
typedef struct wgm_graph wgm_graph;
/* Upload a CSR (and optional uint32 priorities) to HBM. Replaces the graph
 * part of init_persistent_app_for_real / reset_persistent_task
 * (src/lonestar/graph_apps/bfs/bfs.h:209-313,
 * src/pannotia/graph_apps/color/color.h:41-150); the graph stays resident
 * instead of being re-uploaded per run. */
int wgm_graph_create(wgm_runtime* rt, const wgm_csr* g, const uint32_t* prio, wgm_graph** out);
int wgm_graph_destroy(wgm_graph* g);
int64_t wgm_graph_bytes(const wgm_graph* g);

/* Launch shape of a standalone persistent kernel. */
typedef struct wgm_launch {
  int32_t wg_size;     /* threads per workgroup (64..1024, multiple of 64) */
  int32_t num_wgs;     /* <= 0: discovered occupancy */
  int32_t handshake;   /* 1: host flag handshake as in the reference's run_persistent */
  int32_t mode;        /* wgm_sssp's relaxation direction: WGM_SSSP_AUTO, _PUSH or _PULL */
} wgm_launch;

/* wgm_sssp relaxes in one of two directions (the same distances either way):
 * push: frontier Bellman-Ford in near-far bands, each frontier vertex's
 *   out-edges relaxed with atomic minimums (the reference's drelax2);
 * pull: Bellman-Ford sweeps over the in-edges, each vertex taking the
 *   minimum over its in-neighbours, until a sweep changes nothing.
 * AUTO pulls on graphs with an out-degree above 4, whose frontiers reach
 * most edges within a few levels, and pushes on the others. */
#define WGM_SSSP_AUTO 0
#define WGM_SSSP_PUSH 1
#define WGM_SSSP_PULL 2

typedef struct wgm_run_stats {
  int32_t participating;   /* discovered workgroups */
  int32_t iterations;      /* levels / rounds / iterations executed */
  int64_t edges_scanned;   /* sum of out-degrees of reached vertices (BFS/SSSP) */
  double kernel_ms;        /* HIP event time of the whole launch */
  double device_ms;        /* device clock, first to last barrier (excludes discovery) */
  double host_ms;          /* host handshake flip -> persistent_flag==0 (handshake only) */
} wgm_run_stats;

/* Standalone persistent apps: run_persistent (src/host_skeleton/skel.cpp:250-385)
 * with one plugin each. Outputs are host arrays of n entries (may be NULL to
 * leave the result on the device).
 * wgm_bfs: Lonestar BFS (src/lonestar/graph_apps/bfs/bfs.h:16-370,
 *   src/lonestar/graph_apps/bfs/device/standalone.cl:218-320);
 * wgm_sssp: Lonestar SSSP (src/lonestar/graph_apps/sssp/sssp.h:16-382,
 *   src/lonestar/graph_apps/sssp/device/standalone.cl), same distances: by
 *   default pull sweeps over the in-edges on graphs with an out-degree above
 *   4, and on the others push in near-far bands of width 1024 (lc->mode;
 *   a push on a graph of out-degree above 4 takes bands of width 128);
 *   st->edges_scanned counts the reached vertices' out-edges either way;
 * wgm_color: Pannotia color (src/pannotia/graph_apps/color/color.h:41-195,
 *   src/pannotia/graph_apps/color/device/standalone.cl:10-126);
 * wgm_mis: Pannotia MIS (src/pannotia/graph_apps/mis/mis.h:40-202,
 *   src/pannotia/graph_apps/mis/device/standalone.cl:12-146);
 * wgm_pagerank: no counterpart in the reference (pull PageRank, FP32). */
int wgm_bfs(wgm_runtime* rt, wgm_graph* g, int32_t src, const wgm_launch* lc,
            uint32_t* dist_out, wgm_run_stats* st);
int wgm_sssp(wgm_runtime* rt, wgm_graph* g, int32_t src, const wgm_launch* lc,
             uint32_t* dist_out, wgm_run_stats* st);
int wgm_color(wgm_runtime* rt, wgm_graph* g, const wgm_launch* lc, int32_t* color_out,
              wgm_run_stats* st);
int wgm_mis(wgm_runtime* rt, wgm_graph* g, const wgm_launch* lc, int32_t* s_out,
            wgm_run_stats* st);
int wgm_pagerank(wgm_runtime* rt, wgm_graph* g, float damping, int32_t iterations,
                 const wgm_launch* lc, float* rank_out, wgm_run_stats* st);

/* wgm_bfs on a caller-owned device CSR that allocates nothing. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). row_ptr[n + 1], col[m] and dist[n] are
 * device pointers; the kernel and the results are those of wgm_bfs. The
 * traversal's working memory (counters, two frontier queues, heavy lists)
 * is one static device array of WGM_NOALLOC_SCRATCH_BYTES per GPU, so
 * n <= WGM_NOALLOC_MAX_N and m <= WGM_NOALLOC_MAX_M (WGM_ERR_INVALID beyond).
 * Every call on a GPU uses that GPU's array and leaves its contents behind:
 * calls that overlap on one GPU race. lc->num_wgs <= 0 takes the occupancy
 * this runtime discovered for wgm_bfs at that workgroup size, else the HIP
 * occupancy bound. */
#define WGM_NOALLOC_MAX_N (1 << 21)
#define WGM_NOALLOC_MAX_M (1LL << 26)
#define WGM_NOALLOC_SCRATCH_BYTES 17565568LL
int wgm_bfs_noalloc(wgm_runtime* rt, const int32_t* row_ptr, const int32_t* col, int32_t n, int64_t m, int32_t src,
                    const wgm_launch* lc, uint32_t* dist, wgm_run_stats* st);

/* Pannotia SSSP: run_persistent with the sssp plugin
 * (src/pannotia/graph_apps/sssp/sssp.h:40-176, kernel sssp_combined,
 * src/pannotia/graph_apps/sssp/device/sssp.cl:11-92). Min-plus sweeps over the
 * in-edges (the reference reads its graph transposed, sssp.h:43; here the
 * in-edge CSR is built from the graph on first use and kept) until no
 * distance changes. dist_out: 0 at src, WGM_BIG_NUM (sssp.h:3) where
 * unreached. st->iterations = sweeps, st->edges_scanned = sweeps * m. */
#define WGM_BIG_NUM 99999999
int wgm_sssp_minplus(wgm_runtime* rt, wgm_graph* g, int32_t src, const wgm_launch* lc, int32_t* dist_out,
                     wgm_run_stats* st);
/* Pannotia BC: run_persistent with the bc plugin (src/pannotia/graph_apps/bc/bc.h,
 * kernel bc_combined, src/pannotia/graph_apps/bc/device/bc.cl:186-235).
 * Brandes' betweenness over sources [0, num_sources) (all when <= 0), in
 * FP64 without the reference's x10000 scaling; bc_out[v] = sum over those
 * sources s != v of the dependency of s on v (directed, unnormalised).
 * Deterministic: equal to the same sums in source order on the CPU.
 * st->iterations = BFS levels over all sources, st->edges_scanned =
 * num_sources * m. */
int wgm_bc(wgm_runtime* rt, wgm_graph* g, int32_t num_sources, const wgm_launch* lc, double* bc_out,
           wgm_run_stats* st);

/* ------------------------------------------ resizing scheduler (mega-kernel)
 * One resident mega-kernel: discovered workgroup 0 runs the scheduler, the
 * others serve a persistent task and non-persistent tasks, moving between
 * them at resizing barriers (offer_kill / offer_fork). Host protocol of
 * CL_Communicator (src/scheduler_rt/rt_host/cl_communicator.h:19-465) and
 * the run modes of src/host_skeleton/skel.cpp:168-631. */
typedef struct wgm_mega wgm_mega;

#define WGM_PERSIST_BFS 0
#define WGM_PERSIST_SSSP 1
#define WGM_PERSIST_COLOR 2     /* Pannotia colour (merged color, color/device/merged.cl:73-205) */
#define WGM_PERSIST_MIS 3       /* Pannotia MIS (merged mis, mis/device/merged.cl:75-224) */
#define WGM_PERSIST_PAGERANK 4  /* pull PageRank, FP32 (no counterpart in the reference) */

/* A persistent task for the resizing scheduler. */
typedef struct wgm_persistent_task {
  int32_t kind;        /* WGM_PERSIST_* */
  int32_t src;         /* BFS / SSSP source vertex */
  float damping;       /* PageRank */
  int32_t iterations;  /* PageRank */
} wgm_persistent_task;

#define WGM_NP_MATMULT_I32 0  /* reference matmult: int32, djb2 hash of the diagonal */
#define WGM_NP_MATMUL_F32 1   /* FP32 dim x dim (dim % 32 == 0), FP64 checksum check */
#define WGM_NP_REDUCE_MIN 2   /* int32 minimum of `length` values */

typedef struct wgm_np_config {
  int32_t kind;    /* WGM_NP_* */
  int32_t dim;     /* matrix dimension (default 256) */
  int32_t length;  /* reduction length (default 1 << 20) */
  int32_t seed;    /* input seed (0: default) */
} wgm_np_config;

typedef struct wgm_task_times {
  double response_us;   /* DEVICE_TO_TASK -> DEVICE_GOT_GROUPS */
  double execution_us;  /* DEVICE_TO_EXECUTE -> DEVICE_WAITING */
  int32_t ok;           /* result self-check */
  int32_t reserved;
  /* device-clock breakdown of the execution (resident kernels only) */
  double dev_assign_us;  /* scheduler saw TO_EXECUTE -> every workgroup assigned */
  double dev_start_us;   /* assigned -> first workgroup began the task */
  double dev_body_us;    /* first begin -> last workgroup finished its share */
  double dev_retire_us;  /* last share done -> scheduler saw the task complete */
} wgm_task_times;

typedef struct wgm_mega_stats {
  int32_t participating, tasks, np_groups, errors, redlines, reserved;
  double persistent_ms;
  double response_mean_us, response_p50_us, response_p99_us, response_max_us;
  double exec_mean_us, exec_p50_us, exec_p99_us, exec_max_us;
  double total_mean_us, total_p50_us, total_p99_us;
  double avg_persistent_groups;  /* mean of the 1 ms samples of executing persistent workgroups */
} wgm_mega_stats;

/* Prepare a mega-kernel context and the non-persistent tasks' inputs.
 * Replaces mk_init_scheduler_ctx / set_scheduler_args
 * (src/scheduler_rt/rt_host/cl_scheduler.h:7-87) and init_non_persistent_app
 * (src/non_persistent_kernels/matmult/matmult.h:94-164,
 * src/non_persistent_kernels/reduce/reduce.h:18-36). */
int wgm_mega_create(wgm_runtime* rt, const wgm_np_config* np, wgm_mega** out);
int wgm_mega_destroy(wgm_mega* m);
/* Launch the resident mega-kernel (num_wgs <= 0: discovered occupancy) and
 * wait until the scheduler is ready; *participating excludes the scheduler.
 * Replaces CL_Communicator::launch_mega_kernel / number_of_discovered_groups
 * (src/scheduler_rt/rt_host/cl_communicator.h:71-99). */
int wgm_mega_launch(wgm_mega* m, int32_t wg_size, int32_t num_wgs, int32_t query_barrier, int32_t* participating);
/* Start a persistent task on `groups` workgroups (returns once assigned; a
 * monitor thread samples the executing workgroups every 1 ms). Replaces
 * send_persistent_task / monitor_persistent_task
 * (src/scheduler_rt/rt_host/cl_communicator.h:168-224). Every task gives up
 * and regains workgroups at its resizing barriers: BFS / SSSP one per level
 * (merged drelax2, src/lonestar/graph_apps/bfs/device/merged.cl:310-396),
 * colour one per round, MIS two per round, PageRank one per iteration.
 * Colour and MIS need the graph's priorities. */
int wgm_mega_send_persistent_task(wgm_mega* m, wgm_graph* g, const wgm_persistent_task* t, int32_t groups);
/* The same with kind and source (PageRank: damping 0.85, 50 iterations). */
int wgm_mega_send_persistent(wgm_mega* m, wgm_graph* g, int32_t kind, int32_t src, int32_t groups);
/* is_executing_persistent (src/scheduler_rt/rt_host/cl_communicator.h:226-228). */
int wgm_mega_persistent_running(wgm_mega* m);
/* Bounded wait for the persistent task; get_persistent_time
 * (src/scheduler_rt/rt_host/cl_communicator.h:230-239). */
int wgm_mega_wait_persistent(wgm_mega* m, double timeout_s, double* persistent_ms);
/* One non-persistent task on `groups` workgroups, synchronously; checked.
 * The kind is the one given at wgm_mega_create, or `kind` (WGM_NP_*).
 * Replaces send_task_synchronous (src/scheduler_rt/rt_host/cl_communicator.h:123-166)
 * followed by check_non_persistent_task (src/non_persistent_kernels/matmult/matmult.h:190-194). */
int wgm_mega_send_task(wgm_mega* m, int32_t groups, wgm_task_times* out);
int wgm_mega_send_task_kind(wgm_mega* m, int32_t kind, int32_t groups, wgm_task_times* out);
/* DEVICE_TO_QUIT and a bounded wait for the kernel to finish: send_quit_signal
 * (src/scheduler_rt/rt_host/cl_communicator.h:101-105) + the final queue finish. */
int wgm_mega_quit(wgm_mega* m);
/* The summary of print_summary (src/scheduler_rt/rt_host/cl_communicator.h:296-348),
 * plus p50 / p99. */
int wgm_mega_get_stats(wgm_mega* m, wgm_mega_stats* out);
/* Result of the last persistent task (after wgm_mega_quit), n 4-byte values:
 * uint32 distances (BFS / SSSP), int32 colours, int32 MIS membership (2 in
 * the set, 0 not) or FP32 ranks; *levels = levels / rounds / iterations. */
int wgm_mega_result(wgm_mega* m, void* out, int32_t* levels);
/* The non-persistent task launched on its own, `iterations` times:
 * run_non_persistent (src/host_skeleton/skel.cpp:168-221). */
int wgm_np_run_solo(wgm_mega* m, int32_t wg_size, int32_t num_wgs, int32_t iterations, double* mean_us,
                    int32_t* errors);

/* The non-persistent tasks' resident buffers, copied to the host (after
 * wgm_mega_quit; `bytes` must be the buffer's size): WGM_NP_IN_A is A (int32
 * or FP32, dim x dim) or the reduction's `length` int32 values, WGM_NP_IN_B is
 * B, WGM_NP_OUT_C the last product C. The reference's host keeps the same
 * buffers to check a task (matmult.h:120-149, reduce.h:18-35). */
#define WGM_NP_IN_A 0
#define WGM_NP_IN_B 1
#define WGM_NP_OUT_C 2
int wgm_np_copy(wgm_mega* m, int32_t kind, int32_t which, void* host, int64_t bytes);
/* The summary the last task published to the host (check_non_persistent_task,
 * matmult.h:190-194): the djb2 hash of C's diagonal (WGM_NP_MATMULT_I32), the
 * minimum (WGM_NP_REDUCE_MIN) or 1 with *checksum = sum_ij C[i][j] u(i) v(j),
 * u(i) = i % 7 + 1, v(j) = j % 5 + 1, in FP64 (WGM_NP_MATMUL_F32). */
int wgm_np_result(wgm_mega* m, int32_t* value, double* checksum);
/* The values the library checks that summary against, computed on the host
 * when the inputs were made (FP32: the checksum u^T A . B v in FP64 and its
 * a-priori FP32 error bound). */
int wgm_np_expected(wgm_mega* m, int32_t kind, int32_t* value, double* checksum, double* tol);

/* --------------------------------------------- the caller's own kernels
 * Persistent kernels written against include/wgm_device_api.h (the
 * reference's scheduler API: kernel_merge/stubs/stubs.h:4-7,
 * src/scheduler_rt/rt_device/iw_barrier.cl, scheduler_1.cl, kernel_ctx.cl) and
 * compiled by the caller for gfx950.
 * `kernel` is the kernel's host-side handle (its __global__ function). */
typedef struct wgm_args wgm_args;
/* Ordered kernel arguments (after the runtime's own first argument), in
 * place of cl::Kernel::setArg with a running index
 * (src/lonestar/graph_apps/bfs/bfs.h:265-283). Each push copies `bytes`. */
int wgm_args_create(wgm_args** out);
int wgm_args_push(wgm_args* a, const void* value, int32_t bytes);
int wgm_args_destroy(wgm_args* a);
/* Standalone persistent kernel: first parameter wgm_kernel_args, then the
 * pushed arguments. num_wgs <= 0: the workgroups a discovery probe of the
 * kernel finds co-resident. handshake: the host flag handshake of
 * run_persistent (src/host_skeleton/skel.cpp:300-345). */
int wgm_launch_standalone(wgm_runtime* rt, const void* kernel, int32_t wg_size, int32_t num_wgs, int32_t lds_bytes,
                          const wgm_args* args, int32_t handshake, wgm_run_stats* st);
/* A caller's merged mega-kernel (first parameter wgm_mega_args, body
 * WGM_MEGA_BODY): launch_mega_kernel (cl_communicator.h:71-95) on the
 * scheduler state of `m`. */
int wgm_mega_launch_custom(wgm_mega* m, const void* kernel, int32_t wg_size, int32_t num_wgs, int32_t lds_bytes,
                           const wgm_args* args, int32_t* participating);
/* send_persistent_task for the kernel's own persistent task
 * (cl_communicator.h:207-224). */
int wgm_mega_send_persistent_custom(wgm_mega* m, int32_t groups);
/* send_task_synchronous for the kernel's own non-persistent task
 * (cl_communicator.h:123-166); the caller checks the task's output. */
int wgm_mega_send_task_custom(wgm_mega* m, int32_t groups, wgm_task_times* out);

/* Workgroups of `kernel` (the caller's standalone kernel, first parameter
 * wgm_kernel_args; or with mega = 1 a merged mega-kernel, first parameter
 * wgm_mega_args) that a discovery probe finds co-resident, launched with
 * `args`: get_occupancy_d_ctx on the app's own kernel (skel.cpp:223-240). */
int wgm_discover_occupancy_custom(wgm_runtime* rt, const void* kernel, int32_t wg_size, int32_t lds_bytes,
                                  const wgm_args* args, int32_t mega, int32_t* count);
/* Resizing barriers of the next wgm_mega_launch_custom in the query form
 * (__global_barrier_resize_query, src/scheduler_rt/rt_device/iw_barrier.cl:229-305):
 * the reference's --use_query_barrier. */
int wgm_mega_set_query_barrier(wgm_mega* m, int32_t query_barrier);

/* ---------------------------------------------------------- host skeleton
 * The reference's driver picks one persistent app and one non-persistent
 * task at compile time (src/host_skeleton/skel.cpp:47-77) and calls their
 * free functions. Here the two plugin interfaces are vtables of the same
 * entry points, each taking the plugin's own state; kernel arguments go
 * through the ordered wgm_args builder instead of cl::Kernel::setArg with a
 * running index, and every entry point returns a status. */
typedef struct wgm_persistent_app {
  void* self;
  /* persistent_app_name (src/lonestar/graph_apps/bfs/bfs.h:192-194) */
  const char* (*name)(void* self);
  /* persistent_kernel_name (bfs.h:196-198): the host handle of the
   * standalone kernel (first parameter wgm_kernel_args) */
  const void* (*kernel)(void* self);
  /* init_persistent_app_for_occupancy (bfs.h:209-246): read and upload */
  int (*init_for_occupancy)(void* self, wgm_runtime* rt);
  /* init_persistent_app_for_real (bfs.h:248-263): size per-run state for
   * `occupancy` workgroups */
  int (*init_for_real)(void* self, wgm_runtime* rt, int32_t occupancy);
  /* set_persistent_app_args_for_occupancy / _for_real (bfs.h:265-283) */
  int (*set_args_for_occupancy)(void* self, wgm_args* args);
  int (*set_args_for_real)(void* self, wgm_args* args);
  /* reset_persistent_task (bfs.h:284-313) */
  int (*reset)(void* self, wgm_runtime* rt);
  /* check_persistent_task (bfs.h:315-334): WGM_OK, or WGM_ERR_MISMATCH */
  int (*check)(void* self, wgm_runtime* rt);
  /* clean_persistent_task (bfs.h:366-370) */
  int (*clean)(void* self, wgm_runtime* rt);
} wgm_persistent_app;

typedef struct wgm_np_task {
  void* self;
  /* non_persistent_app_name (src/non_persistent_kernels/matmult/matmult.h:26-28) */
  const char* (*name)(void* self);
  /* non_persistent_kernel_name (matmult.h:30-32): the kernel's standalone
   * handle, or NULL (the merged mode runs the task inside the mega-kernel) */
  const void* (*kernel)(void* self);
  /* init_non_persistent_app (matmult.h:94-164) */
  int (*init)(void* self, wgm_runtime* rt);
  /* set_non_persistent_app_args (matmult.h:166-184) */
  int (*set_args)(void* self, wgm_args* args);
  /* reset_non_persistent (matmult.h:186-188); runs while the mega-kernel
   * holds the GPU, so it may touch host memory only */
  int (*reset)(void* self);
  /* check_non_persistent_task (matmult.h:190-194): WGM_OK or WGM_ERR_MISMATCH */
  int (*check)(void* self);
  /* clean_non_persistent_task (matmult.h:196-200) */
  int (*clean)(void* self, wgm_runtime* rt);
} wgm_np_task;

/* The reference's flags for these run modes (skel.cpp:21-44). */
typedef struct wgm_skel_config {
  int32_t threads_per_wg;       /* FLAGS_threads_per_wg (default 128) */
  int32_t num_wgs;              /* FLAGS_num_wgs: at most this many (<= 0: the discovered occupancy) */
  int32_t iterations;           /* FLAGS_run_persistent / FLAGS_merged_iterations (default 1) */
  int32_t non_persistent_wgs;   /* FLAGS_non_persistent_wgs: -1 occupancy - 2, -2 one, k occupancy / k */
  double non_persistent_frequency_ms;  /* FLAGS_non_persistent_frequency: task period */
  int32_t skip_tasks;           /* FLAGS_skip_tasks */
  int32_t use_query_barrier;    /* FLAGS_use_query_barrier */
  int32_t lds_bytes;            /* dynamic LDS of the kernel */
  int32_t max_tasks;            /* cap on tasks per merged iteration (<= 0: none) */
} wgm_skel_config;

typedef struct wgm_skel_report {
  int32_t occupancy;            /* discovered co-resident workgroups */
  int32_t num_wgs;              /* final size: min(num_wgs, occupancy) */
  int32_t np_wgs;               /* workgroups per non-persistent task */
  int32_t iterations;
  int32_t persistent_errors;    /* iterations whose check_persistent_task failed */
  int32_t np_errors;            /* tasks whose check_non_persistent_task failed */
  int32_t tasks;                /* non-persistent tasks sent */
  int32_t redlines;             /* tasks sent late (the period had already passed) */
  double total_ms, mean_ms, min_ms, max_ms, std_ms;  /* persistent task time per iteration */
  double response_p50_us, response_p99_us, exec_p50_us, exec_p99_us;  /* non-persistent tasks (merged) */
} wgm_skel_report;

/* run_persistent (skel.cpp:250-385): init for occupancy, set args, discover
 * the occupancy of the app's kernel, init for real on min(num_wgs,
 * occupancy), set args, then per iteration reset, launch with the host
 * handshake (time = flag flip -> every workgroup retired), check; clean. */
int wgm_run_persistent(wgm_runtime* rt, const wgm_persistent_app* app, const wgm_skel_config* cfg,
                       wgm_skel_report* out);
/* run_merged + execute_merged_iteration (skel.cpp:405-631): the same set-up
 * with the non-persistent task's arguments first, on `merged_kernel` (the
 * kernel_merge output of the two kernels, first parameter wgm_mega_args);
 * per iteration reset both, launch, send the persistent task to every
 * participant, then one non-persistent task per period (redlines counted)
 * with a reset before and a check after each, until the persistent task
 * ends; quit and check. */
int wgm_run_merged(wgm_runtime* rt, const wgm_persistent_app* app, const wgm_np_task* np, const void* merged_kernel,
                   const wgm_skel_config* cfg, wgm_skel_report* out);

/* --------------------------------------------------------- query server
 * Serving (no counterpart in the reference, whose driver sends one
 * non-persistent task at a time beside one persistent task,
 * src/host_skeleton/skel.cpp:405-463): the resident kernel's persistent task
 * is a query server. Graphs stay resident in HBM; BFS / SSSP requests are
 * admitted into request slots while their state fits the HBM budget, are
 * published through a ring in coherent host memory and run concurrently,
 * one workgroup per query. A query yields its workgroup at any level
 * boundary when the scheduler needs workgroups for a non-persistent task and
 * resumes on another; freed workgroups rejoin the server. A request's state,
 * charged to the budget, is 4n + 2 words for BFS and 7n + 4 for SSSP (its
 * near-far queues and flags), each rounded up to 64 words. */
typedef struct wgm_serve wgm_serve;

typedef struct wgm_serve_config {
  int32_t wg_size;       /* default 1024 */
  int32_t num_wgs;       /* <= 0: discovered occupancy */
  int32_t max_inflight;  /* request slots (default 256) */
  int32_t reserve;       /* free workgroups kept for non-persistent tasks (0: 16, < 0: none) */
  int64_t budget_bytes;  /* HBM admission budget: graphs + in-flight state (default 64 GiB) */
  wgm_np_config np;      /* non-persistent task inputs (every kind is resident) */
} wgm_serve_config;

/* One finished query. */
typedef struct wgm_serve_done {
  int64_t id;             /* from wgm_serve_submit */
  int32_t graph, query;   /* graph index, WGM_QUERY_* */
  int32_t src, levels;    /* source vertex, frontier levels */
  int64_t reached;        /* vertices with a finite distance */
  int64_t edges_scanned;  /* sum of out-degrees of reached vertices */
  uint64_t digest;        /* wgm_dist_digest of the distance array */
  double queue_us;        /* submit -> published to the device (admission wait) */
  double service_us;      /* published -> completion seen by the host */
  double device_us;       /* first level start -> end, device clock */
  int32_t preempted;      /* times the query yielded its workgroup */
  int32_t reserved;
} wgm_serve_done;

/* Upload `count` graphs (they count against the budget) and prepare the
 * server; no kernel is launched yet. */
int wgm_serve_create(wgm_runtime* rt, int32_t count, const wgm_csr* graphs, const wgm_serve_config* cfg,
                     wgm_serve** out);
int wgm_serve_destroy(wgm_serve* s);
/* Resident graph bytes (charged to the budget). */
int64_t wgm_serve_graph_bytes(const wgm_serve* s);
/* Memory plan of a server, computed on the host without a GPU: the graph pool
 * and the request-state arena. The arena is max_inflight times the largest
 * request's state, cut to the budget the graphs leave, so graph_bytes +
 * arena_bytes <= budget_bytes; every admitted request's state is carved out of
 * it (first fit) and returned when the request finishes. WGM_ERR_BUDGET if the
 * graphs alone exceed the budget. */
int wgm_serve_plan(int32_t count, const wgm_csr* graphs, const wgm_serve_config* cfg, int64_t* graph_bytes,
                   int64_t* arena_bytes);
/* Bytes of the request-state arena the server allocated (its plan). */
int64_t wgm_serve_arena_bytes(const wgm_serve* s);
/* Launch the resident kernel and start the server on all workgroups but the
 * scheduler and the reserve; *participating excludes the scheduler. */
int wgm_serve_start(wgm_serve* s, int32_t* participating);
/* Queue one request (returns at once). WGM_ERR_BUDGET if its state could
 * never fit the budget next to the resident graphs. */
int wgm_serve_submit(wgm_serve* s, int32_t graph, int32_t query, int32_t src, int64_t* id);
/* Admit queued requests into free slots within the budget, and return up to
 * `max` finished queries. */
int wgm_serve_poll(wgm_serve* s, wgm_serve_done* out, int32_t max, int32_t* count);
/* Requests submitted and not yet returned by wgm_serve_poll. */
int32_t wgm_serve_outstanding(const wgm_serve* s);
/* Bytes of in-flight request state currently charged to the budget. */
int64_t wgm_serve_inflight_bytes(const wgm_serve* s);
/* One non-persistent task (WGM_NP_*) on `groups` workgroups through the
 * scheduler, synchronously; checked. May be called from another thread than
 * submit/poll. */
int wgm_serve_task(wgm_serve* s, int32_t kind, int32_t groups, wgm_task_times* out);
/* Drain outstanding requests (bounded wait), stop the server, quit the kernel. */
int wgm_serve_stop(wgm_serve* s);
/* Order-independent digest of a distance array:
 * sum over v of mix64(((uint64_t)v << 32) | dist[v]) mod 2^64, mix64 being
 * splitmix64's output function. */
uint64_t wgm_dist_digest(const uint32_t* dist, int32_t n);

/* Open-loop load: Poisson arrivals of queries (graph uniform over the pool,
 * BFS or SSSP as in wgm_batch_spec) and non-persistent tasks (FP32 matmul and
 * int32 reduction alternating). */
typedef struct wgm_serve_load {
  double rate_per_s;     /* arrival rate, all classes */
  int32_t requests;      /* arrivals to generate */
  int32_t np_groups;     /* workgroups per non-persistent task */
  double np_fraction;    /* share of arrivals that are non-persistent tasks */
  uint64_t seed;         /* arrival stream (default 6) */
} wgm_serve_load;

typedef struct wgm_serve_report {
  int32_t queries, tasks, task_errors, reserved;
  double wall_s, throughput_rps, gteps;
  double query_p50_us, query_p99_us, query_mean_us;
  double task_p50_us, task_p99_us, task_mean_us;        /* arrival -> done */
  double task_queue_p99_us, task_response_p99_us, task_exec_p99_us;
  double max_inflight_bytes;
} wgm_serve_report;

/* Run one open-loop load to completion. `records` (optional) receives the
 * finished queries (up to load->requests entries) and *nrecords their count. */
int wgm_serve_run_load(wgm_serve* s, const wgm_serve_load* load, wgm_serve_report* rep, wgm_serve_done* records,
                       int32_t* nrecords);

/* ------------------------------------------------------- batched graphs
 * Many BFS / SSSP queries in one persistent launch that pulls work from a
 * device queue (no counterpart in the reference, which runs one graph per
 * process: src/host_skeleton/skel.cpp:250-385). */
typedef struct wgm_batch wgm_batch;

typedef struct wgm_batch_config {
  int32_t wg_size;     /* threads per workgroup, 64..1024 */
  int32_t wgs_per_cu;  /* resident workgroups per CU requested (capped by discovery) */
  int32_t chunk;       /* frontier entries a team workgroup claims at a time */
  int32_t team_edges;  /* edges per team workgroup: team = ceil(m / team_edges), capped at max_team */
  int32_t max_team;    /* largest team */
  int32_t use_graph;   /* 1: capture the launch in a hipGraph and replay it */
  int32_t delta;       /* SSSP near-far band width (0: plain frontier Bellman-Ford) */
  int32_t sssp_team_edges; /* team granularity of SSSP queries (0: team_edges); an SSSP edge
                              costs two to three BFS edges, so its teams can be larger */
  int32_t delta_high_degree; /* SSSP band width on graphs with an out-degree above 4 (0: delta):
                                their few wide levels favour narrow bands (less rework), the
                                long narrow traversals of low-degree graphs wide ones (fewer levels) */
  int32_t far_flags_lds;     /* 1: a low-degree query on one 1024-thread workgroup keeps one bit per vertex
                                in an LDS bitmap (graphs up to 2^18 vertices): SSSP its near-far flags
                                (instead of a word per vertex in HBM), BFS its visited marks (instead of
                                an atomic minimum on the distance per edge) */
  int32_t low_ell;           /* 1: a low-degree query on one workgroup reads its graph from packed rows,
                                4 words per vertex of (target << 8) | weight, built at create for graphs
                                of out-degree <= 4, fewer than 2^24 - 1 vertices and weights in [0, 255];
                                2: as 1, and an SSSP query among them that also has the LDS bitmap
                                (far_flags_lds, near-far bands on) updates its distances with plain loads
                                and stores that stay in its XCD's L2, its candidates folded per target in
                                LDS hash tables first, instead of atomics (performed beyond the L2) */
} wgm_batch_config;

typedef struct wgm_batch_stats {
  double wall_ms;          /* HIP event time around the whole batch */
  double device_ms;        /* device clock, first workgroup start to last exit */
  double gteps;            /* edges_scanned / wall time */
  int64_t edges_scanned;   /* sum over graphs of out-degrees of reached vertices */
  int32_t participating;   /* discovered workgroups (persistent launch) */
  int32_t launches;        /* kernel launches that served the queries */
  int32_t grid;            /* launch grid of the persistent kernel */
  int32_t reserved;
  int64_t edges_relaxed;   /* edges relaxed from live frontier entries (work: BFS = edges_scanned, SSSP >= it) */
} wgm_batch_stats;

/* Upload `count` graphs (host CSR; weights needed for SSSP queries) into
 * pooled HBM arrays. queries[i] is WGM_QUERY_BFS / WGM_QUERY_SSSP; sources
 * may be NULL (vertex 0). */
int wgm_batch_create(wgm_runtime* rt, int32_t count, const wgm_csr* graphs, const int32_t* queries,
                     const int32_t* sources, wgm_batch** out);
int wgm_batch_destroy(wgm_batch* b);
int64_t wgm_batch_bytes(const wgm_batch* b); /* resident graph bytes: CSR, weights, packed rows and edges */
/* One persistent launch serving every query (cfg may be NULL for defaults). */
int wgm_batch_run(wgm_runtime* rt, wgm_batch* b, const wgm_batch_config* cfg, wgm_batch_stats* st);
/* Baseline: one persistent launch per graph, back to back. */
int wgm_batch_run_relaunch(wgm_runtime* rt, wgm_batch* b, const wgm_batch_config* cfg, wgm_batch_stats* st);
/* Baseline: one launch per query of the standalone app kernels of wgm_bfs /
 * wgm_sssp (every discovered workgroup as one team; lc->num_wgs <= 0: the
 * discovered occupancy), back to back, on the batch's resident graphs. */
int wgm_batch_run_standalone(wgm_runtime* rt, wgm_batch* b, const wgm_launch* lc, wgm_batch_stats* st);
/* Result of query `index` of the last run (any output may be NULL). */
int wgm_batch_result(wgm_batch* b, int32_t index, uint32_t* dist_out, int32_t* iterations, int64_t* edges_scanned);
/* Per-query start and end of the last run (wgm_batch_run or _relaunch), in
 * microseconds of the device wall clock from the first query's start; each
 * array holds `count` values in caller order (-1: the query did not run). */
int wgm_batch_job_times(const wgm_batch* b, double* start_us, double* end_us);
/* Search: wg_size {64..1024} x wgs_per_cu {1,2,4,8} from a start suited to
 * mixed batches (team_edges 2^20, sssp_team_edges 786432, bands 512 / 12,
 * LDS bits on, packed rows with L2 distances (low_ell 2)), then one knob at a time at the best so far:
 * chunk {32..4096}; twice over team_edges {2^14..2^22}, sssp_team_edges
 * {0, 2^16..2^20}, the SSSP band {0, 256..4096} and the band of graphs with an
 * out-degree above 4 {0, 4..128}; then far_flags_lds {0,1} x low_ell {0,1,2}.
 * Each candidate runs `reps` times; returns the fastest configuration. */
int wgm_batch_autotune(wgm_runtime* rt, wgm_batch* b, int32_t reps, wgm_batch_config* best,
                       wgm_batch_stats* best_stats);

#ifdef __cplusplus
}
#endif
#endif /* WGM_H */
