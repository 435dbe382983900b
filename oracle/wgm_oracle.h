/*
 * wgm_oracle.h -- CPU restatement of the reference's hot-path algorithms.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for parity tests,
 * the smoke check and the CPU baseline of bench.py. No product code path
 * links, loads or calls it.
 *
 * Each function names the reference file:line whose semantics it restates.
 * How each one is pinned (closed forms, library routines, reference KATs) is
 * listed beside its tests.
 */
#ifndef WGM_ORACLE_H
#define WGM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_INF 1000000000u

/* Serial textbook BFS: hop levels from src, OR_INF for unreached. */
int or_bfs(int32_t n, const int32_t* row_ptr, const int32_t* col, int32_t src, uint32_t* dist);
/* Dijkstra (binary heap, 64-bit sums): shortest distances, OR_INF unreached. */
int or_sssp(int32_t n, const int32_t* row_ptr, const int32_t* col, const int32_t* wt, int32_t src,
            uint32_t* dist);
/* Frontier Bellman-Ford with the reference's level structure; returns the
 * number of levels and the relaxation count. */
int or_sssp_frontier(int32_t n, const int32_t* row_ptr, const int32_t* col, const int32_t* wt, int32_t src,
                     uint32_t* dist, int64_t* relaxations);
/* Pannotia colouring, two phases per round; returns rounds, colour[v] >= 1. */
int or_color(int32_t n, const int32_t* row_ptr, const int32_t* col, const uint32_t* prio, int32_t* color);
/* Pannotia MIS, three phases per round; returns rounds, s[v] in {0, 2}. */
int or_mis(int32_t n, const int32_t* row_ptr, const int32_t* col, const uint32_t* prio, int32_t* s);
/* PageRank r' = (1-d)/n + d * sum r(u)/deg(u): FP64, and FP32 with the
 * kernel's operation order. */
int or_pagerank_f64(int32_t n, const int32_t* row_ptr, const int32_t* col, double d, int32_t iters, double* r);
int or_pagerank_f32(int32_t n, const int32_t* row_ptr, const int32_t* col, float d, int32_t iters, float* r);
/* matmult non-persistent task: middle-square fill, int32 product, djb2 diagonal hash. */
int32_t or_middle_square_fill(int32_t* m, int32_t rows, int32_t cols, int32_t seed);
void or_matmul_i32(const int32_t* a, const int32_t* b, int32_t* c, int32_t a_rows, int32_t a_cols, int32_t b_cols);
int32_t or_hash_diag(const int32_t* m, int32_t rows, int32_t cols);
/* FP32 inputs, FP64 accumulation reference product. */
void or_matmul_f64(const float* a, const float* b, double* c, int32_t m, int32_t k, int32_t n);
/* reduce non-persistent task: minimum. */
int32_t or_reduce_min(const int32_t* buf, int64_t len);
/* Multi-threaded frontier BFS / Bellman-Ford (threads play workgroups,
 * one barrier per level); the CPU baseline. */
/* Pin the threaded restatement's threads (thread t on the t-th CPU of the
 * caller's affinity set) when on != 0; off by default. */
int or_threads_pin(int32_t on);
int or_bfs_threads(int32_t n, const int32_t* row_ptr, const int32_t* col, int32_t src, uint32_t* dist,
                   int32_t threads);
int or_sssp_threads(int32_t n, const int32_t* row_ptr, const int32_t* col, const int32_t* wt, int32_t src,
                    uint32_t* dist, int32_t threads);

/* In-edge (transposed) CSR: the in-edges of w listed by ascending source,
 * each source's edges in CSR order; rwt may be NULL (then wt is ignored). */
int or_reverse_csr(int32_t n, const int32_t* row_ptr, const int32_t* col, const int32_t* wt, int32_t* rrow,
                   int32_t* rcol, int32_t* rwt);
/* Pannotia SSSP: Jacobi min-plus sweeps over in-edges until nothing changes;
 * x[src] = 0, others start at OR_BIG_NUM. Returns the sweeps, including the
 * last one that changed nothing. */
#define OR_BIG_NUM 99999999
int or_sssp_minplus(int32_t n, const int32_t* rrow, const int32_t* rcol, const int32_t* rwt, int32_t src,
                    int32_t* dist);
/* Pannotia BC (Brandes) in FP64 over sources [0, num_sources): levels by
 * BFS, path counts pulled over in-edges (reverse-CSR order), dependencies
 * pulled over out-edges (CSR order), bc[v] += delta_s(v) for v != s with
 * sources ascending. Returns the total number of BFS levels. */
int64_t or_bc(int32_t n, const int32_t* row_ptr, const int32_t* col, const int32_t* rrow, const int32_t* rcol,
              int32_t num_sources, double* bc);

/* FNV-1a 64-bit digest of a byte range (fixture hashes). */
uint64_t or_fnv1a64(const void* data, int64_t len);

#ifdef __cplusplus
}
#endif
#endif
