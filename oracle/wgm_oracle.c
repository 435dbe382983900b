/*
 * wgm_oracle.c -- CPU restatement of the reference's hot-path algorithms.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for parity tests, smoke() and the
 * bench CPU baseline. Nothing in the product links or calls it.
 *
 * Plain C99 + pthreads. Build: oracle/Makefile -> oracle/_build/liboracle.so.
 */
#define _GNU_SOURCE  /* pthread_setaffinity_np / CPU_SET for the pinned threaded restatement */
#include "wgm_oracle.h"

#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ BFS
 * Hop levels from src; the source has level 0 and unreached vertices keep
 * MYINFINITY = 1,000,000,000 (src/lonestar/graph_apps/bfs/device/bfs.cl:33,
 * src/lonestar/graph_apps/bfs/bfs.h:250-262). The reference's worklist kernel
 * (src/lonestar/graph_apps/bfs/device/bfs.cl:104-302) reaches the same unique
 * fixpoint: dist[v] = iteration of first discovery. */
int or_bfs(int32_t n, const int32_t* row_ptr, const int32_t* col, int32_t src, uint32_t* dist) {
  if (n <= 0 || src < 0 || src >= n) return -1;
  int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  if (!q) return -2;
  for (int32_t v = 0; v < n; v++) dist[v] = OR_INF;
  int32_t head = 0, tail = 0;
  dist[src] = 0;
  q[tail++] = src;
  int levels = 0;
  while (head < tail) {
    int32_t u = q[head++];
    if ((int)dist[u] + 1 > levels) levels = (int)dist[u] + 1;
    for (int32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) {
      int32_t v = col[e];
      if (dist[v] == OR_INF) {
        dist[v] = dist[u] + 1;
        q[tail++] = v;
      }
    }
  }
  free(q);
  return levels;
}

/* ----------------------------------------------------------------- SSSP
 * Dijkstra with a binary heap. Distances are the unique shortest-path
 * fixpoint the reference's frontier Bellman-Ford converges to
 * (src/lonestar/graph_apps/sssp/device/sssp.cl:104-134: alt = dist[src] + wt,
 * atomic_min into dist[dst]); INF = 1e9 as in the reference. */
typedef struct { uint64_t d; int32_t v; } heap_item;

static void heap_push(heap_item* h, int64_t* size, heap_item it) {
  int64_t i = (*size)++;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (h[p].d <= it.d) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = it;
}

static heap_item heap_pop(heap_item* h, int64_t* size) {
  heap_item top = h[0];
  heap_item last = h[--(*size)];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = l + 1, m = i;
    uint64_t md = last.d;
    if (l < *size && h[l].d < md) { m = l; md = h[l].d; }
    if (r < *size && h[r].d < md) { m = r; }
    if (m == i) break;
    h[i] = h[m];
    i = m;
  }
  if (*size > 0) h[i] = last;
  return top;
}

int or_sssp(int32_t n, const int32_t* row_ptr, const int32_t* col, const int32_t* wt, int32_t src,
            uint32_t* dist) {
  if (n <= 0 || src < 0 || src >= n) return -1;
  int64_t m = row_ptr[n];
  uint64_t* d = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
  heap_item* h = (heap_item*)malloc(sizeof(heap_item) * (size_t)(m + 1));
  if (!d || !h) { free(d); free(h); return -2; }
  for (int32_t v = 0; v < n; v++) d[v] = UINT64_MAX;
  int64_t hs = 0;
  d[src] = 0;
  heap_item s0 = {0, src};
  heap_push(h, &hs, s0);
  while (hs > 0) {
    heap_item it = heap_pop(h, &hs);
    if (it.d != d[it.v]) continue;
    for (int32_t e = row_ptr[it.v]; e < row_ptr[it.v + 1]; e++) {
      uint64_t alt = it.d + (uint64_t)(uint32_t)wt[e];
      int32_t v = col[e];
      if (alt < d[v]) {
        d[v] = alt;
        heap_item ni = {alt, v};
        heap_push(h, &hs, ni);
      }
    }
  }
  for (int32_t v = 0; v < n; v++) dist[v] = d[v] == UINT64_MAX ? OR_INF : (uint32_t)d[v];
  free(d);
  free(h);
  return 0;
}

/* Frontier Bellman-Ford in the reference's level structure: level 0 pushes
 * the source, each level relaxes every edge of every queued vertex with
 * dist[u] read at relaxation time, and a vertex whose distance drops is queued
 * once for the next level (src/lonestar/graph_apps/sssp/device/sssp.cl:104-318). */
int or_sssp_frontier(int32_t n, const int32_t* row_ptr, const int32_t* col, const int32_t* wt, int32_t src,
                     uint32_t* dist, int64_t* relaxations) {
  if (n <= 0 || src < 0 || src >= n) return -1;
  int32_t* qa = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* qb = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* stamp = (int32_t*)calloc((size_t)n, sizeof(int32_t));
  if (!qa || !qb || !stamp) { free(qa); free(qb); free(stamp); return -2; }
  for (int32_t v = 0; v < n; v++) dist[v] = OR_INF;
  dist[src] = 0;
  int32_t na = 1, level = 0;
  int64_t relax = 0;
  qa[0] = src;
  while (na > 0) {
    int32_t nb = 0;
    for (int32_t i = 0; i < na; i++) {
      int32_t u = qa[i];
      for (int32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) {
        int32_t v = col[e];
        uint32_t alt = dist[u] + (uint32_t)wt[e];
        relax++;
        if (alt < dist[v]) {
          dist[v] = alt;
          if (stamp[v] < level + 1) { stamp[v] = level + 1; qb[nb++] = v; }
        }
      }
    }
    int32_t* t = qa; qa = qb; qb = t;
    na = nb;
    level++;
  }
  if (relaxations) *relaxations = relax;
  free(qa); free(qb); free(stamp);
  return level;
}

/* ------------------------------------------------------------ colouring
 * Pannotia color_combined (src/pannotia/graph_apps/color/device/color.cl:3-97)
 * restated serially: phase 1 computes, for every uncoloured vertex, the max
 * priority over uncoloured neighbours, skipping the single neighbour of a
 * degree-1 vertex (`start != end - 1`, line 46), and raises the stop flag of
 * this round; phase 2 gives colour c to each uncoloured vertex above that max;
 * the loop ends after the first round whose phase 1 raised nothing. Colours
 * start at -1 (src/pannotia/graph_apps/color/color.h:53-56); priorities are
 * uint32 here, so the -1 sentinel is held in 64 bits. */
int or_color(int32_t n, const int32_t* row_ptr, const int32_t* col, const uint32_t* prio, int32_t* color) {
  if (n <= 0) return -1;
  int64_t* max_d = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  if (!max_d) return -2;
  for (int32_t v = 0; v < n; v++) color[v] = -1;
  int stop1 = 0, stop2 = 0;
  int* write_stop = &stop1;
  int* read_stop = &stop2;
  int graph_color = 1;
  for (;;) {
    for (int32_t i = 0; i < n; i++) {
      if (color[i] != -1) continue;
      int32_t start = row_ptr[i], end = row_ptr[i + 1];
      int64_t maximum = -1;
      for (int32_t e = start; e < end; e++) {
        if (color[col[e]] == -1 && start != end - 1) {
          *write_stop = 1;
          if ((int64_t)prio[col[e]] > maximum) maximum = (int64_t)prio[col[e]];
        }
      }
      max_d[i] = maximum;
    }
    int* swap = read_stop;
    read_stop = write_stop;
    write_stop = swap;
    for (int32_t i = 0; i < n; i++)
      if (color[i] == -1 && (int64_t)prio[i] > max_d[i]) color[i] = graph_color;
    if (*read_stop == 0) break;
    graph_color++;
    *write_stop = 0;
  }
  free(max_d);
  return graph_color;
}

/* ------------------------------------------------------------------ MIS
 * Pannotia mis_combined (src/pannotia/graph_apps/mis/device/mis.cl:5-125)
 * restated serially. mis1: every vertex with c == -1 raises stop and records
 * the min priority over neighbours with c == -1 (sentinel above any uint32);
 * mis2: c == -1 vertices below that min enter the set (s = 2, c = -2) and
 * mark their c == -1 neighbours cu = -2; the loop ends after mis2 of a round
 * whose mis1 raised nothing; mis3 copies cu == -2 into c. Initial state
 * s = 0, c = cu = -1 (src/pannotia/graph_apps/mis/mis.h:55-85). */
int or_mis(int32_t n, const int32_t* row_ptr, const int32_t* col, const uint32_t* prio, int32_t* s) {
  if (n <= 0) return -1;
  int32_t* c = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* cu = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int64_t* min_array = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  if (!c || !cu || !min_array) { free(c); free(cu); free(min_array); return -2; }
  const int64_t big = (int64_t)1 << 32;
  for (int32_t v = 0; v < n; v++) { s[v] = 0; c[v] = -1; cu[v] = -1; min_array[v] = big; }
  int stop = 0, rounds = 0;
  for (;;) {
    rounds++;
    for (int32_t v = 0; v < n; v++) {
      if (c[v] != -1) continue;
      stop = 1;
      int64_t mn = big;
      for (int32_t e = row_ptr[v]; e < row_ptr[v + 1]; e++)
        if (c[col[e]] == -1 && (int64_t)prio[col[e]] < mn) mn = (int64_t)prio[col[e]];
      min_array[v] = mn;
    }
    int local_stop = stop;
    for (int32_t v = 0; v < n; v++) {
      if ((int64_t)prio[v] < min_array[v] && c[v] == -1) {
        s[v] = 2;
        c[v] = -2;
        for (int32_t e = row_ptr[v]; e < row_ptr[v + 1]; e++)
          if (c[col[e]] == -1) cu[col[e]] = -2;
      }
    }
    if (local_stop == 0) break;
    stop = 0;
    for (int32_t v = 0; v < n; v++)
      if (cu[v] == -2) c[v] = cu[v];
  }
  free(c); free(cu); free(min_array);
  return rounds;
}

/* ------------------------------------------------------------- PageRank
 * Not in the reference (no PageRank anywhere under src/). Definition used:
 * r0 = 1/n, r'(v) = (1-d)/n + d * sum_{u in N(v)} r(u)/deg(u), deg-0
 * vertices contribute nothing. The FP32 version repeats the kernel's exact
 * operation order: contrib = r/deg (correctly rounded), CSR-order sum from
 * 0, r = fmaf(d, sum, base) with base = (1-d)/n and r0 = 1/n in FP32. */
int or_pagerank_f64(int32_t n, const int32_t* row_ptr, const int32_t* col, double d, int32_t iters, double* r) {
  if (n <= 0 || iters < 1) return -1;
  double* c0 = (double*)malloc(sizeof(double) * (size_t)n);
  double* c1 = (double*)malloc(sizeof(double) * (size_t)n);
  if (!c0 || !c1) { free(c0); free(c1); return -2; }
  const double base = (1.0 - d) / (double)n, r0 = 1.0 / (double)n;
  for (int32_t v = 0; v < n; v++) {
    int32_t deg = row_ptr[v + 1] - row_ptr[v];
    c0[v] = deg ? r0 / (double)deg : 0.0;
  }
  for (int32_t it = 0; it < iters; it++) {
    const double* cin = (it & 1) ? c1 : c0;
    double* cout = (it & 1) ? c0 : c1;
    for (int32_t v = 0; v < n; v++) {
      double sum = 0.0;
      for (int32_t e = row_ptr[v]; e < row_ptr[v + 1]; e++) sum += cin[col[e]];
      double rv = base + d * sum;
      int32_t deg = row_ptr[v + 1] - row_ptr[v];
      if (it == iters - 1) r[v] = rv;
      else cout[v] = deg ? rv / (double)deg : 0.0;
    }
  }
  free(c0); free(c1);
  return 0;
}

int or_pagerank_f32(int32_t n, const int32_t* row_ptr, const int32_t* col, float d, int32_t iters, float* r) {
  if (n <= 0 || iters < 1) return -1;
  float* c0 = (float*)malloc(sizeof(float) * (size_t)n);
  float* c1 = (float*)malloc(sizeof(float) * (size_t)n);
  if (!c0 || !c1) { free(c0); free(c1); return -2; }
  const float nf = (float)n;
  const float base = (1.0f - d) / nf;
  const float r0 = 1.0f / nf;
  for (int32_t v = 0; v < n; v++) {
    int32_t deg = row_ptr[v + 1] - row_ptr[v];
    c0[v] = deg ? r0 / (float)deg : 0.0f;
  }
  for (int32_t it = 0; it < iters; it++) {
    const float* cin = (it & 1) ? c1 : c0;
    float* cout = (it & 1) ? c0 : c1;
    for (int32_t v = 0; v < n; v++) {
      int32_t beg = row_ptr[v], end = row_ptr[v + 1];
      float sum = 0.0f;
      for (int32_t e = beg; e < end; e++) sum = sum + cin[col[e]];
      float rv = fmaf(d, sum, base);
      if (it == iters - 1) r[v] = rv;
      else cout[v] = (end > beg) ? rv / (float)(end - beg) : 0.0f;
    }
  }
  free(c0); free(c1);
  return 0;
}

/* ---------------------------------------------------- non-persistent tasks
 * matmult (src/non_persistent_kernels/matmult/matmult.h:38-90 and
 * src/non_persistent_kernels/matmult/device/matmult.cl:9-67): middle-square
 * digit fill, int32 product with wrap-around, djb2 hash of the diagonal. */
int32_t or_middle_square_fill(int32_t* m, int32_t rows, int32_t cols, int32_t seed) {
  for (int32_t r = 0; r < rows; r++) {
    for (int32_t c = 0; c < cols; c++) {
      seed = (int32_t)((uint32_t)seed * (uint32_t)seed);
      seed = (seed / 1000) % 1000000;
      m[r * cols + c] = seed % 10;
    }
  }
  return seed;
}

void or_matmul_i32(const int32_t* a, const int32_t* b, int32_t* c, int32_t a_rows, int32_t a_cols, int32_t b_cols) {
  for (int32_t r = 0; r < a_rows; r++) {
    for (int32_t j = 0; j < b_cols; j++) {
      uint32_t acc = 0;
      for (int32_t k = 0; k < a_cols; k++)
        acc += (uint32_t)a[r * a_cols + k] * (uint32_t)b[k * b_cols + j];
      c[r * b_cols + j] = (int32_t)acc;
    }
  }
}

int32_t or_hash_diag(const int32_t* m, int32_t rows, int32_t cols) {
  uint32_t h = 5381;
  for (int32_t i = 0; i < rows && i < cols; i++) h = h * 33u + (uint32_t)m[i * cols + i];
  return (int32_t)h;
}

void or_matmul_f64(const float* a, const float* b, double* c, int32_t m, int32_t k, int32_t n) {
  for (int32_t i = 0; i < m; i++)
    for (int32_t j = 0; j < n; j++) {
      double acc = 0.0;
      for (int32_t t = 0; t < k; t++) acc += (double)a[i * k + t] * (double)b[t * n + j];
      c[i * n + j] = acc;
    }
}

/* MY_reduce (src/non_persistent_kernels/reduce/device/reduce.cl:4-41): min. */
int32_t or_reduce_min(const int32_t* buf, int64_t len) {
  int32_t mn = INT32_MAX;
  for (int64_t i = 0; i < len; i++)
    if (buf[i] < mn) mn = buf[i];
  return mn;
}

/* ------------------------------------------------- threaded CPU baseline
 * Threads play workgroups: each level every thread takes a slice of the
 * frontier, claims vertices with CAS (BFS) or lowers them with a CAS-min loop
 * (SSSP), appends to the next frontier through one fetch-add per thread, and
 * a pthread barrier ends the level -- the reference's round structure with
 * one barrier per level. */
typedef struct {
  int32_t n;
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* wt;
  uint32_t* dist;
  int32_t* stamp;
  int32_t* q[2];
  int32_t count[3];
  int32_t threads;
  pthread_barrier_t bar;
} th_shared;

typedef struct { th_shared* sh; int32_t tid; int32_t cpu; } th_arg;

/* Pinning of the threaded restatement: 0 leaves threads to the scheduler;
 * 1 pins thread t to the t-th CPU of the caller's affinity set. */
static int g_pin = 0;
int or_threads_pin(int32_t on) {
  g_pin = on ? 1 : 0;
  return 0;
}

static void pin_to(int32_t cpu) {
  if (cpu < 0) return;
  cpu_set_t set;
  CPU_ZERO(&set);
  CPU_SET(cpu, &set);
  pthread_setaffinity_np(pthread_self(), sizeof(set), &set);
}

static void* th_frontier(void* p) {
  th_arg* a = (th_arg*)p;
  th_shared* s = a->sh;
  pin_to(a->cpu);
  const int sssp = s->wt != NULL;
  int32_t* local = (int32_t*)malloc(sizeof(int32_t) * 4096);
  int cap = 4096;
  for (int level = 0;; level++) {
    int32_t in_n = __atomic_load_n(&s->count[level % 3], __ATOMIC_ACQUIRE);
    if (in_n == 0) break;
    const int32_t* in_q = s->q[level & 1];
    int32_t* out_q = s->q[(level + 1) & 1];
    int32_t lo = (int32_t)((int64_t)in_n * a->tid / s->threads);
    int32_t hi = (int32_t)((int64_t)in_n * (a->tid + 1) / s->threads);
    int nl = 0;
    for (int32_t i = lo; i < hi; i++) {
      int32_t u = in_q[i];
      uint32_t du = __atomic_load_n(&s->dist[u], __ATOMIC_RELAXED);
      for (int32_t e = s->row_ptr[u]; e < s->row_ptr[u + 1]; e++) {
        int32_t v = s->col[e];
        int push = 0;
        if (!sssp) {
          uint32_t expect = OR_INF;
          if (__atomic_load_n(&s->dist[v], __ATOMIC_RELAXED) == OR_INF &&
              __atomic_compare_exchange_n(&s->dist[v], &expect, (uint32_t)(level + 1), 0, __ATOMIC_RELAXED,
                                          __ATOMIC_RELAXED))
            push = 1;
        } else {
          uint32_t alt = du + (uint32_t)s->wt[e];
          uint32_t cur = __atomic_load_n(&s->dist[v], __ATOMIC_RELAXED);
          int lowered = 0;
          while (alt < cur) {
            if (__atomic_compare_exchange_n(&s->dist[v], &cur, alt, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
              lowered = 1;
              break;
            }
          }
          if (lowered) {
            int32_t st = __atomic_load_n(&s->stamp[v], __ATOMIC_RELAXED);
            while (st < level + 1) {
              if (__atomic_compare_exchange_n(&s->stamp[v], &st, level + 1, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                push = 1;
                break;
              }
            }
          }
        }
        if (push) {
          if (nl == cap) {
            int32_t base = __atomic_fetch_add(&s->count[(level + 1) % 3], nl, __ATOMIC_RELAXED);
            memcpy(out_q + base, local, sizeof(int32_t) * (size_t)nl);
            nl = 0;
          }
          local[nl++] = v;
        }
      }
    }
    if (nl) {
      int32_t base = __atomic_fetch_add(&s->count[(level + 1) % 3], nl, __ATOMIC_RELAXED);
      memcpy(out_q + base, local, sizeof(int32_t) * (size_t)nl);
    }
    if (a->tid == 0) __atomic_store_n(&s->count[(level + 2) % 3], 0, __ATOMIC_RELAXED);
    pthread_barrier_wait(&s->bar);
  }
  free(local);
  return NULL;
}

static int run_threads(int32_t n, const int32_t* row_ptr, const int32_t* col, const int32_t* wt, int32_t src,
                       uint32_t* dist, int32_t threads) {
  if (n <= 0 || src < 0 || src >= n || threads < 1) return -1;
  th_shared s;
  memset(&s, 0, sizeof(s));
  s.n = n; s.row_ptr = row_ptr; s.col = col; s.wt = wt; s.dist = dist; s.threads = threads;
  s.q[0] = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  s.q[1] = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  s.stamp = wt ? (int32_t*)calloc((size_t)n, sizeof(int32_t)) : NULL;
  th_arg* args = (th_arg*)malloc(sizeof(th_arg) * (size_t)threads);
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  if (!s.q[0] || !s.q[1] || (wt && !s.stamp) || !args || !th) {
    free(s.q[0]); free(s.q[1]); free(s.stamp); free(args); free(th);
    return -2;
  }
  for (int32_t v = 0; v < n; v++) dist[v] = OR_INF;
  dist[src] = 0;
  s.q[0][0] = src;
  s.count[0] = 1;
  pthread_barrier_init(&s.bar, NULL, (unsigned)threads);
  /* pinned: thread t on the t-th CPU of the caller's set (the caller's own
   * mask is restored afterwards: thread 0 is the calling thread) */
  cpu_set_t caller;
  int have_mask = g_pin && pthread_getaffinity_np(pthread_self(), sizeof(caller), &caller) == 0;
  int32_t ncpu = have_mask ? CPU_COUNT(&caller) : 0;
  for (int32_t t = 0, c = -1, seen = 0; t < threads; t++) {
    args[t].sh = &s;
    args[t].tid = t;
    args[t].cpu = -1;
    if (ncpu > 0) {
      /* the (t mod ncpu)-th CPU present in the caller's mask */
      int32_t want = t % ncpu;
      if (want < seen) { c = -1; seen = 0; }
      while (seen <= want) {
        c++;
        if (CPU_ISSET(c, &caller)) seen++;
      }
      args[t].cpu = c;
    }
    if (t > 0) pthread_create(&th[t], NULL, th_frontier, &args[t]);
  }
  th_frontier(&args[0]);
  for (int32_t t = 1; t < threads; t++) pthread_join(th[t], NULL);
  if (have_mask) pthread_setaffinity_np(pthread_self(), sizeof(caller), &caller);
  pthread_barrier_destroy(&s.bar);
  free(s.q[0]); free(s.q[1]); free(s.stamp); free(args); free(th);
  return 0;
}

int or_bfs_threads(int32_t n, const int32_t* row_ptr, const int32_t* col, int32_t src, uint32_t* dist,
                   int32_t threads) {
  return run_threads(n, row_ptr, col, NULL, src, dist, threads);
}

int or_sssp_threads(int32_t n, const int32_t* row_ptr, const int32_t* col, const int32_t* wt, int32_t src,
                    uint32_t* dist, int32_t threads) {
  if (!wt) return -1;
  return run_threads(n, row_ptr, col, wt, src, dist, threads);
}

/* ---------------------------------------------------------------- digest
 * FNV-1a, 64-bit: offset basis 0xcbf29ce484222325, prime 0x100000001b3
 * (Fowler/Noll/Vo, published parameters). */
/* ------------------------------------------------------ reverse CSR
 * Counting sort by destination, sources visited in ascending order: the
 * in-edge lists that parseCOO_transpose builds
 * (src/pannotia/graph_parser/parse.cpp:692-797, stable sort by row). */
int or_reverse_csr(int32_t n, const int32_t* row_ptr, const int32_t* col, const int32_t* wt, int32_t* rrow,
                   int32_t* rcol, int32_t* rwt) {
  if (n <= 0) return -1;
  const int32_t m = row_ptr[n];
  memset(rrow, 0, sizeof(int32_t) * ((size_t)n + 1));
  for (int32_t e = 0; e < m; e++) rrow[col[e] + 1]++;
  for (int32_t v = 0; v < n; v++) rrow[v + 1] += rrow[v];
  int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  if (!pos) return -2;
  memcpy(pos, rrow, sizeof(int32_t) * (size_t)n);
  for (int32_t u = 0; u < n; u++)
    for (int32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) {
      const int32_t p = pos[col[e]]++;
      rcol[p] = u;
      if (rwt && wt) rwt[p] = wt[e];
    }
  free(pos);
  return 0;
}

/* ---------------------------------------------------- Pannotia SSSP
 * sssp_combined (src/pannotia/graph_apps/sssp/device/sssp.cl:11-92): per
 * iteration x = y; y[v] = min(x[v], min_j data[j] + x[col[j]]) over the
 * in-edge row of v (the graph is read transposed, sssp.h:43); stop when y
 * equals x. Start values: 0 at the source, BIG_NUM elsewhere (sssp.h:3,
 * 61-68). */
int or_sssp_minplus(int32_t n, const int32_t* rrow, const int32_t* rcol, const int32_t* rwt, int32_t src,
                    int32_t* dist) {
  if (n <= 0 || src < 0 || src >= n) return -1;
  int32_t* x = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  if (!x) return -2;
  for (int32_t v = 0; v < n; v++) dist[v] = v == src ? 0 : OR_BIG_NUM;
  int sweeps = 0;
  for (;;) {
    memcpy(x, dist, sizeof(int32_t) * (size_t)n);
    int changed = 0;
    for (int32_t v = 0; v < n; v++) {
      int32_t mn = x[v];
      for (int32_t j = rrow[v]; j < rrow[v + 1]; j++) {
        const int32_t c = rwt[j] + x[rcol[j]];
        if (c < mn) mn = c;
      }
      dist[v] = mn;
      changed |= mn != x[v];
    }
    sweeps++;
    if (!changed) break;
  }
  free(x);
  return sweeps;
}

/* ------------------------------------------------------- Pannotia BC
 * bc_combined (src/pannotia/graph_apps/bc/device/bc.cl:186-235) for each
 * source s: level-synchronous BFS accumulating path counts
 * (mega_bfs_kernel_func, bc.cl:17-104), then the backward sweep
 * accumulating dependencies and adding them to bc (mega_backtrack_kernel_func,
 * bc.cl:106-183), which here pulls over out-edges instead of pushing over
 * the reverse graph. The reference scales rho / sigma by 10000 in FP32;
 * this is the unscaled FP64 form of the same sums:
 *   sigma[w] = sum over in-edges (u, w) with d[u] = d[w] - 1 of sigma[u]
 *   delta[v] = sum over out-edges (v, w) with d[w] = d[v] + 1 of
 *              sigma[v] / sigma[w] * (1 + delta[w])
 *   bc[v]   += delta[v] for v != s. */
int64_t or_bc(int32_t n, const int32_t* row_ptr, const int32_t* col, const int32_t* rrow, const int32_t* rcol,
              int32_t num_sources, double* bc) {
  if (n <= 0 || num_sources < 0 || num_sources > n) return -1;
  int32_t* d = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* lvl = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 2));
  double* sig = (double*)malloc(sizeof(double) * (size_t)n);
  double* del = (double*)malloc(sizeof(double) * (size_t)n);
  if (!d || !q || !lvl || !sig || !del) {
    free(d), free(q), free(lvl), free(sig), free(del);
    return -2;
  }
  for (int32_t v = 0; v < n; v++) bc[v] = 0.0;
  int64_t levels_total = 0;
  for (int32_t s = 0; s < num_sources; s++) {
    for (int32_t v = 0; v < n; v++) {
      d[v] = -1;
      del[v] = 0.0;
    }
    d[s] = 0;
    sig[s] = 1.0;
    q[0] = s;
    lvl[0] = 0;
    lvl[1] = 1;
    int32_t L = 0, tail = 1;
    while (lvl[L] < lvl[L + 1]) {
      for (int32_t i = lvl[L]; i < lvl[L + 1]; i++) {
        const int32_t v = q[i];
        for (int32_t e = row_ptr[v]; e < row_ptr[v + 1]; e++)
          if (d[col[e]] < 0) {
            d[col[e]] = L + 1;
            q[tail++] = col[e];
          }
      }
      lvl[L + 2] = tail;
      for (int32_t i = lvl[L + 1]; i < tail; i++) {
        const int32_t w = q[i];
        double acc = 0.0;
        for (int32_t j = rrow[w]; j < rrow[w + 1]; j++)
          if (d[rcol[j]] == L) acc += sig[rcol[j]];
        sig[w] = acc;
      }
      L++;
    }
    levels_total += L;
    for (int32_t l = L - 1; l >= 0; l--)
      for (int32_t i = lvl[l]; i < lvl[l + 1]; i++) {
        const int32_t v = q[i];
        double acc = 0.0;
        for (int32_t e = row_ptr[v]; e < row_ptr[v + 1]; e++) {
          const int32_t w = col[e];
          if (d[w] == l + 1) acc += sig[v] / sig[w] * (1.0 + del[w]);
        }
        del[v] = acc;
      }
    for (int32_t v = 0; v < n; v++)
      if (v != s) bc[v] += del[v];
  }
  free(d), free(q), free(lvl), free(sig), free(del);
  return levels_total;
}

uint64_t or_fnv1a64(const void* data, int64_t len) {
  const unsigned char* p = (const unsigned char*)data;
  uint64_t h = 0xcbf29ce484222325ull;
  for (int64_t i = 0; i < len; i++) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}
