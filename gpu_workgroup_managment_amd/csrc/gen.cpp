// Seeded synthetic graph generators: Erdos-Renyi G(n, M), RMAT, 2-D grids,
// edge weights and tie-free vertex priorities, all producing 0-based CSR
// (int32 row_ptr[n+1], int32 col[m], neighbours ascending).
//
// The reference reads graphs from Galois .gr / METIS files instead
// (src/lonestar/graph_apps/bfs/bfs.h:95-169, src/pannotia/graph_parser/parse.cpp:58-175);
// those files are external downloads, so every workload here is generated.
// One deliberate difference from the Lonestar loader: it stores psrc[n] = nedges
// on a 1-based edge array (src/lonestar/graph_apps/bfs/bfs.h:41-43), which makes
// the kernels drop the last out-edge of vertex n-1. Our CSR is complete.

#include "wgm.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ULL;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next() { s += kGolden; return mix64(s); }
  // uniform double in [0, 1) from the top 53 bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // integer in [0, bound) by multiply-high of the top 32 bits
  uint32_t below(uint32_t bound) { return (uint32_t)(((next() >> 32) * (uint64_t)bound) >> 32); }
};

// LSD radix sort of 64-bit keys on their low `bits` bits (11-bit digits).
void radix_sort_u64(std::vector<uint64_t>& keys, int bits) {
  std::vector<uint64_t> tmp(keys.size());
  const int kDigit = 11;
  const size_t kBuckets = size_t(1) << kDigit;
  std::vector<size_t> count(kBuckets);
  for (int shift = 0; shift < bits; shift += kDigit) {
    std::fill(count.begin(), count.end(), 0);
    for (uint64_t k : keys) count[(k >> shift) & (kBuckets - 1)]++;
    size_t sum = 0;
    for (size_t b = 0; b < kBuckets; b++) { size_t c = count[b]; count[b] = sum; sum += c; }
    for (uint64_t k : keys) tmp[count[(k >> shift) & (kBuckets - 1)]++] = k;
    keys.swap(tmp);
  }
}

int bits_for(uint64_t n) {  // bits needed to hold values < n
  int b = 1;
  while ((uint64_t(1) << b) < n) b++;
  return b;
}

// Build CSR from packed (src << 32 | dst) keys that are sorted and unique.
int csr_from_sorted(int32_t n, const std::vector<uint64_t>& keys, wgm_csr* out) {
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  out->m = (int64_t)keys.size();
  if (out->m > INT32_MAX) return WGM_ERR_INVALID;
  out->row_ptr = (int32_t*)std::malloc(sizeof(int32_t) * ((size_t)n + 1));
  out->col = (int32_t*)std::malloc(sizeof(int32_t) * (keys.empty() ? 1 : keys.size()));
  if (!out->row_ptr || !out->col) { wgm_csr_free(out); return WGM_ERR_NOMEM; }
  std::memset(out->row_ptr, 0, sizeof(int32_t) * ((size_t)n + 1));
  for (size_t e = 0; e < keys.size(); e++) {
    out->row_ptr[(keys[e] >> 32) + 1]++;
    out->col[e] = (int32_t)(keys[e] & 0xffffffffULL);
  }
  for (int32_t v = 0; v < n; v++) out->row_ptr[v + 1] += out->row_ptr[v];
  return WGM_OK;
}

void sort_unique(std::vector<uint64_t>& keys, int32_t n) {
  radix_sort_u64(keys, 32 + bits_for((uint64_t)n));
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
}

}  // namespace

extern "C" {

uint64_t wgm_splitmix64_next(uint64_t* state) {
  *state += kGolden;
  return mix64(*state);
}

void wgm_free(void* p) { std::free(p); }

void wgm_csr_free(wgm_csr* g) {
  if (!g) return;
  std::free(g->row_ptr);
  std::free(g->col);
  std::free(g->wt);
  g->row_ptr = nullptr;
  g->col = nullptr;
  g->wt = nullptr;
}

int wgm_gen_er(int32_t n, int64_t pairs, uint64_t seed, wgm_csr* out) {
  if (!out || n < 2) return WGM_ERR_INVALID;
  if ((double)pairs > 0.5 * (double)n * (double)(n - 1)) return WGM_ERR_INVALID;
  // Open-addressing set of canonical pairs; the first `pairs` distinct
  // non-loop pairs of the stream are kept, in stream order.
  size_t cap = 1;
  while (cap < (size_t)pairs * 2 + 16) cap <<= 1;
  std::vector<uint64_t> table(cap, ~0ULL);
  std::vector<uint64_t> keys;
  keys.reserve((size_t)pairs * 2);
  SplitMix rng(seed);
  int64_t have = 0;
  while (have < pairs) {
    uint32_t u = rng.below((uint32_t)n), v = rng.below((uint32_t)n);
    if (u == v) continue;
    uint32_t a = std::min(u, v), b = std::max(u, v);
    uint64_t key = ((uint64_t)a << 32) | b;
    size_t h = (size_t)(mix64(key) & (cap - 1));
    bool dup = false;
    while (table[h] != ~0ULL) {
      if (table[h] == key) { dup = true; break; }
      h = (h + 1) & (cap - 1);
    }
    if (dup) continue;
    table[h] = key;
    keys.push_back(((uint64_t)a << 32) | b);
    keys.push_back(((uint64_t)b << 32) | a);
    have++;
  }
  sort_unique(keys, n);
  return csr_from_sorted(n, keys, out);
}

int wgm_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                 uint64_t seed, int32_t truncate_n, wgm_csr* out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1) return WGM_ERR_INVALID;
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return WGM_ERR_INVALID;
  const uint64_t full = uint64_t(1) << scale;
  const int32_t n = truncate_n > 0 ? truncate_n : (int32_t)full;
  if ((uint64_t)n > full) return WGM_ERR_INVALID;
  const uint64_t edges = (uint64_t)edge_factor << scale;
  std::vector<uint64_t> keys;
  keys.reserve(edges);
  SplitMix rng(seed);
  const double ab = a + b, abc = a + b + c;
  for (uint64_t e = 0; e < edges; e++) {
    uint64_t u = 0, v = 0;
    for (int level = 0; level < scale; level++) {
      double r = rng.uniform();
      uint64_t bu, bv;
      if (r < a) { bu = 0; bv = 0; }
      else if (r < ab) { bu = 0; bv = 1; }
      else if (r < abc) { bu = 1; bv = 0; }
      else { bu = 1; bv = 1; }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    if (u == v || u >= (uint64_t)n || v >= (uint64_t)n) continue;
    keys.push_back((u << 32) | v);
  }
  sort_unique(keys, n);
  return csr_from_sorted(n, keys, out);
}

int wgm_gen_grid(int32_t rows, int32_t cols, wgm_csr* out) {
  if (!out || rows < 1 || cols < 1) return WGM_ERR_INVALID;
  const int64_t n64 = (int64_t)rows * cols;
  if (n64 > INT32_MAX) return WGM_ERR_INVALID;
  const int32_t n = (int32_t)n64;
  const int64_t m = 2 * ((int64_t)rows * (cols - 1) + (int64_t)(rows - 1) * cols);
  if (m > INT32_MAX) return WGM_ERR_INVALID;
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  out->m = m;
  out->row_ptr = (int32_t*)std::malloc(sizeof(int32_t) * ((size_t)n + 1));
  out->col = (int32_t*)std::malloc(sizeof(int32_t) * (m ? (size_t)m : 1));
  if (!out->row_ptr || !out->col) { wgm_csr_free(out); return WGM_ERR_NOMEM; }
  int64_t e = 0;
  for (int32_t r = 0; r < rows; r++) {
    for (int32_t c = 0; c < cols; c++) {
      const int32_t v = r * cols + c;
      out->row_ptr[v] = (int32_t)e;
      // ascending ids: up, left, right, down
      if (r > 0) out->col[e++] = v - cols;
      if (c > 0) out->col[e++] = v - 1;
      if (c + 1 < cols) out->col[e++] = v + 1;
      if (r + 1 < rows) out->col[e++] = v + cols;
    }
  }
  out->row_ptr[n] = (int32_t)e;
  return WGM_OK;
}

int wgm_gen_weights(wgm_csr* g, int32_t lo, int32_t hi, uint64_t seed) {
  if (!g || lo < 0 || hi < lo) return WGM_ERR_INVALID;
  std::free(g->wt);
  g->wt = (int32_t*)std::malloc(sizeof(int32_t) * (g->m ? (size_t)g->m : 1));
  if (!g->wt) return WGM_ERR_NOMEM;
  SplitMix rng(seed);
  const uint32_t span = (uint32_t)(hi - lo + 1);
  for (int64_t e = 0; e < g->m; e++) g->wt[e] = lo + (int32_t)rng.below(span);
  return WGM_OK;
}

int wgm_gen_priorities(const wgm_csr* g, uint64_t seed, uint32_t* prio) {
  if (!g || !prio) return WGM_ERR_INVALID;
  // Draw k of vertex v: top 32 bits of the first splitmix64 output for the
  // state seed ^ (v + (k << 32)). Vertices are finalised in id order; a
  // vertex re-draws (k += 1) while it ties with an already-final neighbour,
  // so every edge joins two different priorities.
  for (int32_t v = 0; v < g->n; v++) {
    for (uint64_t k = 0;; k++) {
      uint64_t st = seed ^ ((uint64_t)(uint32_t)v + (k << 32));
      uint32_t p = (uint32_t)(wgm_splitmix64_next(&st) >> 32);
      bool tie = false;
      for (int32_t e = g->row_ptr[v]; e < g->row_ptr[v + 1]; e++) {
        int32_t u = g->col[e];
        if (u < v && prio[u] == p) { tie = true; break; }
      }
      if (!tie) { prio[v] = p; break; }
    }
  }
  return WGM_OK;
}

// Batched workload: graph i draws n_i = round(2^(10 + 8u)) and a per-graph
// seed from one splitmix64 stream; even i is RMAT(scale ceil(log2 n_i),
// edge factor 16, truncated to n_i vertices), odd i is a floor(sqrt(n_i))^2
// grid. Every graph gets weights in [1, 255].
int wgm_batch_spec(int32_t count, uint64_t seed, int32_t lo_exp, int32_t hi_exp,
                   wgm_batch_item* items) {
  if (count < 0 || !items || lo_exp < 1 || hi_exp < lo_exp || hi_exp > 28) return WGM_ERR_INVALID;
  SplitMix rng(seed);
  for (int32_t i = 0; i < count; i++) {
    double u = rng.uniform();
    double e = lo_exp + (hi_exp - lo_exp) * u;
    int64_t n = (int64_t)std::llround(std::pow(2.0, e));
    uint64_t gseed = rng.next();
    wgm_batch_item& it = items[i];
    it.index = i;
    it.kind = (i % 2 == 0) ? WGM_GRAPH_RMAT : WGM_GRAPH_GRID;
    if (it.kind == WGM_GRAPH_RMAT) {
      it.n = (int32_t)n;
      it.scale = bits_for((uint64_t)n);  // ceil(log2 n) for n >= 2
      it.rows = it.cols = 0;
    } else {
      int32_t side = (int32_t)std::floor(std::sqrt((double)n));
      while ((int64_t)(side + 1) * (side + 1) <= n) side++;
      while ((int64_t)side * side > n) side--;
      it.rows = it.cols = side;
      it.n = side * side;
      it.scale = 0;
    }
    it.query = (i % 4 == 0 || i % 4 == 3) ? WGM_QUERY_BFS : WGM_QUERY_SSSP;
    it.graph_seed = gseed;
    it.weight_seed = gseed ^ 3ULL;
  }
  return WGM_OK;
}

int wgm_batch_generate(int32_t count, const wgm_batch_item* items, wgm_csr* graphs, int32_t threads) {
  if (count < 0 || !items || !graphs) return WGM_ERR_INVALID;
  if (threads < 1) threads = 1;
  std::vector<int> status(count, WGM_OK);
  auto work = [&](int t) {
    for (int32_t i = t; i < count; i += threads) {
      const wgm_batch_item& it = items[i];
      int rc;
      if (it.kind == WGM_GRAPH_RMAT)
        rc = wgm_gen_rmat(it.scale, 16, 0.57, 0.19, 0.19, it.graph_seed, it.n, &graphs[i]);
      else
        rc = wgm_gen_grid(it.rows, it.cols, &graphs[i]);
      if (rc == WGM_OK) rc = wgm_gen_weights(&graphs[i], 1, 255, it.weight_seed);
      status[i] = rc;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; t++) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  for (int32_t i = 0; i < count; i++)
    if (status[i] != WGM_OK) return status[i];
  return WGM_OK;
}

}  // extern "C"
