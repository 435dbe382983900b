// formats.cpp -- graph and solution files of the reference's datasets, read
// into (and written from) the library's 0-based CSR.
//
//   Galois .gr v1   src/lonestar/graph_apps/bfs/bfs.h:95-169 (readFromGR)
//   DIMACS sp       src/pannotia/graph_parser/parse.cpp:177-281 (parseCOO),
//                   692-797 (parseCOO_transpose)
//   METIS           parse.cpp:58-175 (parseMetis)
//   MatrixMarket    parse.cpp:467-580 (parseMM)
//   edge list       src/lonestar/graph_apps/sssp/sssp.h:49-94 (readFromEdges)
//   solutions       "%d:%d" (bfs.h:336-365), "%d: %d" (color.h:149-190,
//                   pannotia sssp.h:127-170), "%d" (mis.h:170-195), "%f" (bc.h:176-203)
//
// Differences from the reference, all on purpose:
//   * every reader builds a complete row_ptr (n + 1 entries, empty rows
//     included); the reference emits a row start only when a new head appears,
//     which misaligns rows after an isolated vertex (parse.cpp:148-161);
//   * the .gr CSR keeps every edge (the reference's psrc[n] = nedges drops the
//     last vertex's last out-edge, bfs.h:41-43);
//   * the undirected DIMACS transpose holds each arc and its reverse (the
//     reference inserts the transposed arc twice, parse.cpp:737-747);
//   * METIS files with edge weights (fmt 1 / 001) are read with their
//     weights; the reference reads every token as a neighbour;
//   * malformed files are errors with a message, never exit().
// Within a row, edges keep the order of the file (the reference's
// std::stable_sort by row), so CSR order is the reference's.

#include <algorithm>
#include <cerrno>
#include <cinttypes>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "wgm.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

struct Arc {
  int32_t row, col, wt;
};

// Counting sort by row, stable: CSR in file order within each row.
int arcs_to_csr(int64_t n, const std::vector<Arc>& arcs, bool weighted, wgm_csr* out) {
  if (n < 0 || n > INT32_MAX - 1) return fail(WGM_ERR_FORMAT, "vertex count %" PRId64 " out of range", n);
  if ((int64_t)arcs.size() > INT32_MAX) return fail(WGM_ERR_FORMAT, "more than 2^31-1 edges");
  std::memset(out, 0, sizeof *out);
  out->n = (int32_t)n;
  out->m = (int64_t)arcs.size();
  out->row_ptr = (int32_t*)std::calloc((size_t)n + 1, sizeof(int32_t));
  out->col = (int32_t*)std::malloc(sizeof(int32_t) * (arcs.empty() ? 1 : arcs.size()));
  if (weighted) out->wt = (int32_t*)std::malloc(sizeof(int32_t) * (arcs.empty() ? 1 : arcs.size()));
  if (!out->row_ptr || !out->col || (weighted && !out->wt)) {
    wgm_csr_free(out);
    return fail(WGM_ERR_NOMEM, "out of host memory");
  }
  for (const Arc& a : arcs) out->row_ptr[a.row + 1]++;
  for (int64_t v = 0; v < n; v++) out->row_ptr[v + 1] += out->row_ptr[v];
  std::vector<int32_t> pos(out->row_ptr, out->row_ptr + n);
  for (const Arc& a : arcs) {
    const int32_t p = pos[a.row]++;
    out->col[p] = a.col;
    if (weighted) out->wt[p] = a.wt;
  }
  return WGM_OK;
}

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// Whole-file read (the .gr reader maps the file the same way, bfs.h:97-110).
int slurp(const char* path, std::vector<unsigned char>* buf) {
  File fl(path, "rb");
  if (!fl.f) return fail(WGM_ERR_IO, "cannot open %s: %s", path, std::strerror(errno));
  if (std::fseek(fl.f, 0, SEEK_END) != 0) return fail(WGM_ERR_IO, "cannot seek %s", path);
  const long len = std::ftell(fl.f);
  if (len < 0) return fail(WGM_ERR_IO, "cannot size %s", path);
  std::fseek(fl.f, 0, SEEK_SET);
  buf->resize((size_t)len);
  if (len > 0 && std::fread(buf->data(), (size_t)len, 1, fl.f) != 1) return fail(WGM_ERR_IO, "short read on %s", path);
  return WGM_OK;
}

// Line reader over a text buffer.
struct Lines {
  const char* p;
  const char* end;
  int64_t lineno = 0;
  explicit Lines(const std::vector<unsigned char>& b) : p((const char*)b.data()), end((const char*)b.data() + b.size()) {}
  bool next(std::string* line) {
    if (p >= end) return false;
    const char* q = (const char*)std::memchr(p, '\n', (size_t)(end - p));
    if (!q) q = end;
    line->assign(p, (size_t)(q - p));
    if (!line->empty() && line->back() == '\r') line->pop_back();
    p = q < end ? q + 1 : end;
    lineno++;
    return true;
  }
};

// Parse a signed integer at *s (leading blanks skipped); false if none.
bool parse_int(const char** s, int64_t* v) {
  const char* c = *s;
  while (*c == ' ' || *c == '\t') c++;
  if (!*c) return false;
  char* e = nullptr;
  errno = 0;
  const long long x = std::strtoll(c, &e, 10);
  if (e == c || errno) return false;
  *v = x;
  *s = e;
  return true;
}

bool blank(const std::string& s) {
  for (char c : s)
    if (c != ' ' && c != '\t') return false;
  return true;
}

// ----------------------------------------------------------------- .gr v1
int read_gr(const char* path, int32_t flags, wgm_csr* out) {
  std::vector<unsigned char> b;
  int rc = slurp(path, &b);
  if (rc) return rc;
  if (b.size() < 32) return fail(WGM_ERR_FORMAT, "%s: shorter than the .gr header", path);
  uint64_t h[4];
  std::memcpy(h, b.data(), sizeof h);  // little-endian host (x86-64), as the file
  const uint64_t version = h[0], esize = h[1], nn = h[2], ne = h[3];
  if (version != 1) return fail(WGM_ERR_FORMAT, "%s: .gr version %" PRIu64 ", expected 1", path, version);
  if (nn > (uint64_t)INT32_MAX - 1 || ne > (uint64_t)INT32_MAX)
    return fail(WGM_ERR_FORMAT, "%s: %" PRIu64 " nodes / %" PRIu64 " edges out of range", path, nn, ne);
  if (esize != 0 && esize != 4) return fail(WGM_ERR_FORMAT, "%s: edge data of %" PRIu64 " bytes (0 or 4 supported)", path, esize);
  const size_t idx_off = 32, dst_off = idx_off + 8 * nn, dat_off = dst_off + 4 * (ne + (ne & 1));
  const size_t need = esize ? dat_off + 4 * ne : dst_off + 4 * ne;
  if (b.size() < need) return fail(WGM_ERR_FORMAT, "%s: %zu bytes, the header needs %zu", path, b.size(), need);
  std::vector<uint64_t> idx(nn);
  if (nn) std::memcpy(idx.data(), b.data() + idx_off, 8 * nn);
  uint64_t prev = 0;
  for (uint64_t v = 0; v < nn; v++) {
    if (idx[v] < prev || idx[v] > ne) return fail(WGM_ERR_FORMAT, "%s: edge index of node %" PRIu64 " not monotone", path, v);
    prev = idx[v];
  }
  if (nn && idx[nn - 1] != ne) return fail(WGM_ERR_FORMAT, "%s: last edge index %" PRIu64 " != %" PRIu64, path, idx[nn - 1], ne);
  if (!nn && ne) return fail(WGM_ERR_FORMAT, "%s: edges without nodes", path);
  const uint32_t* dst = (const uint32_t*)(b.data() + dst_off);
  const int32_t* dat = (const int32_t*)(b.data() + dat_off);
  std::vector<Arc> arcs;
  arcs.reserve((size_t)ne);
  for (uint64_t v = 0, e = 0; v < nn; v++) {
    for (; e < idx[v]; e++) {
      uint32_t d;
      std::memcpy(&d, dst + e, 4);
      if (d >= nn) return fail(WGM_ERR_FORMAT, "%s: edge %" PRIu64 " from %" PRIu64 " to %u out of range", path, e, v, d);
      int32_t w = 0;
      if (esize) std::memcpy(&w, dat + e, 4);
      if (flags & WGM_READ_TRANSPOSE)
        arcs.push_back({(int32_t)d, (int32_t)v, w});
      else
        arcs.push_back({(int32_t)v, (int32_t)d, w});
    }
  }
  return arcs_to_csr((int64_t)nn, arcs, esize != 0, out);
}

int write_gr(const char* path, const wgm_csr* g) {
  File fl(path, "wb");
  if (!fl.f) return fail(WGM_ERR_IO, "cannot create %s: %s", path, std::strerror(errno));
  const uint64_t h[4] = {1, g->wt ? 4u : 0u, (uint64_t)g->n, (uint64_t)g->m};
  bool ok = std::fwrite(h, sizeof h, 1, fl.f) == 1;
  for (int32_t v = 0; ok && v < g->n; v++) {
    const uint64_t e = (uint64_t)g->row_ptr[v + 1];
    ok = std::fwrite(&e, 8, 1, fl.f) == 1;
  }
  if (ok && g->m) ok = std::fwrite(g->col, 4, (size_t)g->m, fl.f) == (size_t)g->m;
  const uint32_t pad = 0;
  if (ok && (g->m & 1)) ok = std::fwrite(&pad, 4, 1, fl.f) == 1;
  if (ok && g->wt && g->m) ok = std::fwrite(g->wt, 4, (size_t)g->m, fl.f) == (size_t)g->m;
  if (!ok) return fail(WGM_ERR_IO, "write to %s failed", path);
  return WGM_OK;
}

// ------------------------------------------------------------------ DIMACS
// "c ..." comments, "p sp <n> <m>", "a <u> <v> <w>" (1-based). Undirected:
// every arc is followed by its reverse (parse.cpp:217-230).
int read_dimacs(const char* path, int32_t flags, wgm_csr* out) {
  std::vector<unsigned char> b;
  int rc = slurp(path, &b);
  if (rc) return rc;
  Lines ln(b);
  std::string s;
  int64_t n = -1, m = -1;
  std::vector<Arc> arcs;
  const bool und = flags & WGM_READ_UNDIRECTED, tr = flags & WGM_READ_TRANSPOSE;
  while (ln.next(&s)) {
    if (s.empty() || s[0] == 'c' || blank(s)) continue;
    const char* c = s.c_str() + 1;
    if (s[0] == 'p') {
      while (*c == ' ' || *c == '\t') c++;
      while (*c && *c != ' ' && *c != '\t') c++;  // problem type ("sp")
      if (n >= 0 || !parse_int(&c, &n) || !parse_int(&c, &m) || n < 0 || m < 0)
        return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": bad or repeated problem line", path, ln.lineno);
      arcs.reserve((size_t)(und ? 2 * m : m));
    } else if (s[0] == 'a') {
      int64_t u, v, w;
      if (n < 0) return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": arc before the problem line", path, ln.lineno);
      if (!parse_int(&c, &u) || !parse_int(&c, &v) || !parse_int(&c, &w))
        return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": arc needs <u> <v> <w>", path, ln.lineno);
      if (u < 1 || u > n || v < 1 || v > n)
        return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": vertex out of 1..%" PRId64, path, ln.lineno, n);
      if (w < INT32_MIN || w > INT32_MAX) return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": weight out of range", path, ln.lineno);
      const int32_t a = (int32_t)(u - 1), d = (int32_t)(v - 1);
      arcs.push_back(tr ? Arc{d, a, (int32_t)w} : Arc{a, d, (int32_t)w});
      if (und) arcs.push_back(tr ? Arc{a, d, (int32_t)w} : Arc{d, a, (int32_t)w});
    } else {
      return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": unknown line type '%c'", path, ln.lineno, s[0]);
    }
  }
  if (n < 0) return fail(WGM_ERR_FORMAT, "%s: no problem line", path);
  if ((int64_t)arcs.size() != (und ? 2 * m : m))
    return fail(WGM_ERR_FORMAT, "%s: %zu arcs read, the problem line says %" PRId64, path,
                und ? arcs.size() / 2 : arcs.size(), m);
  return arcs_to_csr(n, arcs, true, out);
}

// Writes every CSR edge as an arc (a directed file: read it back without
// WGM_READ_UNDIRECTED to get the same CSR).
int write_dimacs(const char* path, const wgm_csr* g) {
  File fl(path, "w");
  if (!fl.f) return fail(WGM_ERR_IO, "cannot create %s: %s", path, std::strerror(errno));
  bool ok = std::fprintf(fl.f, "c written by libwgm\np sp %d %" PRId64 "\n", g->n, g->m) > 0;
  for (int32_t v = 0; ok && v < g->n; v++)
    for (int32_t e = g->row_ptr[v]; ok && e < g->row_ptr[v + 1]; e++)
      ok = std::fprintf(fl.f, "a %d %d %d\n", v + 1, g->col[e] + 1, g->wt ? g->wt[e] : 1) > 0;
  if (!ok) return fail(WGM_ERR_IO, "write to %s failed", path);
  return WGM_OK;
}

// ------------------------------------------------------------------- METIS
// "%" comments; header "<n> <m> [fmt]"; then line i lists the (1-based)
// neighbours of vertex i; fmt 1 / 001: "<v> <w>" pairs. Undirected files
// list each edge at both ends (the reader checks 2m entries); the
// reference's `directed` flag only changes the expected count.
int read_metis(const char* path, int32_t flags, wgm_csr* out) {
  std::vector<unsigned char> b;
  int rc = slurp(path, &b);
  if (rc) return rc;
  Lines ln(b);
  std::string s;
  int64_t n = -1, m = -1, fmt = 0, v = 0;
  std::vector<Arc> arcs;
  const bool tr = flags & WGM_READ_TRANSPOSE;
  bool weighted = false;
  while (ln.next(&s)) {
    if (!s.empty() && s[0] == '%') continue;
    const char* c = s.c_str();
    if (n < 0) {
      if (blank(s)) continue;
      if (!parse_int(&c, &n) || !parse_int(&c, &m) || n < 0 || m < 0)
        return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": header needs <n> <m>", path, ln.lineno);
      if (parse_int(&c, &fmt) && fmt != 0 && fmt != 1)
        return fail(WGM_ERR_FORMAT, "%s: METIS fmt %03" PRId64 " (vertex sizes / weights) is not supported", path, fmt);
      weighted = fmt == 1;
      arcs.reserve((size_t)(flags & WGM_READ_UNDIRECTED ? 2 * m : m));
      continue;
    }
    if (v >= n) {
      if (blank(s)) continue;
      return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": more adjacency lines than %" PRId64 " vertices", path, ln.lineno, n);
    }
    int64_t u, w = 0;
    while (parse_int(&c, &u)) {
      if (u < 1 || u > n) return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": neighbour %" PRId64 " out of 1..%" PRId64, path, ln.lineno, u, n);
      if (weighted && !parse_int(&c, &w))
        return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": neighbour %" PRId64 " without a weight", path, ln.lineno, u);
      arcs.push_back(tr ? Arc{(int32_t)(u - 1), (int32_t)v, (int32_t)w} : Arc{(int32_t)v, (int32_t)(u - 1), (int32_t)w});
    }
    while (*c == ' ' || *c == '\t') c++;
    if (*c) return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": unexpected '%c'", path, ln.lineno, *c);
    v++;
  }
  if (n < 0) return fail(WGM_ERR_FORMAT, "%s: no header", path);
  const int64_t expect = flags & WGM_READ_UNDIRECTED ? 2 * m : m;
  if ((int64_t)arcs.size() != expect)
    return fail(WGM_ERR_FORMAT, "%s: %zu adjacency entries, the header implies %" PRId64, path, arcs.size(), expect);
  return arcs_to_csr(n, arcs, weighted, out);
}

// Writes the CSR as adjacency lines; m in the header is the CSR edge count
// halved when `undirected` (each edge appears at both ends).
int write_metis(const char* path, const wgm_csr* g, bool undirected) {
  File fl(path, "w");
  if (!fl.f) return fail(WGM_ERR_IO, "cannot create %s: %s", path, std::strerror(errno));
  if (undirected && (g->m & 1)) return fail(WGM_ERR_INVALID, "an undirected METIS file needs an even edge count");
  bool ok = std::fprintf(fl.f, "%d %" PRId64 "%s\n", g->n, undirected ? g->m / 2 : g->m, g->wt ? " 1" : "") > 0;
  for (int32_t v = 0; ok && v < g->n; v++) {
    for (int32_t e = g->row_ptr[v]; ok && e < g->row_ptr[v + 1]; e++) {
      ok = (g->wt ? std::fprintf(fl.f, e == g->row_ptr[v] ? "%d %d" : " %d %d", g->col[e] + 1, g->wt[e])
                  : std::fprintf(fl.f, e == g->row_ptr[v] ? "%d" : " %d", g->col[e] + 1)) > 0;
    }
    ok = ok && std::fputc('\n', fl.f) != EOF;
  }
  if (!ok) return fail(WGM_ERR_IO, "write to %s failed", path);
  return WGM_OK;
}

// ------------------------------------------------------------ MatrixMarket
// "%" comments (the banner included); "<rows> <cols> <nnz>"; then
// "<i> <j> [w]". Self-loops are dropped, as parse.cpp:517-520 does.
int read_mm(const char* path, int32_t flags, wgm_csr* out) {
  std::vector<unsigned char> b;
  int rc = slurp(path, &b);
  if (rc) return rc;
  Lines ln(b);
  std::string s;
  int64_t n = -1, cols = 0, nnz = 0, entries = 0;
  std::vector<Arc> arcs;
  const bool und = flags & WGM_READ_UNDIRECTED, tr = flags & WGM_READ_TRANSPOSE;
  int weighted = -1;
  while (ln.next(&s)) {
    if ((!s.empty() && s[0] == '%') || blank(s)) continue;
    const char* c = s.c_str();
    if (n < 0) {
      if (!parse_int(&c, &n) || !parse_int(&c, &cols) || !parse_int(&c, &nnz) || n < 0 || cols != n || nnz < 0)
        return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": size line needs <n> <n> <nnz>", path, ln.lineno);
      continue;
    }
    int64_t i, j, w = 0;
    if (!parse_int(&c, &i) || !parse_int(&c, &j))
      return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": entry needs <i> <j>", path, ln.lineno);
    const bool has_w = parse_int(&c, &w);
    if (weighted < 0) weighted = has_w;
    if (has_w != (weighted == 1)) return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": weights on some entries only", path, ln.lineno);
    if (i < 1 || i > n || j < 1 || j > n) return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": index out of 1..%" PRId64, path, ln.lineno, n);
    entries++;
    if (i == j) continue;
    const int32_t a = (int32_t)(i - 1), d = (int32_t)(j - 1);
    arcs.push_back(tr ? Arc{d, a, (int32_t)w} : Arc{a, d, (int32_t)w});
    if (und) arcs.push_back(tr ? Arc{a, d, (int32_t)w} : Arc{d, a, (int32_t)w});
  }
  if (n < 0) return fail(WGM_ERR_FORMAT, "%s: no size line", path);
  if (entries != nnz) return fail(WGM_ERR_FORMAT, "%s: %" PRId64 " entries, the size line says %" PRId64, path, entries, nnz);
  return arcs_to_csr(n, arcs, weighted == 1, out);
}

// ------------------------------------------------------ Lonestar edge list
// "<n> <m>" then m lines "<src> <dst> [w]" (0-based).
int read_edges(const char* path, int32_t flags, wgm_csr* out) {
  std::vector<unsigned char> b;
  int rc = slurp(path, &b);
  if (rc) return rc;
  Lines ln(b);
  std::string s;
  int64_t n = -1, m = 0;
  std::vector<Arc> arcs;
  int weighted = -1;
  const bool und = flags & WGM_READ_UNDIRECTED, tr = flags & WGM_READ_TRANSPOSE;
  while (ln.next(&s)) {
    if (blank(s)) continue;
    const char* c = s.c_str();
    if (n < 0) {
      if (!parse_int(&c, &n) || !parse_int(&c, &m) || n < 0 || m < 0)
        return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": header needs <n> <m>", path, ln.lineno);
      continue;
    }
    int64_t u, v, w = 0;
    if (!parse_int(&c, &u) || !parse_int(&c, &v)) return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": edge needs <src> <dst>", path, ln.lineno);
    const bool has_w = parse_int(&c, &w);
    if (weighted < 0) weighted = has_w;
    if (has_w != (weighted == 1)) return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": weights on some edges only", path, ln.lineno);
    if (u < 0 || u >= n || v < 0 || v >= n) return fail(WGM_ERR_FORMAT, "%s:%" PRId64 ": vertex out of 0..%" PRId64, path, ln.lineno, n - 1);
    arcs.push_back(tr ? Arc{(int32_t)v, (int32_t)u, (int32_t)w} : Arc{(int32_t)u, (int32_t)v, (int32_t)w});
    if (und) arcs.push_back(tr ? Arc{(int32_t)u, (int32_t)v, (int32_t)w} : Arc{(int32_t)v, (int32_t)u, (int32_t)w});
  }
  if (n < 0) return fail(WGM_ERR_FORMAT, "%s: no header", path);
  if ((int64_t)arcs.size() != (und ? 2 * m : m)) return fail(WGM_ERR_FORMAT, "%s: edge count differs from the header", path);
  return arcs_to_csr(n, arcs, weighted == 1, out);
}

// --------------------------------------------------------------- solutions
const char* sol_fmt(int32_t style) {
  switch (style) {
    case WGM_SOL_COLON: return "%d:%d\n";
    case WGM_SOL_COLON_SPACE: return "%d: %d\n";
    default: return nullptr;
  }
}

// One value per line, the way the reference's fscanf patterns read it: an
// optional "<index>:" (any blanks around ':') before the value in the colon
// styles. Returns false on a malformed or missing line.
bool next_value(Lines& ln, int32_t style, std::string* s, double* out, bool as_float) {
  while (ln.next(s)) {
    if (blank(*s)) continue;
    const char* c = s->c_str();
    if (style == WGM_SOL_COLON || style == WGM_SOL_COLON_SPACE) {
      int64_t idx;
      if (!parse_int(&c, &idx)) return false;
      while (*c == ' ' || *c == '\t') c++;
      if (*c != ':') return false;
      c++;
    }
    while (*c == ' ' || *c == '\t') c++;
    char* e = nullptr;
    errno = 0;
    const double v = as_float ? std::strtod(c, &e) : (double)std::strtoll(c, &e, 10);
    if (e == c || errno) return false;
    *out = v;
    return true;
  }
  return false;
}

}  // namespace

extern "C" {

const char* wgm_format_error(void) { return g_err.c_str(); }

int wgm_read_graph(const char* path, int32_t format, int32_t flags, wgm_csr* out) {
  if (!path || !out) return fail(WGM_ERR_INVALID, "NULL argument");
  if (flags & ~(WGM_READ_UNDIRECTED | WGM_READ_TRANSPOSE)) return fail(WGM_ERR_INVALID, "unknown flags 0x%x", flags);
  std::memset(out, 0, sizeof *out);
  switch (format) {
    case WGM_FMT_GR:
      if (flags & WGM_READ_UNDIRECTED) return fail(WGM_ERR_INVALID, ".gr files hold every arc; no undirected flag");
      return read_gr(path, flags, out);
    case WGM_FMT_DIMACS: return read_dimacs(path, flags, out);
    case WGM_FMT_METIS: return read_metis(path, flags, out);
    case WGM_FMT_MM: return read_mm(path, flags, out);
    case WGM_FMT_EDGES: return read_edges(path, flags, out);
    default: return fail(WGM_ERR_INVALID, "unknown graph format %d", format);
  }
}

int wgm_write_graph(const char* path, int32_t format, int32_t flags, const wgm_csr* g) {
  if (!path || !g || (g->n > 0 && !g->row_ptr) || (g->m > 0 && !g->col)) return fail(WGM_ERR_INVALID, "NULL argument");
  if (g->n < 0 || g->m < 0 || g->m > INT32_MAX) return fail(WGM_ERR_INVALID, "bad CSR shape");
  switch (format) {
    case WGM_FMT_GR: return write_gr(path, g);
    case WGM_FMT_DIMACS: return write_dimacs(path, g);
    case WGM_FMT_METIS: return write_metis(path, g, flags & WGM_READ_UNDIRECTED);
    default: return fail(WGM_ERR_INVALID, "format %d cannot be written", format);
  }
}

int wgm_write_solution(const char* path, int32_t style, const int32_t* values, int32_t n) {
  if (!path || (n > 0 && !values) || n < 0) return fail(WGM_ERR_INVALID, "bad argument");
  if (style != WGM_SOL_COLON && style != WGM_SOL_COLON_SPACE && style != WGM_SOL_PLAIN)
    return fail(WGM_ERR_INVALID, "unknown solution style %d", style);
  File fl(path, "w");
  if (!fl.f) return fail(WGM_ERR_IO, "cannot create %s: %s", path, std::strerror(errno));
  bool ok = true;
  // Lonestar files number from 0, Pannotia's writers from 1 (color.h:156)
  for (int32_t i = 0; ok && i < n; i++)
    ok = (style == WGM_SOL_PLAIN ? std::fprintf(fl.f, "%d\n", values[i])
                                 : std::fprintf(fl.f, sol_fmt(style), style == WGM_SOL_COLON ? i : i + 1, values[i])) > 0;
  if (!ok) return fail(WGM_ERR_IO, "write to %s failed", path);
  return WGM_OK;
}

int wgm_check_solution(const char* path, int32_t style, const int32_t* values, int32_t n, int32_t* first_bad) {
  if (first_bad) *first_bad = -1;
  if (!path || (n > 0 && !values) || n < 0) return fail(WGM_ERR_INVALID, "bad argument");
  if (style != WGM_SOL_COLON && style != WGM_SOL_COLON_SPACE && style != WGM_SOL_PLAIN)
    return fail(WGM_ERR_INVALID, "unknown solution style %d", style);
  std::vector<unsigned char> b;
  int rc = slurp(path, &b);
  if (rc) return rc;
  Lines ln(b);
  std::string s;
  for (int32_t i = 0; i < n; i++) {
    double v;
    if (!next_value(ln, style, &s, &v, false)) {
      if (first_bad) *first_bad = i;
      return fail(WGM_ERR_FORMAT, "%s: no value for vertex %d", path, i);
    }
    if ((int64_t)v != (int64_t)values[i]) {
      if (first_bad) *first_bad = i;
      return fail(WGM_ERR_MISMATCH, "%s: vertex %d has %d, the file %" PRId64, path, i, values[i], (int64_t)v);
    }
  }
  return WGM_OK;
}

int wgm_write_solution_f32(const char* path, const float* values, int32_t n) {
  if (!path || (n > 0 && !values) || n < 0) return fail(WGM_ERR_INVALID, "bad argument");
  File fl(path, "w");
  if (!fl.f) return fail(WGM_ERR_IO, "cannot create %s: %s", path, std::strerror(errno));
  bool ok = true;
  for (int32_t i = 0; ok && i < n; i++) ok = std::fprintf(fl.f, "%.9g\n", (double)values[i]) > 0;
  if (!ok) return fail(WGM_ERR_IO, "write to %s failed", path);
  return WGM_OK;
}

int wgm_check_solution_f32(const char* path, const float* values, int32_t n, float tol, int32_t* first_bad) {
  if (first_bad) *first_bad = -1;
  if (!path || (n > 0 && !values) || n < 0 || !(tol >= 0)) return fail(WGM_ERR_INVALID, "bad argument");
  std::vector<unsigned char> b;
  int rc = slurp(path, &b);
  if (rc) return rc;
  Lines ln(b);
  std::string s;
  for (int32_t i = 0; i < n; i++) {
    double v;
    if (!next_value(ln, WGM_SOL_PLAIN, &s, &v, true)) {
      if (first_bad) *first_bad = i;
      return fail(WGM_ERR_FORMAT, "%s: no value for vertex %d", path, i);
    }
    // |file - value| > tol fails (bc.h:194); NaN never matches
    const double d = v - (double)values[i];
    if (!(d <= tol && -d <= tol)) {
      if (first_bad) *first_bad = i;
      return fail(WGM_ERR_MISMATCH, "%s: vertex %d has %g, the file %g", path, i, (double)values[i], v);
    }
  }
  return WGM_OK;
}

}  // extern "C"
