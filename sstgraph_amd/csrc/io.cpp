// Graph files on the host: the reference's readers for .adj and .mtx
// (include/SSTGraph/internal/helpers.hpp:363-631, get_edges_from_file) and
// its AdjacencyGraph writer (graph_utilities.cpp:151-183), over the C ABI.
//
// Formats, as the reference reads them:
//  * .adj / .adj.shuf: "AdjacencyGraph" (or "WeightedAdjacencyGraph"), n, m,
//    n row offsets, m destinations (then m weights), whitespace separated;
//    row i holds destinations [offset[i], offset[i+1]) (the last row ends at
//    m).
//  * .mtx: "%%MatrixMarket matrix coordinate <integer|pattern|real>
//    <symmetric|general>", comment lines starting with '%' after the first
//    line, then "rows cols entries" and one "i j [value]" line per entry,
//    1-indexed.  Every entry is returned in both directions, whether the
//    file says symmetric or general (helpers.hpp:505-521), and the matrix
//    must be square.
// Where the reference prints and exits on a malformed file these return
// SSTG_E_INVAL, and SSTG_E_IO when the file cannot be opened.
//
// Tokens are parsed by up to 16 threads, each over a slice of the file cut
// at whitespace.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "sstg.h"

namespace {

bool is_space(char c) {
  return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\f' ||
         c == '\v';
}

bool ends_with(const std::string &s, const char *suf) {
  const size_t k = std::strlen(suf);
  return s.size() >= k && s.compare(s.size() - k, k, suf) == 0;
}

int read_file(const char *path, std::vector<char> *buf) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return SSTG_E_IO;
  std::fseek(f, 0, SEEK_END);
  long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (len < 0) {
    std::fclose(f);
    return SSTG_E_IO;
  }
  buf->resize((size_t)len + 1);
  size_t got = len ? std::fread(buf->data(), 1, (size_t)len, f) : 0;
  std::fclose(f);
  if (got != (size_t)len) return SSTG_E_IO;
  (*buf)[(size_t)len] = 0;
  buf->resize((size_t)len);
  return SSTG_OK;
}

// Start offsets of the whitespace-separated tokens of buf[from, to).
std::vector<uint64_t> tokenize(const std::vector<char> &buf, size_t from) {
  const size_t n = buf.size();
  unsigned nt = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (n - from < (1u << 20)) nt = 1;
  std::vector<size_t> cut(nt + 1);
  cut[0] = from;
  cut[nt] = n;
  for (unsigned t = 1; t < nt; t++) {
    size_t p = from + (n - from) * t / nt;
    while (p < n && !is_space(buf[p])) p++;  // cut only at whitespace
    cut[t] = std::max(p, cut[t - 1]);
  }
  std::vector<std::vector<uint64_t>> part(nt);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      for (size_t p = cut[t]; p < cut[t + 1]; p++)
        if (!is_space(buf[p]) && (p == from || is_space(buf[p - 1])))
          part[t].push_back(p);
    });
  for (auto &x : th) x.join();
  std::vector<uint64_t> out;
  size_t total = 0;
  for (auto &p : part) total += p.size();
  out.reserve(total);
  for (auto &p : part) out.insert(out.end(), p.begin(), p.end());
  return out;
}

bool token_is(const std::vector<char> &buf, uint64_t at, const char *word) {
  const size_t k = std::strlen(word);
  if (at + k > buf.size()) return false;
  if (std::memcmp(buf.data() + at, word, k) != 0) return false;
  return at + k == buf.size() || is_space(buf[at + k]);
}

// unsigned decimal token; false on anything else
bool parse_u64(const std::vector<char> &buf, uint64_t at, uint64_t *out) {
  uint64_t v = 0;
  size_t p = at;
  if (p >= buf.size() || buf[p] < '0' || buf[p] > '9') return false;
  for (; p < buf.size() && !is_space(buf[p]); p++) {
    if (buf[p] < '0' || buf[p] > '9') return false;
    v = v * 10 + (uint64_t)(buf[p] - '0');
  }
  *out = v;
  return true;
}

// value token of an integer or real file, as a uint32 weight
bool parse_weight(const std::vector<char> &buf, uint64_t at, bool real,
                  uint32_t *out) {
  if (!real) {
    uint64_t v = 0;
    if (!parse_u64(buf, at, &v)) {
      // negative integers: keep the low 32 bits as strtol would
      char *end = nullptr;
      long long s = std::strtoll(buf.data() + at, &end, 10);
      if (end == buf.data() + at) return false;
      v = (uint64_t)s;
    }
    *out = (uint32_t)v;
    return true;
  }
  char *end = nullptr;
  double d = std::strtod(buf.data() + at, &end);
  if (end == buf.data() + at) return false;
  *out = (uint32_t)d;
  return true;
}

template <class F>
bool parallel_ok(uint64_t n, F fn) {
  unsigned nt = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (n < (1u << 16)) nt = 1;
  std::vector<char> ok(nt, 1);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      for (uint64_t i = n * t / nt; i < n * (t + 1) / nt; i++)
        if (!fn(i)) {
          ok[t] = 0;
          return;
        }
    });
  for (auto &x : th) x.join();
  return std::all_of(ok.begin(), ok.end(), [](char c) { return c != 0; });
}

struct Edges {
  uint32_t n = 0;
  uint64_t m = 0;
  uint32_t *src = nullptr, *dst = nullptr, *w = nullptr;
  void release() {
    std::free(src);
    std::free(dst);
    std::free(w);
    src = dst = w = nullptr;
  }
  bool alloc(bool weights) {
    const size_t b = (m ? m : 1) * sizeof(uint32_t);
    src = (uint32_t *)std::malloc(b);
    dst = (uint32_t *)std::malloc(b);
    if (weights) w = (uint32_t *)std::malloc(b);
    return src && dst && (!weights || w);
  }
};

int read_adj(const std::vector<char> &buf, bool want_w, Edges *e) {
  std::vector<uint64_t> tok = tokenize(buf, 0);
  if (tok.empty()) return SSTG_E_INVAL;
  bool weighted_file = token_is(buf, tok[0], "WeightedAdjacencyGraph");
  if (!weighted_file && !token_is(buf, tok[0], "AdjacencyGraph"))
    return SSTG_E_INVAL;
  const uint64_t len = tok.size() - 1;
  uint64_t n = 0, m = 0;
  if (len < 2 || !parse_u64(buf, tok[1], &n) || !parse_u64(buf, tok[2], &m))
    return SSTG_E_INVAL;
  if (n == 0 || m == 0 || n > 0xFFFFFFFFull) return SSTG_E_INVAL;
  const bool has_w = len == n + 2 * m + 2;
  if (len != n + m + 2 && !has_w) return SSTG_E_INVAL;
  if (weighted_file && want_w && !has_w) return SSTG_E_INVAL;
  std::vector<uint64_t> off(n + 1);
  off[n] = m;
  for (uint64_t i = 0; i < n; i++)
    if (!parse_u64(buf, tok[3 + i], &off[i])) return SSTG_E_INVAL;
  for (uint64_t i = 0; i < n; i++)
    if (off[i] > off[i + 1]) return SSTG_E_INVAL;
  e->n = (uint32_t)n;
  e->m = m;
  if (!e->alloc(want_w)) return SSTG_E_NOMEM;
  const uint64_t eb = 3 + n, wb = 3 + n + m;
  bool ok = parallel_ok(n, [&](uint64_t i) {
    for (uint64_t j = off[i]; j < off[i + 1]; j++) {
      uint64_t d = 0;
      if (!parse_u64(buf, tok[eb + j], &d) || d > 0xFFFFFFFFull) return false;
      e->src[j] = (uint32_t)i;
      e->dst[j] = (uint32_t)d;
      if (e->w) {
        if (weighted_file && has_w) {
          if (!parse_weight(buf, tok[wb + j], false, &e->w[j])) return false;
        } else {
          e->w[j] = 1;  // unweighted file read as weighted: weight 1
        }
      }
    }
    return true;
  });
  return ok ? SSTG_OK : SSTG_E_INVAL;
}

int read_mtx(std::vector<char> &buf, bool want_w, Edges *e) {
  // the first line is the banner; '%' lines after it are comments
  size_t p = 0;
  while (p < buf.size() && buf[p] != '\n') p++;
  if (p == buf.size()) return SSTG_E_INVAL;
  const size_t banner_end = p;
  p++;
  while (p < buf.size() && buf[p] == '%') {
    while (p < buf.size() && buf[p] != '\n') buf[p++] = ' ';
    if (p < buf.size()) p++;
  }
  std::vector<uint64_t> tok = tokenize(buf, 0);
  if (tok.size() < 8) return SSTG_E_INVAL;
  if (!token_is(buf, tok[0], "%%MatrixMarket") ||
      !token_is(buf, tok[1], "matrix") || !token_is(buf, tok[2], "coordinate"))
    return SSTG_E_INVAL;
  bool pattern = false, real = false;
  if (token_is(buf, tok[3], "pattern"))
    pattern = true;
  else if (token_is(buf, tok[3], "real"))
    real = true;
  else if (!token_is(buf, tok[3], "integer"))
    return SSTG_E_INVAL;
  if (!token_is(buf, tok[4], "symmetric") && !token_is(buf, tok[4], "general"))
    return SSTG_E_INVAL;
  if (tok[4] >= banner_end || tok[5] < banner_end) return SSTG_E_INVAL;
  uint64_t r = 0, c = 0, m = 0;
  if (!parse_u64(buf, tok[5], &r) || !parse_u64(buf, tok[6], &c) ||
      !parse_u64(buf, tok[7], &m))
    return SSTG_E_INVAL;
  const uint64_t per = pattern ? 2 : 3;
  if (tok.size() - 8 != m * per) return SSTG_E_INVAL;
  if (r != c || r > 0xFFFFFFFFull) return SSTG_E_INVAL;
  e->n = (uint32_t)r;
  e->m = 2 * m;
  if (!e->alloc(want_w)) return SSTG_E_NOMEM;
  bool ok = parallel_ok(m, [&](uint64_t i) {
    uint64_t s = 0, d = 0;
    if (!parse_u64(buf, tok[8 + i * per], &s) ||
        !parse_u64(buf, tok[9 + i * per], &d))
      return false;
    if (s == 0 || d == 0 || s > r || d > r) return false;  // 1-indexed
    e->src[2 * i] = (uint32_t)(s - 1);
    e->dst[2 * i] = (uint32_t)(d - 1);
    e->src[2 * i + 1] = (uint32_t)(d - 1);
    e->dst[2 * i + 1] = (uint32_t)(s - 1);
    if (e->w) {
      uint32_t v = 1;
      if (!pattern && !parse_weight(buf, tok[10 + i * per], real, &v))
        return false;
      e->w[2 * i] = e->w[2 * i + 1] = v;
    }
    return true;
  });
  return ok ? SSTG_OK : SSTG_E_INVAL;
}

}  // namespace

extern "C" {

int sstg_edges_from_file(const char *path, uint32_t *n, uint64_t *m,
                         uint32_t **src, uint32_t **dst, uint32_t **weights) {
  if (!path || !n || !m || !src || !dst) return SSTG_E_INVAL;
  *src = *dst = nullptr;
  if (weights) *weights = nullptr;
  const std::string name(path);
  const bool adj = ends_with(name, ".adj") || ends_with(name, ".adj.shuf");
  const bool mtx = ends_with(name, ".mtx");
  if (!adj && !mtx) return SSTG_E_INVAL;  // "can't read that format"
  std::vector<char> buf;
  int rc = read_file(path, &buf);
  if (rc != SSTG_OK) return rc;
  Edges e;
  rc = adj ? read_adj(buf, weights != nullptr, &e)
           : read_mtx(buf, weights != nullptr, &e);
  if (rc != SSTG_OK) {
    e.release();
    return rc;
  }
  *n = e.n;
  *m = e.m;
  *src = e.src;
  *dst = e.dst;
  if (weights) *weights = e.w;
  return SSTG_OK;
}

void sstg_host_free(void *p) { std::free(p); }

int sstg_load_file(const char *path, uint32_t flags, sstg_graph **out) {
  if (!out || (flags & ~SSTG_LOAD_WEIGHTS)) return SSTG_E_INVAL;
  *out = nullptr;
  const bool weighted = (flags & SSTG_LOAD_WEIGHTS) != 0;
  uint32_t n = 0, *s = nullptr, *d = nullptr, *w = nullptr;
  uint64_t m = 0;
  int rc = sstg_edges_from_file(path, &n, &m, &s, &d, weighted ? &w : nullptr);
  if (rc != SSTG_OK) return rc;
  sstg_graph *g = nullptr;
  rc = sstg_create(n, n, 0, &g);
  if (rc == SSTG_OK)
    rc = weighted ? sstg_insert_batch_weighted(g, s, d, w, m, 0)
                  : sstg_insert_batch(g, s, d, m, 0);
  std::free(s);
  std::free(d);
  std::free(w);
  if (rc != SSTG_OK) {
    if (g) sstg_destroy(g);
    return rc;
  }
  *out = g;
  return SSTG_OK;
}

int sstg_write_adj(sstg_graph *g, const char *path) {
  if (!g || !path) return SSTG_E_INVAL;
  uint32_t n = 0;
  int rc = sstg_num_nodes(g, &n);
  if (rc != SSTG_OK) return rc;
  std::vector<uint64_t> off((size_t)n + 1);
  rc = sstg_export_csr(g, off.data(), nullptr, 0);
  if (rc != SSTG_OK) return rc;
  std::vector<uint32_t> cols(off[n] ? off[n] : 1);
  rc = sstg_export_csr(g, off.data(), cols.data(), 0);
  if (rc != SSTG_OK) return rc;
  FILE *f = std::fopen(path, "w");
  if (!f) return SSTG_E_IO;
  bool ok = std::fprintf(f, "AdjacencyGraph\n%u\n%llu\n", n,
                         (unsigned long long)off[n]) > 0;
  for (uint32_t i = 0; ok && i < n; i++)
    ok = std::fprintf(f, "%llu\n", (unsigned long long)off[i]) > 0;
  for (uint64_t j = 0; ok && j < off[n]; j++)
    ok = std::fprintf(f, "%u\n", cols[j]) > 0;
  if (std::fclose(f) != 0) ok = false;
  return ok ? SSTG_OK : SSTG_E_IO;
}

}  // extern "C"
