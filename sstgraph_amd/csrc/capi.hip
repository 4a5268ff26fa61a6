// C ABI of sstgraph_amd (include/sstg.h): argument checking, host/device
// staging of arrays, and the runtime pieces shared by every call (workspace,
// kernel profiler, error notes).
#include <cstdlib>

#include "algo.hpp"
#include "graph.hpp"

namespace sstg {

static thread_local char g_last_error[256] = {0};

void note_hip_error(hipError_t e, const char *what, const char *file,
                    int line) {
  std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s (%s:%d)",
                hipGetErrorString(e), what, file, line);
  if (std::getenv("SSTG_VERBOSE"))
    std::fprintf(stderr, "sstgraph_amd: %s\n", g_last_error);
}

// ------------------------------------------------------------ profiler
Profiler &profiler() {
  static Profiler p;
  return p;
}

hipEvent_t Profiler::get_event() {
  if (!free_events.empty()) {
    hipEvent_t e = free_events.back();
    free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void Profiler::begin(const char *, hipStream_t s, hipEvent_t *a) {
  std::lock_guard<std::mutex> lk(mu);
  *a = get_event();
  if (*a) hipEventRecord(*a, s);
}

void Profiler::end(const char *name, hipStream_t s, hipEvent_t a) {
  std::lock_guard<std::mutex> lk(mu);
  if (!a) return;
  hipEvent_t b = get_event();
  if (!b) return;
  hipEventRecord(b, s);
  pending.push_back(Rec{name, a, b});
  if (pending.size() > 4096) {
    // keep the event pool bounded
    for (auto &r : pending) {
      hipEventSynchronize(r.b);
      float ms = 0;
      hipEventElapsedTime(&ms, r.a, r.b);
      bool found = false;
      for (auto &x : acc)
        if (x.name == r.name) {
          x.ms += ms;
          x.n++;
          found = true;
          break;
        }
      if (!found) acc.push_back(Acc{r.name, (double)ms, 1});
      free_events.push_back(r.a);
      free_events.push_back(r.b);
    }
    pending.clear();
  }
}

void Profiler::flush() {
  std::lock_guard<std::mutex> lk(mu);
  for (auto &r : pending) {
    hipEventSynchronize(r.b);
    float ms = 0;
    hipEventElapsedTime(&ms, r.a, r.b);
    bool found = false;
    for (auto &x : acc)
      if (x.name == r.name) {
        x.ms += ms;
        x.n++;
        found = true;
        break;
      }
    if (!found) acc.push_back(Acc{r.name, (double)ms, 1});
    free_events.push_back(r.a);
    free_events.push_back(r.b);
  }
  pending.clear();
}

void Profiler::reset() {
  flush();
  std::lock_guard<std::mutex> lk(mu);
  acc.clear();
}

// ----------------------------------------------------------- workspace
int Workspace::get(int slot, size_t bytes, void **out) {
  if (slot < 0) return SSTG_E_INVAL;
  if (arena) {
    if ((size_t)slot >= arena_slots.size()) arena_slots.resize(slot + 1);
    Slot &a = arena_slots[slot];
    if (a.bytes < bytes) {
      const size_t want = (bytes + 255) & ~(size_t)255;
      if (want > arena_bytes - arena_used) return SSTG_E_NOMEM;
      a.p = arena + arena_used;
      a.bytes = want;
      arena_used += want;
      generation++;
    }
    *out = a.p;
    return SSTG_OK;
  }
  if ((size_t)slot >= slots.size()) slots.resize(slot + 1);
  Slot &s = slots[slot];
  if (s.bytes < bytes) {
    if (s.p) {
      if (stream) hipStreamSynchronize(stream);
      hipFree(s.p);
      total -= s.bytes;
      s.p = nullptr;
      s.bytes = 0;
    }
    size_t want = bytes + bytes / 4;
    want = (want + 4095) & ~(size_t)4095;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      want = (bytes + 4095) & ~(size_t)4095;
      e = hipMalloc(&p, want);
      if (e != hipSuccess) {
        note_hip_error(e, "workspace hipMalloc", __FILE__, __LINE__);
        return SSTG_E_NOMEM;
      }
    }
    s.p = p;
    s.bytes = want;
    total += want;
    generation++;
  }
  *out = s.p;
  return SSTG_OK;
}

void Workspace::release_all() {
  if (stream) hipStreamSynchronize(stream);
  for (auto &s : slots)
    if (s.p) hipFree(s.p);
  slots.clear();
  total = 0;
  generation++;
}

static bool device_ok(int *dev) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return false;
  }
  if (hipGetDevice(dev) != hipSuccess) return false;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, *dev) != hipSuccess) return false;
  return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

// Output staging: runs fn(device_ptr) and copies the result to host memory
// unless the caller passed device memory.
template <class T, class F>
static int with_output(sstg_graph *g, T *out, uint64_t count, uint32_t flags,
                       int slot, F fn) {
  if (!out) return fn((T *)nullptr);
  if (flags & SSTG_DEVICE_PTRS) return fn(out);
  T *d = nullptr;
  SSTG_TRY(g->ws.get(slot, count, &d));
  SSTG_TRY(fn(d));
  SSTG_HIP(hipMemcpyAsync(out, d, count * sizeof(T), hipMemcpyDeviceToHost,
                          g->stream));
  SSTG_HIP(hipStreamSynchronize(g->stream));
  return SSTG_OK;
}

}  // namespace sstg

using namespace sstg;

extern "C" {

const char *sstg_strerror(int status) {
  switch (status) {
    case SSTG_OK:
      return "ok";
    case SSTG_E_INVAL:
      return "invalid argument";
    case SSTG_E_RANGE:
      return "vertex id out of range";
    case SSTG_E_NOMEM:
      return "out of device memory";
    case SSTG_E_DEVICE:
      return g_last_error[0] ? g_last_error : "HIP runtime error";
    case SSTG_E_NODEVICE:
      return "no gfx950 device";
    case SSTG_E_IO:
      return "file could not be opened, read or written";
    case SSTG_E_UNSUPPORTED:
      return "unsupported";
    case SSTG_E_TOO_LARGE:
      return "input larger than the fixed scratch of an allocation-free call";
    default:
      return "unknown status";
  }
}

int sstg_abi_version(void) { return SSTG_ABI_VERSION; }

int sstg_device_count(int *out) {
  if (!out) return SSTG_E_INVAL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  *out = n;
  return SSTG_OK;
}

int sstg_create(uint32_t n_rows, uint32_t n_cols, uint32_t flags,
                sstg_graph **out) {
  (void)flags;
  if (!out) return SSTG_E_INVAL;
  *out = nullptr;
  int dev = 0;
  if (!device_ok(&dev)) return SSTG_E_NODEVICE;
  sstg_graph *g = new (std::nothrow) sstg_graph();
  if (!g) return SSTG_E_NOMEM;
  g->device = dev;
  g->n_rows = n_rows;
  g->n_cols = n_cols;
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) !=
      hipSuccess) {
    delete g;
    return SSTG_E_DEVICE;
  }
  g->ws.stream = g->stream;
  int rc = graph_alloc(g);
  if (rc != SSTG_OK) {
    graph_free(g);
    hipStreamDestroy(g->stream);
    delete g;
    return rc;
  }
  *out = g;
  return SSTG_OK;
}

int sstg_destroy(sstg_graph *g) {
  if (!g) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  graph_free(g);
  hipStreamDestroy(g->stream);
  delete g;
  return SSTG_OK;
}

int sstg_insert_batch(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                      uint64_t n, uint32_t flags) {
  if (!g || (n && (!src || !dst))) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  return graph_update(g, src, dst, n, true,
                      !(flags & SSTG_KEEP_SELF_LOOPS),
                      (flags & SSTG_DEVICE_PTRS) != 0);
}

// Runs fn(src, dst, w, count) on device copies of chunks of a host or
// device batch (w may be null).
extern "C++" template <class F>
static int device_chunks(sstg_graph *g, const uint32_t *src,
                         const uint32_t *dst, const uint32_t *w, uint64_t n,
                         bool device_ptrs, F fn) {
  const uint64_t kChunkPairs = 1ull << 28;
  for (uint64_t b = 0; b < n; b += kChunkPairs) {
    const uint64_t c = std::min(kChunkPairs, n - b);
    if (device_ptrs) {
      SSTG_TRY(fn(src + b, dst + b, w ? w + b : nullptr, c));
      continue;
    }
    uint32_t *ds = nullptr, *dd = nullptr, *dw = nullptr;
    SSTG_TRY(g->ws.get(kWsWtSrc, c, &ds));
    SSTG_TRY(g->ws.get(kWsWtDst, c, &dd));
    SSTG_HIP(hipMemcpyAsync(ds, src + b, c * 4, hipMemcpyHostToDevice,
                            g->stream));
    SSTG_HIP(hipMemcpyAsync(dd, dst + b, c * 4, hipMemcpyHostToDevice,
                            g->stream));
    if (w) {
      SSTG_TRY(g->ws.get(kWsWtVal, c, &dw));
      SSTG_HIP(hipMemcpyAsync(dw, w + b, c * 4, hipMemcpyHostToDevice,
                              g->stream));
    }
    SSTG_TRY(fn(ds, dd, dw, c));
  }
  return SSTG_OK;
}

int sstg_insert_batch_weighted(sstg_graph *g, const uint32_t *src,
                               const uint32_t *dst, const uint32_t *w,
                               uint64_t n, uint32_t flags) {
  if (!g || (n && (!src || !dst || !w))) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  const bool dev = (flags & SSTG_DEVICE_PTRS) != 0;
  const bool keep_self = (flags & SSTG_KEEP_SELF_LOOPS) != 0;
  // the edges first: an id out of range leaves graph and weights unchanged
  SSTG_TRY(graph_update(g, src, dst, n, true, !keep_self, dev));
  return device_chunks(g, src, dst, w, n, dev,
                       [&](const uint32_t *s, const uint32_t *d,
                           const uint32_t *ww, uint64_t c) {
                         return weights_upsert(g, s, d, ww, c, keep_self);
                       });
}

int sstg_remove_batch(sstg_graph *g, const uint32_t *src, const uint32_t *dst,
                      uint64_t n, uint32_t flags) {
  if (!g || (n && (!src || !dst))) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  const bool dev = (flags & SSTG_DEVICE_PTRS) != 0;
  SSTG_TRY(graph_update(g, src, dst, n, false, false, dev));
  if (!g->wt_keys) return SSTG_OK;
  return device_chunks(g, src, dst, nullptr, n, dev,
                       [&](const uint32_t *s, const uint32_t *d,
                           const uint32_t *, uint64_t c) {
                         return weights_erase(g, s, d, c);
                       });
}

int sstg_edge_weights(sstg_graph *g, const uint32_t *rows,
                      const uint32_t *cols, uint64_t n, uint32_t *out,
                      uint32_t flags) {
  if (!g || (n && (!rows || !cols || !out))) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  if (flags & SSTG_DEVICE_PTRS) {
    SSTG_TRY(weights_get(g, rows, cols, n, out));
    SSTG_HIP(hipStreamSynchronize(g->stream));
    return SSTG_OK;
  }
  uint32_t *dr = nullptr, *dc = nullptr, *dout = nullptr;
  SSTG_TRY(g->ws.get(kWsWtSrc, n, &dr));
  SSTG_TRY(g->ws.get(kWsWtDst, n, &dc));
  SSTG_TRY(g->ws.get(kWsWtVal, n, &dout));
  SSTG_HIP(hipMemcpyAsync(dr, rows, n * 4, hipMemcpyHostToDevice, g->stream));
  SSTG_HIP(hipMemcpyAsync(dc, cols, n * 4, hipMemcpyHostToDevice, g->stream));
  SSTG_TRY(weights_get(g, dr, dc, n, dout));
  SSTG_HIP(hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, g->stream));
  SSTG_HIP(hipStreamSynchronize(g->stream));
  return SSTG_OK;
}

int sstg_set_degree_thresholds(sstg_graph *g, const uint32_t t[3]) {
  if (!g || !t) return SSTG_E_INVAL;
  return graph_set_thresholds(g, kBinsPull, t);
}

int sstg_get_degree_thresholds(sstg_graph *g, uint32_t t[3]) {
  if (!g || !t) return SSTG_E_INVAL;
  for (int i = 0; i < 3; i++) t[i] = g->bin_hi[kBinsPull][i];
  return SSTG_OK;
}

int sstg_set_bfs_degree_thresholds(sstg_graph *g, const uint32_t t[3]) {
  if (!g || !t) return SSTG_E_INVAL;
  return graph_set_thresholds(g, kBinsBfs, t);
}

int sstg_get_bfs_degree_thresholds(sstg_graph *g, uint32_t t[3]) {
  if (!g || !t) return SSTG_E_INVAL;
  for (int i = 0; i < 3; i++) t[i] = g->bin_hi[kBinsBfs][i];
  return SSTG_OK;
}

int sstg_set_bfs_switch(sstg_graph *g, double alpha, double beta) {
  if (!g) return SSTG_E_INVAL;
  return graph_set_bfs_switch(g, alpha, beta);
}

int sstg_get_bfs_switch(sstg_graph *g, double *alpha, double *beta) {
  if (!g || !alpha || !beta) return SSTG_E_INVAL;
  *alpha = g->bfs_alpha;
  *beta = g->bfs_beta;
  return SSTG_OK;
}

int sstg_set_pagerank_segments(sstg_graph *g, uint32_t bits) {
  if (!g) return SSTG_E_INVAL;
  if (bits != 0 && (bits < 8 || bits > 30)) return SSTG_E_INVAL;
  g->pr_seg_bits = bits;
  return SSTG_OK;
}

int sstg_get_pagerank_segments(sstg_graph *g, uint32_t *bits) {
  if (!g || !bits) return SSTG_E_INVAL;
  *bits = g->pr_seg_bits;
  return SSTG_OK;
}

int sstg_autotune(sstg_graph *g, uint32_t what, uint32_t src,
                  uint32_t pr_iters, sstg_tune_report *out) {
  if (!g) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  return graph_autotune(g, what, src, pr_iters, out);
}

int sstg_insert(sstg_graph *g, uint32_t row, uint32_t col) {
  if (!g) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  return graph_update(g, &row, &col, 1, true, false, false);
}

int sstg_remove(sstg_graph *g, uint32_t row, uint32_t col) {
  return sstg_remove_batch(g, &row, &col, 1, 0);
}

int sstg_last_update_stats(sstg_graph *g, sstg_update_stats *out) {
  if (!g || !out) return SSTG_E_INVAL;
  *out = g->last;
  return SSTG_OK;
}

int sstg_has(sstg_graph *g, uint32_t row, uint32_t col, int *out) {
  if (!g || !out) return SSTG_E_INVAL;
  if (row >= g->n_rows || col >= g->n_cols) return SSTG_E_RANGE;
  hipSetDevice(g->device);
  uint8_t h = 0;
  SSTG_TRY(graph_has_batch(g, &row, &col, 1, &h, false));
  *out = h;
  return SSTG_OK;
}

int sstg_has_batch(sstg_graph *g, const uint32_t *rows, const uint32_t *cols,
                   uint64_t n, uint8_t *out, uint32_t flags) {
  if (!g || (n && (!rows || !cols || !out))) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  return graph_has_batch(g, rows, cols, n, out,
                         (flags & SSTG_DEVICE_PTRS) != 0);
}

int sstg_num_nodes(sstg_graph *g, uint32_t *out) {
  if (!g || !out) return SSTG_E_INVAL;
  *out = g->n_rows;
  return SSTG_OK;
}

int sstg_num_edges(sstg_graph *g, uint64_t *out) {
  if (!g || !out) return SSTG_E_INVAL;
  *out = g->m;
  return SSTG_OK;
}

int sstg_degrees(sstg_graph *g, uint32_t *out, uint32_t flags) {
  if (!g || !out) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  SSTG_HIP(hipMemcpyAsync(out, g->deg, (uint64_t)g->n_rows * sizeof(uint32_t),
                          (flags & SSTG_DEVICE_PTRS) ? hipMemcpyDeviceToDevice
                                                     : hipMemcpyDeviceToHost,
                          g->stream));
  SSTG_HIP(hipStreamSynchronize(g->stream));
  return SSTG_OK;
}

int sstg_export_csr(sstg_graph *g, uint64_t *offsets, uint32_t *cols,
                    uint32_t flags) {
  if (!g || !offsets) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  return graph_export(g, offsets, cols, (flags & SSTG_DEVICE_PTRS) != 0);
}

int sstg_memory_size(sstg_graph *g, uint64_t *out) {
  if (!g || !out) return SSTG_E_INVAL;
  uint64_t n = g->n_rows;
  *out = sizeof(sstg_graph) + n * (sizeof(uint64_t) + 2 * sizeof(uint32_t)) +
         g->pool_cap * sizeof(uint32_t);
  return SSTG_OK;
}

int sstg_degree_order_map(sstg_graph *g, uint32_t *map, uint32_t flags) {
  if (!g || !map) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  return with_output<uint32_t>(g, map, g->n_rows, flags, kWsOut,
                               [&](uint32_t *d) {
                                 return graph_degree_order_map(g, d);
                               });
}

int sstg_rewrite_graph(sstg_graph *g, const uint32_t *map, uint32_t flags,
                       sstg_graph **out) {
  if (!g || !map || !out) return SSTG_E_INVAL;
  *out = nullptr;
  // map relabels rows and columns alike, so it has n_rows == n_cols entries
  if (g->n_rows != g->n_cols) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  const uint32_t *dmap = map;
  if (!(flags & SSTG_DEVICE_PTRS)) {
    uint32_t *d = nullptr;
    SSTG_TRY(g->ws.get(kWsOut, g->n_rows, &d));
    SSTG_HIP(hipMemcpyAsync(d, map, (uint64_t)g->n_rows * 4,
                            hipMemcpyHostToDevice, g->stream));
    dmap = d;
  }
  sstg_graph *ng = nullptr;
  SSTG_TRY(sstg_create(g->n_rows, g->n_cols, 0, &ng));
  int rc = graph_rewrite(g, dmap, ng);
  if (rc != SSTG_OK) {
    sstg_destroy(ng);
    return rc;
  }
  *out = ng;
  return SSTG_OK;
}

int sstg_statistics(sstg_graph *g, sstg_stats *out) {
  if (!g || !out) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  *out = sstg_stats{};
  out->edges = g->m;
  out->rows = g->n_rows;
  out->cols = g->n_cols;
  SSTG_TRY(sstg_memory_size(g, &out->memory_bytes));
  out->pool_capacity = g->pool_cap;
  out->pool_used = g->pool_tail;
  out->pool_garbage = g->garbage;
  out->rebuilds = g->rebuilds;
  return graph_statistics(g, out->dest_bits, &out->empty_rows,
                          &out->max_degree, &out->slot_capacity);
}

int sstg_touch_all_sum(sstg_graph *g, uint64_t *out) {
  if (!g || !out) return SSTG_E_INVAL;
  hipSetDevice(g->device);
  return graph_touch_sum(g, out);
}

int sstg_touch_all(sstg_graph *g, uint64_t *out) {
  return sstg_touch_all_sum(g, out);
}

int sstg_common_neighbors(sstg_graph *g, uint32_t u, uint32_t v,
                          int early_end, uint64_t *out) {
  if (!g || !out) return SSTG_E_INVAL;
  if (u >= g->n_rows || v >= g->n_rows) return SSTG_E_RANGE;
  hipSetDevice(g->device);
  return algo_common_neighbors(g, u, v, early_end, out);
}

int sstg_bfs(sstg_graph *g, uint32_t src, int32_t *parent, uint32_t *level,
             uint32_t flags) {
  if (!g) return SSTG_E_INVAL;
  if (g->n_rows != g->n_cols) return SSTG_E_INVAL;  // graphs are square
  if (src >= g->n_rows) return SSTG_E_RANGE;
  hipSetDevice(g->device);
  return with_output<int32_t>(
      g, parent, g->n_rows, flags, kWsOut, [&](int32_t *dp) {
        return with_output<uint32_t>(
            g, level, g->n_rows, flags, kWsOut2,
            [&](uint32_t *dl) { return algo_bfs(g, src, dp, dl); });
      });
}

int sstg_pagerank_f64(sstg_graph *g, uint32_t iters, double damping,
                      double *out, uint32_t flags) {
  if (!g || !out) return SSTG_E_INVAL;
  if (g->n_rows != g->n_cols) return SSTG_E_INVAL;  // graphs are square
  hipSetDevice(g->device);
  return with_output<double>(g, out, g->n_rows, flags, kWsOut, [&](double *d) {
    return algo_pagerank(g, iters, damping, d);
  });
}

int sstg_pagerank_f32k(sstg_graph *g, uint32_t iters, double damping,
                       float *out, uint32_t flags) {
  if (!g || !out) return SSTG_E_INVAL;
  if (g->n_rows != g->n_cols) return SSTG_E_INVAL;  // graphs are square
  hipSetDevice(g->device);
  return with_output<float>(g, out, g->n_rows, flags, kWsOut, [&](float *d) {
    return algo_pagerank_f32k(g, iters, damping, d);
  });
}

int sstg_bc(sstg_graph *g, uint32_t src, double *out, uint32_t flags) {
  if (!g || !out) return SSTG_E_INVAL;
  if (g->n_rows != g->n_cols) return SSTG_E_INVAL;  // graphs are square
  if (src >= g->n_rows) return SSTG_E_RANGE;
  hipSetDevice(g->device);
  return with_output<double>(g, out, g->n_rows, flags, kWsOut,
                             [&](double *d) { return algo_bc(g, src, d); });
}

int sstg_cc(sstg_graph *g, uint32_t *out, uint32_t flags) {
  if (!g || !out) return SSTG_E_INVAL;
  if (g->n_rows != g->n_cols) return SSTG_E_INVAL;  // graphs are square
  hipSetDevice(g->device);
  return with_output<uint32_t>(g, out, g->n_rows, flags, kWsOut,
                               [&](uint32_t *d) { return algo_cc(g, d); });
}

int sstg_tc(sstg_graph *g, uint64_t *out) {
  if (!g || !out) return SSTG_E_INVAL;
  if (g->n_rows != g->n_cols) return SSTG_E_INVAL;  // graphs are square
  hipSetDevice(g->device);
  return algo_tc(g, out);
}

int sstg_bf(sstg_graph *g, uint32_t src, int32_t *out, uint32_t flags) {
  if (!g || !out) return SSTG_E_INVAL;
  if (g->n_rows != g->n_cols) return SSTG_E_INVAL;  // graphs are square
  if (src >= g->n_rows) return SSTG_E_RANGE;
  hipSetDevice(g->device);
  return with_output<int32_t>(g, out, g->n_rows, flags, kWsOut,
                              [&](int32_t *d) {
                                return g->wt_keys ? algo_bf_weighted(g, src, d)
                                                  : algo_bf(g, src, d);
                              });
}

int sstg_rmat(uint32_t n, uint32_t seed, double a, double b, double c,
              uint64_t first, uint64_t count, uint32_t *src, uint32_t *dst,
              uint32_t flags) {
  if (count && (!src || !dst)) return SSTG_E_INVAL;
  if (flags & ~(SSTG_DEVICE_PTRS | SSTG_RMAT_NORMALIZED)) return SSTG_E_INVAL;
  const bool norm = (flags & SSTG_RMAT_NORMALIZED) != 0;
  int dev = 0;
  if (!device_ok(&dev)) return SSTG_E_NODEVICE;
  hipStream_t s = nullptr;
  if (flags & SSTG_DEVICE_PTRS) {
    SSTG_TRY(rmat_generate(n, seed, a, b, c, first, count, src, dst, s, norm));
    SSTG_HIP(hipStreamSynchronize(s));
    return SSTG_OK;
  }
  uint32_t *ds = nullptr, *dd = nullptr;
  SSTG_HIP(hipMalloc(&ds, (count ? count : 1) * sizeof(uint32_t)));
  SSTG_HIP(hipMalloc(&dd, (count ? count : 1) * sizeof(uint32_t)));
  int rc = rmat_generate(n, seed, a, b, c, first, count, ds, dd, s, norm);
  if (rc == SSTG_OK && count) {
    hipMemcpy(src, ds, count * sizeof(uint32_t), hipMemcpyDeviceToHost);
    hipMemcpy(dst, dd, count * sizeof(uint32_t), hipMemcpyDeviceToHost);
  }
  hipFree(ds);
  hipFree(dd);
  return rc;
}

int sstg_er(uint32_t n, uint32_t seed, uint64_t first, uint64_t count,
            uint32_t *src, uint32_t *dst, uint32_t flags) {
  if (count && (!src || !dst)) return SSTG_E_INVAL;
  int dev = 0;
  if (!device_ok(&dev)) return SSTG_E_NODEVICE;
  if (flags & SSTG_DEVICE_PTRS) {
    SSTG_TRY(er_generate(n, seed, first, count, src, dst, nullptr));
    SSTG_HIP(hipStreamSynchronize(nullptr));
    return SSTG_OK;
  }
  uint32_t *ds = nullptr, *dd = nullptr;
  SSTG_HIP(hipMalloc(&ds, (count ? count : 1) * sizeof(uint32_t)));
  SSTG_HIP(hipMalloc(&dd, (count ? count : 1) * sizeof(uint32_t)));
  int rc = er_generate(n, seed, first, count, ds, dd, nullptr);
  if (rc == SSTG_OK && count) {
    hipMemcpy(src, ds, count * sizeof(uint32_t), hipMemcpyDeviceToHost);
    hipMemcpy(dst, dd, count * sizeof(uint32_t), hipMemcpyDeviceToHost);
  }
  hipFree(ds);
  hipFree(dd);
  return rc;
}

int sstg_device_alloc(uint64_t bytes, void **out) {
  if (!out) return SSTG_E_INVAL;
  SSTG_HIP(hipMalloc(out, bytes ? bytes : 1));
  return SSTG_OK;
}

int sstg_device_free(void *p) {
  SSTG_HIP(hipFree(p));
  return SSTG_OK;
}

// The copies return once the data is in place for every stream: a graph's
// stream does not wait for the null stream, and hipMemcpy may return before a
// device-to-device copy (or the device side of a copy from pageable host
// memory) has finished, so an update started right after a copy into its
// input could read a part not yet written.
int sstg_memcpy_h2d(void *dst, const void *src, uint64_t bytes) {
  SSTG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  SSTG_HIP(hipDeviceSynchronize());
  return SSTG_OK;
}

int sstg_memcpy_d2h(void *dst, const void *src, uint64_t bytes) {
  SSTG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return SSTG_OK;
}

int sstg_memcpy_d2d(void *dst, const void *src, uint64_t bytes) {
  SSTG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  SSTG_HIP(hipDeviceSynchronize());
  return SSTG_OK;
}

int sstg_device_synchronize(void) {
  SSTG_HIP(hipDeviceSynchronize());
  return SSTG_OK;
}

int sstg_mem_info(uint64_t *free_bytes, uint64_t *total_bytes) {
  size_t f = 0, t = 0;
  SSTG_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return SSTG_OK;
}

int sstg_profile_enable(int on) {
  profiler().on = on != 0;
  return SSTG_OK;
}

int sstg_profile_reset(void) {
  profiler().reset();
  return SSTG_OK;
}

int sstg_profile_query(const char *name, double *total_ms,
                       uint64_t *launches) {
  if (!name) return SSTG_E_INVAL;
  Profiler &p = profiler();
  p.flush();
  std::lock_guard<std::mutex> lk(p.mu);
  double ms = 0;
  uint64_t n = 0;
  for (auto &a : p.acc)
    if (a.name.find(name) != std::string::npos) {
      ms += a.ms;
      n += a.n;
    }
  if (total_ms) *total_ms = ms;
  if (launches) *launches = n;
  return SSTG_OK;
}

int sstg_profile_dump(char *buf, uint64_t buf_bytes) {
  if (!buf || buf_bytes == 0) return SSTG_E_INVAL;
  Profiler &p = profiler();
  p.flush();
  std::lock_guard<std::mutex> lk(p.mu);
  std::string s;
  for (auto &a : p.acc) {
    char line[256];
    std::snprintf(line, sizeof(line), "%s %llu %.6f\n", a.name.c_str(),
                  (unsigned long long)a.n, a.ms);
    s += line;
  }
  std::snprintf(buf, buf_bytes, "%s", s.c_str());
  return SSTG_OK;
}

}  // extern "C"
