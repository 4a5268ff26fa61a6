// SSTGraph.hpp -- C++ drop-in for SSTGraph's SparseMatrixV<true> and the
// EdgeMapVertexMap algorithm calls, backed by the MI355X library (sstg.h).
//
// A caller of the reference switches by including this header instead of
// "SSTGraph/SparseMatrix.hpp" and the EdgeMapVertexMap algorithm headers, and
// linking libsstg.so.  Names, argument meaning and result ownership follow the
// reference:
//   SSTGraph::SparseMatrixV<true>      include/SSTGraph/SparseMatrix.hpp:24-516
//   EdgeMapVertexMap::BFS/PR_S/BC/CC/TC/BF/TouchAll
//                                      call sites internal/test.hpp:476-618
// Algorithm results are malloc'd arrays the caller free()s, as in the
// reference harness (test.hpp:517,538,545,568,603,615,631).
//
// Differences (kept deliberately): insert_batch/remove_batch do not reorder
// the caller's array; ids out of range throw SSTGraph::sstg_error instead of
// corrupting memory; the CSR form (is_csr = true) is provided, binary or
// with one uint32_t value per edge (SparseMatrixV<true, uint32_t>, the
// weighted graph BF reads, test.hpp:699-846).  map_neighbors/map_range run on the host over a snapshot of the
// rows exported from the device (refreshed after every mutation).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "sstg.h"

namespace SSTGraph {

using el_t = uint32_t;

class sstg_error : public std::runtime_error {
 public:
  sstg_error(int code, const std::string &what)
      : std::runtime_error(what + ": " + sstg_strerror(code)), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

namespace detail {
inline void check(int rc, const char *what) {
  if (rc != SSTG_OK) throw sstg_error(rc, what);
}
}  // namespace detail

namespace detail {
template <typename... Ts> constexpr bool binary_types() {
  if constexpr (sizeof...(Ts) == 0) {
    return true;
  } else if constexpr (sizeof...(Ts) == 1) {
    return (std::is_same_v<bool, Ts> && ...);
  } else {
    return false;
  }
}
template <typename... Ts> constexpr bool weight_types() {
  if constexpr (sizeof...(Ts) == 1) {
    return (std::is_same_v<uint32_t, Ts> && ...);
  } else {
    return false;
  }
}
}  // namespace detail

template <bool is_csr_ = true, typename... Ts> class SparseMatrixV {
  static_assert(is_csr_, "sstgraph_amd stores the CSR orientation only");
  static_assert(detail::binary_types<Ts...>() || detail::weight_types<Ts...>(),
                "sstgraph_amd stores binary graphs or uint32_t edge weights");
  static constexpr bool weighted = detail::weight_types<Ts...>();

 public:
  using element_type = std::conditional_t<weighted,
                                          std::tuple<el_t, el_t, uint32_t>,
                                          std::tuple<el_t, el_t>>;
  using value_type = std::conditional_t<weighted, uint32_t, bool>;

  // SparseMatrix.hpp:73-87
  SparseMatrixV(el_t height, el_t width) : rows_(height), cols_(width) {
    detail::check(sstg_create(height, width, 0, &g_), "SparseMatrixV");
  }
  // adopts a graph handle made by the library (rewrite_graph, load)
  SparseMatrixV(sstg_graph *h, el_t height, el_t width)
      : g_(h), rows_(height), cols_(width) {}
  SparseMatrixV(const SparseMatrixV &o) : rows_(o.rows_), cols_(o.cols_) {
    detail::check(sstg_create(rows_, cols_, 0, &g_), "SparseMatrixV copy");
    copy_from(o);
  }
  SparseMatrixV(SparseMatrixV &&o) noexcept
      : g_(o.g_), rows_(o.rows_), cols_(o.cols_) {
    o.g_ = nullptr;
  }
  SparseMatrixV &operator=(const SparseMatrixV &o) {
    if (this != &o) {
      SparseMatrixV tmp(o);
      std::swap(g_, tmp.g_);
      rows_ = o.rows_;
      cols_ = o.cols_;
      snap_valid_ = false;
    }
    return *this;
  }
  SparseMatrixV &operator=(SparseMatrixV &&o) noexcept {
    if (this != &o) {
      if (g_) sstg_destroy(g_);
      g_ = o.g_;
      rows_ = o.rows_;
      cols_ = o.cols_;
      o.g_ = nullptr;
      snap_valid_ = false;
    }
    return *this;
  }
  ~SparseMatrixV() {
    if (g_) sstg_destroy(g_);
  }

  static constexpr bool is_csr() { return true; }
  static constexpr bool is_csc() { return false; }
  static constexpr bool is_binary() { return !weighted; }

  // SparseMatrix.hpp:54-59, 150-161
  [[nodiscard]] uint32_t get_line_count() const { return rows_; }
  [[nodiscard]] uint32_t get_line_width() const { return cols_; }
  [[nodiscard]] size_t num_nodes() const { return rows_; }
  [[nodiscard]] uint32_t get_rows() const { return rows_; }
  [[nodiscard]] uint32_t get_cols() const { return cols_; }
  [[nodiscard]] uint32_t getDegree(uint32_t node) const {
    refresh();
    return (uint32_t)(snap_off_[node + 1] - snap_off_[node]);
  }

  // SparseMatrix.hpp:61-71
  [[nodiscard]] uint64_t common_neighbors(uint32_t node1, uint32_t node2,
                                          bool early_end = false) const {
    uint64_t out = 0;
    detail::check(sstg_common_neighbors(g_, node1, node2, early_end ? 1 : 0,
                                        &out),
                  "common_neighbors");
    return out;
  }

  // SparseMatrix.hpp:720-726
  [[nodiscard]] bool has(el_t row, el_t col) const {
    int out = 0;
    detail::check(sstg_has(g_, row, col, &out), "has");
    return out != 0;
  }

  // SparseMatrix.hpp:727-735: the edge's weight (1 if it was inserted
  // without one), 0 if absent; has(row, col) on a binary graph
  [[nodiscard]] value_type value(el_t row, el_t col) const {
    if constexpr (weighted) {
      uint32_t w = 0;
      detail::check(sstg_edge_weights(g_, &row, &col, 1, &w, 0), "value");
      return w;
    } else {
      return has(row, col);
    }
  }

  // SparseMatrix.hpp:536-554 (single insert keeps self-loops)
  void insert(element_type e) {
    if constexpr (weighted) {
      el_t r = std::get<0>(e), c = std::get<1>(e);
      uint32_t w = std::get<2>(e);
      detail::check(sstg_insert_batch_weighted(g_, &r, &c, &w, 1,
                                               SSTG_KEEP_SELF_LOOPS),
                    "insert");
      snap_valid_ = false;
    } else {
      insert(std::get<0>(e), std::get<1>(e));
    }
  }
  void insert(el_t row, el_t col) {
    detail::check(sstg_insert(g_, row, col), "insert");
    snap_valid_ = false;
  }
  void insert(el_t row, el_t col, uint32_t w) {
    static_assert(weighted, "a binary graph takes no value");
    insert(element_type(row, col, w));
  }
  void remove(el_t row, el_t col) {
    detail::check(sstg_remove(g_, row, col), "remove");
    snap_valid_ = false;
  }

  // SparseMatrix.hpp:556-643: rows become old U {(x,y) : x != y}
  void insert_batch(element_type *edges, uint64_t n, bool semi_sorted = false) {
    batch(edges, n, semi_sorted, true);
  }
  // SparseMatrix.hpp:645-718: rows become old \ batch
  void remove_batch(std::tuple<el_t, el_t> *edges, uint64_t n,
                    bool semi_sorted = false) {
    batch(edges, n, semi_sorted, false);
  }
  // weighted structure-of-arrays form
  void insert_batch(const el_t *src, const el_t *dst, const uint32_t *w,
                    uint64_t n, uint32_t flags = 0) {
    static_assert(weighted, "a binary graph takes no values");
    detail::check(sstg_insert_batch_weighted(g_, src, dst, w, n, flags),
                  "insert_batch");
    snap_valid_ = false;
  }
  // Structure-of-arrays forms (host or, with SSTG_DEVICE_PTRS, device).
  void insert_batch(const el_t *src, const el_t *dst, uint64_t n,
                    uint32_t flags = 0) {
    detail::check(sstg_insert_batch(g_, src, dst, n, flags), "insert_batch");
    snap_valid_ = false;
  }
  void remove_batch(const el_t *src, const el_t *dst, uint64_t n,
                    uint32_t flags = 0) {
    detail::check(sstg_remove_batch(g_, src, dst, n, flags), "remove_batch");
    snap_valid_ = false;
  }

  // SparseMatrix.hpp:518-534, 436-454
  [[nodiscard]] uint64_t M() const {
    uint64_t m = 0;
    detail::check(sstg_num_edges(g_, &m), "M");
    return m;
  }
  [[nodiscard]] uint64_t get_memory_size() const {
    uint64_t b = 0;
    detail::check(sstg_memory_size(g_, &b), "get_memory_size");
    return b;
  }
  uint64_t touch_all_sum(uint64_t const *const order = nullptr) const {
    (void)order;  // the sum does not depend on the row order
    uint64_t s = 0;
    detail::check(sstg_touch_all_sum(g_, &s), "touch_all_sum");
    return s;
  }

  // SparseMatrix.hpp:218-253: f(row, neighbour) in ascending order; with
  // F::no_early_exit == false a true return stops the scan.
  template <class F, size_t... Is>
  void map_neighbors(uint64_t i, F &f, void *d, bool parallel) const {
    (void)d;
    (void)parallel;
    refresh();
    constexpr bool all = std::remove_reference_t<F>::no_early_exit;
    for (uint64_t e = snap_off_[i]; e < snap_off_[i + 1]; e++) {
      if constexpr (all) {
        f(i, (uint64_t)snap_cols_[e]);
      } else {
        if (f(i, (uint64_t)snap_cols_[e])) break;
      }
    }
  }
  template <class F, size_t... Is>
  void map_neighbors_early_exit(uint64_t i, F &&f, void *d, bool p) const {
    (void)d;
    (void)p;
    refresh();
    for (uint64_t e = snap_off_[i]; e < snap_off_[i + 1]; e++)
      if (f(i, (uint64_t)snap_cols_[e])) break;
  }
  template <class F, size_t... Is>
  void map_neighbors_no_early_exit(uint64_t i, F &&f, void *d, bool p) const {
    (void)d;
    (void)p;
    refresh();
    for (uint64_t e = snap_off_[i]; e < snap_off_[i + 1]; e++)
      f(i, (uint64_t)snap_cols_[e]);
  }
  // SparseMatrix.hpp:306-315
  template <class F, size_t... Is>
  void map_range(F &f, uint64_t start_node, uint64_t end_node, void *d) const {
    (void)d;
    refresh();
    for (uint64_t i = start_node; i < end_node; i++)
      for (uint64_t e = snap_off_[i]; e < snap_off_[i + 1]; e++)
        f(i, (uint64_t)snap_cols_[e]);
  }

  // SparseMatrix.hpp:737-749: new label of vertex i by descending
  // (degree, id), computed on the device
  [[nodiscard]] std::vector<uint64_t> degree_order_map() const {
    std::vector<uint32_t> m(rows_);
    detail::check(sstg_degree_order_map(g_, m.data(), 0), "degree_order_map");
    return std::vector<uint64_t>(m.begin(), m.end());
  }
  // SparseMatrix.hpp:751-766: every edge relabelled on the device, self-
  // loops kept (the reference re-inserts edge by edge with insert())
  [[nodiscard]] SparseMatrixV rewrite_graph(
      std::vector<uint64_t> new_label_map) const {
    // the reference re-inserts two-element tuples: binary graphs only
    static_assert(!weighted, "rewrite_graph relabels binary graphs");
    if (new_label_map.size() != rows_)
      throw sstg_error(SSTG_E_INVAL, "rewrite_graph: one label per row");
    std::vector<uint32_t> m(rows_);
    for (uint64_t i = 0; i < rows_; i++) {
      if (new_label_map[i] > 0xFFFFFFFFull)
        throw sstg_error(SSTG_E_RANGE, "rewrite_graph");
      m[i] = (uint32_t)new_label_map[i];
    }
    sstg_graph *h = nullptr;
    detail::check(sstg_rewrite_graph(g_, m.data(), 0, &h), "rewrite_graph");
    return SparseMatrixV(h, rows_, cols_);
  }

  // SparseMatrix.hpp:340-426: the reference's layout-free lines (its
  // per-tinyset byte-width tables describe a CPU layout this store does
  // not have), then the HBM store's own figures
  void print_statistics() const {
    sstg_stats st;
    detail::check(sstg_statistics(g_, &st), "print_statistics");
    std::printf("matrix stored in csr\n");
    std::printf("there are %llu elements\n", (unsigned long long)st.edges);
    std::printf("there are %u rows and %u columns\n", st.rows, st.cols);
    std::printf("The graph uses %llu bytes of memory\n",
                (unsigned long long)st.memory_bytes);
    std::printf("%llu rows are empty; the largest row has %llu elements\n",
                (unsigned long long)st.empty_rows,
                (unsigned long long)st.max_degree);
    std::printf("slots hold %llu elements; the pool has %llu (%llu below "
                "the tail, %llu abandoned), %llu rebuilds\n",
                (unsigned long long)st.slot_capacity,
                (unsigned long long)st.pool_capacity,
                (unsigned long long)st.pool_used,
                (unsigned long long)st.pool_garbage,
                (unsigned long long)st.rebuilds);
    std::printf("count of outgoing edges which uses x bits\n");
    unsigned long long total = 0;
    for (int i = 0; i < 33; i++) {
      total += st.dest_bits[i];
      std::printf("%d, %llu, total so far = %llu\n", i,
                  (unsigned long long)st.dest_bits[i], total);
    }
  }

  sstg_graph *handle() const { return g_; }

 private:
  template <class E>
  void batch(E *edges, uint64_t n, bool semi_sorted, bool ins) {
    std::vector<el_t> s(n), d(n);
    std::vector<uint32_t> w;
    constexpr bool valued = std::tuple_size_v<E> == 3;
    if constexpr (valued) w.resize(n);
    for (uint64_t i = 0; i < n; i++) {  // std::get, never reinterpret_cast:
      s[i] = std::get<0>(edges[i]);       // libstdc++ stores get<1> first
      d[i] = std::get<1>(edges[i]);
      if constexpr (valued) w[i] = std::get<2>(edges[i]);
    }
    uint32_t flags = semi_sorted ? SSTG_SEMI_SORTED : 0u;
    int rc;
    if constexpr (valued)
      rc = sstg_insert_batch_weighted(g_, s.data(), d.data(), w.data(), n,
                                      flags);
    else
      rc = ins ? sstg_insert_batch(g_, s.data(), d.data(), n, flags)
               : sstg_remove_batch(g_, s.data(), d.data(), n, flags);
    detail::check(rc, ins ? "insert_batch" : "remove_batch");
    snap_valid_ = false;
  }
  void copy_from(const SparseMatrixV &o) {
    o.refresh();
    std::vector<el_t> s(o.snap_cols_.size());
    for (uint64_t i = 0; i < o.rows_; i++)
      for (uint64_t e = o.snap_off_[i]; e < o.snap_off_[i + 1]; e++)
        s[e] = (el_t)i;
    if constexpr (weighted) {
      std::vector<uint32_t> w(s.size());
      detail::check(sstg_edge_weights(o.g_, s.data(), o.snap_cols_.data(),
                                      s.size(), w.data(), 0),
                    "SparseMatrixV copy");
      insert_batch(s.data(), o.snap_cols_.data(), w.data(), s.size(),
                   SSTG_KEEP_SELF_LOOPS);
    } else {
      insert_batch(s.data(), o.snap_cols_.data(), s.size(),
                   SSTG_KEEP_SELF_LOOPS);
    }
  }
  void refresh() const {
    if (snap_valid_) return;
    snap_off_.assign((size_t)rows_ + 1, 0);
    detail::check(sstg_export_csr(g_, snap_off_.data(), nullptr, 0), "export");
    snap_cols_.resize(snap_off_.back());
    detail::check(sstg_export_csr(g_, snap_off_.data(), snap_cols_.data(), 0),
                  "export");
    snap_valid_ = true;
  }

  sstg_graph *g_ = nullptr;
  uint32_t rows_ = 0, cols_ = 0;
  mutable bool snap_valid_ = false;
  mutable std::vector<uint64_t> snap_off_;
  mutable std::vector<el_t> snap_cols_;
};

// helpers.hpp:604-631: the edges of a .adj / .adj.shuf / .mtx file as a
// malloc'd array the caller free()s (.mtx entries come in both directions,
// 1-indexed ids shifted to 0).  value_type bool gives (row, col) tuples,
// uint32_t gives (row, col, weight) with weight 1 from an unweighted file
// (helpers.hpp:557-566).  Throws sstg_error where the reference prints and
// exits.
template <typename value_type = bool>
std::conditional_t<std::is_same_v<value_type, bool>, std::tuple<el_t, el_t>,
                   std::tuple<el_t, el_t, value_type>> *
get_edges_from_file(const std::string &filename, uint64_t *edge_count,
                    uint32_t *node_count, bool print = true) {
  static_assert(std::is_same_v<value_type, bool> ||
                    std::is_same_v<value_type, uint32_t>,
                "sstgraph_amd reads binary graphs or uint32_t weights");
  constexpr bool valued = !std::is_same_v<value_type, bool>;
  using E = std::conditional_t<valued, std::tuple<el_t, el_t, value_type>,
                               std::tuple<el_t, el_t>>;
  (void)print;
  uint32_t *s = nullptr, *d = nullptr, *w = nullptr;
  detail::check(sstg_edges_from_file(filename.c_str(), node_count, edge_count,
                                     &s, &d, valued ? &w : nullptr),
                "get_edges_from_file");
  auto *out = (E *)std::malloc((*edge_count ? *edge_count : 1) * sizeof(E));
  if (!out) {
    sstg_host_free(s);
    sstg_host_free(d);
    sstg_host_free(w);
    throw sstg_error(SSTG_E_NOMEM, "get_edges_from_file");
  }
  for (uint64_t i = 0; i < *edge_count; i++) {
    if constexpr (valued)
      new (&out[i]) E(s[i], d[i], w[i]);
    else
      new (&out[i]) E(s[i], d[i]);
  }
  sstg_host_free(s);
  sstg_host_free(d);
  sstg_host_free(w);
  return out;
}

}  // namespace SSTGraph

namespace EdgeMapVertexMap {

namespace detail {
template <class T> T *alloc(size_t n) {
  T *p = static_cast<T *>(std::malloc((n ? n : 1) * sizeof(T)));
  if (!p) throw std::bad_alloc();
  return p;
}
}  // namespace detail

// BFS(g, src) (test.hpp:486): parent array, parent[src] == src, -1 unreached.
template <class G> int32_t *BFS(G &g, uint32_t src) {
  int32_t *out = detail::alloc<int32_t>(g.num_nodes());
  int rc = sstg_bfs(g.handle(), src, out, nullptr, 0);
  if (rc) {
    std::free(out);
    throw SSTGraph::sstg_error(rc, "BFS");
  }
  return out;
}

// PR_S<T>(g, iters) (test.hpp:527): T = double (FP64) or float (FP32 ranks
// with Kahan-compensated sums).
template <class T, class G> T *PR_S(G &g, uint64_t iters) {
  static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>);
  T *out = detail::alloc<T>(g.num_nodes());
  int rc;
  if constexpr (std::is_same_v<T, double>)
    rc = sstg_pagerank_f64(g.handle(), (uint32_t)iters, 0.85, out, 0);
  else
    rc = sstg_pagerank_f32k(g.handle(), (uint32_t)iters, 0.85, out, 0);
  if (rc) {
    std::free(out);
    throw SSTGraph::sstg_error(rc, "PR_S");
  }
  return out;
}

// BC(g, src) (test.hpp:547): single-source Brandes dependencies.
template <class G> double *BC(G &g, uint32_t src) {
  double *out = detail::alloc<double>(g.num_nodes());
  int rc = sstg_bc(g.handle(), src, out, 0);
  if (rc) {
    std::free(out);
    throw SSTGraph::sstg_error(rc, "BC");
  }
  return out;
}

// CC(g) (test.hpp:571): label = smallest vertex id of the component.
template <class G> uint32_t *CC(G &g) {
  uint32_t *out = detail::alloc<uint32_t>(g.num_nodes());
  int rc = sstg_cc(g.handle(), out, 0);
  if (rc) {
    std::free(out);
    throw SSTGraph::sstg_error(rc, "CC");
  }
  return out;
}

// TC(g) (test.hpp:606): number of triangles.
template <class G> uint64_t TC(G &g) {
  uint64_t out = 0;
  SSTGraph::detail::check(sstg_tc(g.handle(), &out), "TC");
  return out;
}

// BF(g, src) (test.hpp:618, 811): shortest distances, INT32_MAX/2 where
// unreached; edge weights on SparseMatrixV<true, uint32_t>, 1 per edge on a
// binary graph.
template <class G> int32_t *BF(G &g, uint32_t src) {
  int32_t *out = detail::alloc<int32_t>(g.num_nodes());
  int rc = sstg_bf(g.handle(), src, out, 0);
  if (rc) {
    std::free(out);
    throw SSTGraph::sstg_error(rc, "BF");
  }
  return out;
}

// TouchAll(g) (test.hpp:476): sum of all destination ids.
template <class G> uint64_t TouchAll(G &g) {
  uint64_t out = 0;
  SSTGraph::detail::check(sstg_touch_all(g.handle(), &out), "TouchAll");
  return out;
}

}  // namespace EdgeMapVertexMap
