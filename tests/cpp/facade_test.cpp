// Exercises the C++ drop-in (include/sstgraph_amd/SSTGraph.hpp) the way the
// reference harness uses SparseMatrixV<true> and EdgeMapVertexMap
// (include/SSTGraph/internal/test.hpp:396-697), checking every result
// against the CPU oracle.  Exit code 0 = pass, 77 = no GPU (skipped).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <cstdlib>
#include <tuple>
#include <vector>

#include "sstg_oracle.h"
#include "sstgraph_amd/SSTGraph.hpp"

using SSTGraph::el_t;

#define CHECK(cond)                                                        \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      return 1;                                                            \
    }                                                                      \
  } while (0)

struct CountF {  // map_neighbors functor with the reference's contract
  static constexpr bool no_early_exit = true;
  uint64_t *sum;
  bool operator()(uint64_t, uint64_t d) {
    *sum += d;
    return false;
  }
};

int main(int argc, char **argv) {
  int ndev = 0;
  sstg_device_count(&ndev);
  if (ndev == 0) {
    std::printf("no GPU: skipped\n");
    return 77;
  }
  const uint32_t scale = 12, n = 1u << scale;
  const uint64_t cnt = 16ull * n;
  std::vector<uint32_t> s(cnt), d(cnt);
  CHECK(orc_rmat(n, 7, 0.5, 0.1, 0.1, 0, cnt, s.data(), d.data()) == 0);
  // AoS tuples, symmetrised, as get_edges_from_file hands them over
  std::vector<std::tuple<el_t, el_t>> edges;
  for (uint64_t i = 0; i < cnt; i++) {
    edges.emplace_back(s[i], d[i]);
    edges.emplace_back(d[i], s[i]);
  }
  SSTGraph::SparseMatrixV<true> g(n, n);
  // build in chunks like test.hpp:443-450
  const uint64_t chunk = 10000;
  uint64_t i = 0;
  for (; i + chunk < edges.size(); i += chunk) g.insert_batch(edges.data() + i, chunk);
  g.insert_batch(edges.data() + i, edges.size() - i);

  orc_graph *og = orc_graph_create(n, n);
  std::vector<uint32_t> ss(2 * cnt), dd(2 * cnt);
  for (uint64_t k = 0; k < edges.size(); k++) {
    ss[k] = std::get<0>(edges[k]);
    dd[k] = std::get<1>(edges[k]);
  }
  CHECK(orc_insert_batch(og, ss.data(), dd.data(), ss.size()) == 0);
  std::vector<uint64_t> off(n + 1);
  orc_export_csr(og, off.data(), nullptr);
  std::vector<uint32_t> cols(off[n]);
  orc_export_csr(og, off.data(), cols.data());

  CHECK(g.M() == off[n]);
  CHECK(g.num_nodes() == n);
  // contents row by row through map_range and map_neighbors
  std::vector<uint32_t> got;
  auto collect = [&](uint64_t, uint64_t dst) { got.push_back((uint32_t)dst); };
  g.map_range(collect, 0, n, nullptr);
  CHECK(got == cols);
  for (uint32_t v = 0; v < n; v += 97) {
    CHECK(g.getDegree(v) == off[v + 1] - off[v]);
    uint64_t sum = 0, want = 0;
    CountF f{&sum};
    g.map_neighbors(v, f, nullptr, false);
    for (uint64_t e = off[v]; e < off[v + 1]; e++) want += cols[e];
    CHECK(sum == want);
  }
  CHECK(g.touch_all_sum() == orc_touch_all_sum(og));
  CHECK(g.has(std::get<0>(edges[5]), std::get<1>(edges[5])) ||
        std::get<0>(edges[5]) == std::get<1>(edges[5]));
  for (uint32_t u = 1; u < n; u += 211)
    CHECK(g.common_neighbors(u, n - u, true) ==
          orc_common_neighbors(off.data(), cols.data(), u, n - u, 1));

  // BFS -> depths exactly as test.hpp:497-509 derives them
  const uint32_t src = n - 1;
  int32_t *par = EdgeMapVertexMap::BFS(g, src);
  std::vector<uint32_t> olev(n);
  orc_bfs(n, off.data(), cols.data(), src, olev.data(), nullptr);
  for (uint32_t j = 0; j < n; j++) {
    uint32_t depth = UINT32_MAX;
    if (par[j] >= 0) {
      depth = 0;
      int32_t cur = (int32_t)j;
      while (cur != par[cur]) {
        depth++;
        cur = par[cur];
      }
    }
    CHECK(depth == olev[j]);
  }
  std::free(par);

  double *pr = EdgeMapVertexMap::PR_S<double>(g, 10);
  std::vector<double> opr(n);
  orc_pagerank(n, off.data(), cols.data(), 10, 0.85, opr.data());
  for (uint32_t j = 0; j < n; j++)
    CHECK(std::fabs(pr[j] - opr[j]) <= 1e-9 * std::fabs(opr[j]));
  std::free(pr);
  float *prf = EdgeMapVertexMap::PR_S<float>(g, 10);
  for (uint32_t j = 0; j < n; j++)
    CHECK(std::fabs(prf[j] - opr[j]) <= 1e-5 * std::fabs(opr[j]));
  std::free(prf);

  double *bc = EdgeMapVertexMap::BC(g, src);
  std::vector<double> obc(n);
  orc_bc(n, off.data(), cols.data(), src, obc.data());
  for (uint32_t j = 0; j < n; j++)
    CHECK(std::fabs(bc[j] - obc[j]) <= 1e-9 * std::fabs(obc[j]) + 1e-12);
  std::free(bc);

  uint32_t *cc = EdgeMapVertexMap::CC(g);
  std::vector<uint32_t> occ(n);
  orc_cc(n, off.data(), cols.data(), occ.data());
  for (uint32_t j = 0; j < n; j++) CHECK(cc[j] == occ[j]);
  std::free(cc);

  const uint64_t tc = EdgeMapVertexMap::TC(g);
  CHECK(tc == orc_tc(n, off.data(), cols.data()));
  int32_t *bf = EdgeMapVertexMap::BF(g, src);
  std::vector<int32_t> obf(n);
  orc_bf(n, off.data(), cols.data(), nullptr, src, obf.data());
  for (uint32_t j = 0; j < n; j++) CHECK(bf[j] == obf[j]);
  std::free(bf);
  CHECK(EdgeMapVertexMap::TouchAll(g) == orc_touch_all_sum(og));

  // degree-order relabel keeps the triangle count and edge count; the map is
  // the reference's definition (SparseMatrix.hpp:737-749)
  const std::vector<uint64_t> dmap = g.degree_order_map();
  {
    std::vector<std::pair<uint64_t, uint64_t>> lm(n);
    for (uint32_t j = 0; j < n; j++) lm[j] = {off[j + 1] - off[j], j};
    std::sort(lm.begin(), lm.end(), std::greater<>());
    for (uint32_t j = 0; j < n; j++) CHECK(dmap[lm[j].second] == j);
  }
  SSTGraph::SparseMatrixV<true> h = g.rewrite_graph(dmap);
  CHECK(h.M() == g.M());
  CHECK(EdgeMapVertexMap::TC(h) == tc);
  // copies are deep
  SSTGraph::SparseMatrixV<true> c2(g);
  c2.remove(std::get<0>(edges[0]), std::get<1>(edges[0]));
  CHECK(g.M() == off[n]);

  // the update loop of test.hpp:668-677 on the same graph
  std::vector<std::tuple<el_t, el_t>> es;
  std::vector<uint32_t> bs(5000), bd(5000);
  orc_rmat(n / 2, 99, 0.5, 0.1, 0.1, 0, 5000, bs.data(), bd.data());
  for (uint32_t k = 0; k < 5000; k++) es.emplace_back(bs[k], bd[k]);
  g.insert_batch(es.data(), es.size());
  g.remove_batch(es.data(), es.size());
  orc_insert_batch(og, bs.data(), bd.data(), 5000);
  orc_remove_batch(og, bs.data(), bd.data(), 5000);
  CHECK(g.M() == orc_num_edges(og));

  // the weighted form BF reads (test.hpp:699-846): SparseMatrixV<true,
  // uint32_t> built from (row, col, value) tuples, weights a function of the
  // unordered pair so duplicates agree
  {
    auto wf = [](uint32_t a, uint32_t b) {
      const uint32_t lo = std::min(a, b), hi = std::max(a, b);
      return (uint32_t)((lo * 31u + hi) % 97u + 1u);
    };
    std::vector<std::tuple<el_t, el_t, uint32_t>> wedges;
    for (const auto &e : edges)
      wedges.emplace_back(std::get<0>(e), std::get<1>(e),
                          wf(std::get<0>(e), std::get<1>(e)));
    SSTGraph::SparseMatrixV<true, uint32_t> wg(n, n);
    static_assert(!SSTGraph::SparseMatrixV<true, uint32_t>::is_binary());
    wg.insert_batch(wedges.data(), wedges.size());
    CHECK(wg.M() == off[n]);
    std::vector<int32_t> aw(off[n]);
    for (uint32_t r = 0; r < n; r++)
      for (uint64_t e = off[r]; e < off[r + 1]; e++)
        aw[e] = (int32_t)wf(r, cols[e]);
    const uint32_t r0 = std::get<0>(edges[5]), c0 = std::get<1>(edges[5]);
    if (r0 != c0) CHECK(wg.value(r0, c0) == wf(r0, c0));
    CHECK(wg.value(n - 1, n - 1) == 0);
    wg.insert(r0, c0, 1000u);  // a present key's value is overwritten
    if (r0 != c0) CHECK(wg.value(r0, c0) == 1000u);
    wg.insert(r0, c0, wf(r0, c0));
    int32_t *wbf = EdgeMapVertexMap::BF(wg, src);
    std::vector<int32_t> owbf(n);
    orc_bf(n, off.data(), cols.data(), aw.data(), src, owbf.data());
    for (uint32_t j = 0; j < n; j++) CHECK(wbf[j] == owbf[j]);
    std::free(wbf);
    // copies carry the weights
    SSTGraph::SparseMatrixV<true, uint32_t> wc(wg);
    if (r0 != c0) CHECK(wc.value(r0, c0) == wf(r0, c0));
  }

  bool threw = false;
  try {
    g.insert(n, 0);
  } catch (const SSTGraph::sstg_error &e) {
    threw = e.code() == SSTG_E_RANGE;
  }
  CHECK(threw);
  orc_graph_destroy(og);

  // a graph file read as the harness reads it (test.hpp:414-449)
  if (argc > 1) {
    uint64_t fm = 0;
    uint32_t fn = 0;
    auto *fe = SSTGraph::get_edges_from_file(argv[1], &fm, &fn);
    SSTGraph::SparseMatrixV<true> fg(fn, fn);
    fg.insert_batch(fe, fm);
    std::free(fe);
    CHECK(fn == 28 && fm == 420 && fg.M() == 420);
    CHECK(EdgeMapVertexMap::TC(fg) == 420);
    fg.print_statistics();
    // the same file read weighted: a pattern file gives weight 1 throughout
    auto *fw = SSTGraph::get_edges_from_file<uint32_t>(argv[1], &fm, &fn);
    SSTGraph::SparseMatrixV<true, uint32_t> fwg(fn, fn);
    fwg.insert_batch(fw, fm);
    for (uint64_t k = 0; k < fm; k++) CHECK(std::get<2>(fw[k]) == 1u);
    std::free(fw);
    int32_t *d1 = EdgeMapVertexMap::BF(fwg, 0);
    int32_t *d2 = EdgeMapVertexMap::BF(fg, 0);
    for (uint32_t j = 0; j < fn; j++) CHECK(d1[j] == d2[j]);
    std::free(d1);
    std::free(d2);
  }
  std::printf("facade ok: n=%u m=%llu tc=%llu\n", n,
              (unsigned long long)off[n], (unsigned long long)tc);
  return 0;
}
