// Autotuner for the per-graph launch settings.
//
// The vertex-pull kernels (PageRank, the bottom-up BFS steps, the TC
// orientation, the batched BFS levels) give each vertex 1, 8 or 64 lanes or
// a whole 256-thread block by its degree, and BFS switches between top-down
// and bottom-up steps by two ratios.  Where these pay off depends on the
// degree distribution and on how the rows sit in HBM, so they are per-graph
// settings (degree boundaries 8 / 64 / 2048, alpha 15, beta 18 by default)
// that the tuner picks by timing the graph's own algorithms:
//  * objective: PageRank iterations (SSTG_TUNE_PR) and/or one BFS from a
//    given source (SSTG_TUNE_BFS), each the minimum of three timed runs
//    after an untimed one with the candidate's bins built;
//  * search: coordinate descent over a grid of candidates -- the block
//    boundary first, then the 64-lane and the 8-lane ones, then (with a BFS
//    objective) beta and alpha -- at most three sweeps; a candidate replaces
//    the current setting only when it is at least 2% faster, so timing noise
//    does not move it.
// Results do not depend on the settings beyond the FP64 summation order of
// PageRank (a vertex's lane count fixes its shuffle tree) and the choice
// among valid BFS parents.
#include <chrono>

#include "algo.hpp"
#include "graph.hpp"

namespace sstg {

namespace {

constexpr double kCand[5][8] = {
    {1, 2, 4, 8, 16, 32},                         // 1-lane boundary
    {16, 32, 64, 128, 256, 512},                  // 8-lane boundary
    {256, 512, 1024, 2048, 4096, 8192, 16384, 65536},  // wave boundary
    {4, 8, 15, 30, 60},                           // alpha
    {6, 9, 12, 18, 24, 36},                       // beta
};
constexpr int kNCand[5] = {6, 6, 8, 5, 6};
constexpr int kTimedRuns = 3;
constexpr int kMaxSweeps = 3;
constexpr double kMinGain = 0.98;

struct Setting {
  double v[5];  // t0, t1, t2, alpha, beta
  bool valid() const { return v[0] < v[1] && v[1] < v[2]; }
};

Setting current(const sstg_graph *g) {
  return Setting{{(double)g->bin_hi[0], (double)g->bin_hi[1],
                  (double)g->bin_hi[2], g->bfs_alpha, g->bfs_beta}};
}

int apply(sstg_graph *g, const Setting &st) {
  const uint32_t t[3] = {(uint32_t)st.v[0], (uint32_t)st.v[1],
                         (uint32_t)st.v[2]};
  SSTG_TRY(graph_set_thresholds(g, t));
  return graph_set_bfs_switch(g, st.v[3], st.v[4]);
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch())
      .count();
}

struct Objective {
  sstg_graph *g;
  uint32_t what, src, pr_iters;
  double *rank = nullptr;
  uint32_t *level = nullptr;

  int once() {
    if (what & SSTG_TUNE_PR)
      SSTG_TRY(algo_pagerank(g, pr_iters, 0.85, rank));
    if (what & SSTG_TUNE_BFS) SSTG_TRY(algo_bfs(g, src, nullptr, level));
    SSTG_HIP(hipStreamSynchronize(g->stream));
    return SSTG_OK;
  }

  // minimum over kTimedRuns of the objective under setting st
  int time(const Setting &st, double *ms) {
    SSTG_TRY(apply(g, st));
    SSTG_TRY(once());  // builds the bins, warms the caches
    double best = 1e300;
    for (int r = 0; r < kTimedRuns; r++) {
      const double t0 = now_ms();
      SSTG_TRY(once());
      const double dt = now_ms() - t0;
      if (dt < best) best = dt;
    }
    *ms = best;
    return SSTG_OK;
  }
};

}  // namespace

int graph_set_thresholds(sstg_graph *g, const uint32_t t[3]) {
  if (!(t[0] < t[1] && t[1] < t[2])) return SSTG_E_INVAL;
  if (t[0] == g->bin_hi[0] && t[1] == g->bin_hi[1] && t[2] == g->bin_hi[2])
    return SSTG_OK;
  for (int i = 0; i < 3; i++) g->bin_hi[i] = t[i];
  g->bins_version = ~0ull;  // rebuilt on next use
  g->tune_epoch++;
  return SSTG_OK;
}

int graph_set_bfs_switch(sstg_graph *g, double alpha, double beta) {
  if (!(alpha > 0) || !(beta > 0)) return SSTG_E_INVAL;  // NaN fails too
  g->bfs_alpha = alpha;
  g->bfs_beta = beta;
  return SSTG_OK;
}

int graph_autotune(sstg_graph *g, uint32_t what, uint32_t src,
                   uint32_t pr_iters, sstg_tune_report *rep) {
  if (what == 0 || (what & ~(SSTG_TUNE_PR | SSTG_TUNE_BFS))) return SSTG_E_INVAL;
  if ((what & SSTG_TUNE_BFS) && src >= g->n_rows) return SSTG_E_RANGE;
  if (pr_iters == 0) pr_iters = 2;
  Objective obj{g, what, src, pr_iters};
  SSTG_TRY(g->ws.get(kWsOut, g->n_rows ? g->n_rows : 1, &obj.rank));
  SSTG_TRY(g->ws.get(kWsOut2, g->n_rows ? g->n_rows : 1, &obj.level));
  const Setting start = current(g);
  Setting cur = start;
  double cur_ms = 0;
  uint32_t evals = 0;
  int rc = obj.time(cur, &cur_ms);
  evals++;
  const double start_ms = cur_ms;
  // alpha and beta only matter to a BFS objective
  const int ncoord = (what & SSTG_TUNE_BFS) ? 5 : 3;
  static const int order[5] = {2, 1, 0, 4, 3};
  for (int sweep = 0; rc == SSTG_OK && sweep < kMaxSweeps; sweep++) {
    bool moved = false;
    for (int oi = 0; oi < 5 && rc == SSTG_OK; oi++) {
      const int c = order[oi];
      if (c >= ncoord) continue;
      for (int k = 0; k < kNCand[c]; k++) {
        Setting t = cur;
        t.v[c] = kCand[c][k];
        if (!t.valid() || t.v[c] == cur.v[c]) continue;
        double ms = 0;
        rc = obj.time(t, &ms);
        if (rc != SSTG_OK) break;
        evals++;
        if (ms < kMinGain * cur_ms) {
          cur = t;
          cur_ms = ms;
          moved = true;
        }
      }
    }
    if (!moved) break;
  }
  // leave the graph on the best setting found (the start one on error)
  const int rc2 = apply(g, rc == SSTG_OK ? cur : start);
  if (rc != SSTG_OK) return rc;
  if (rc2 != SSTG_OK) return rc2;
  if (rep) {
    for (int i = 0; i < 3; i++) {
      rep->thresholds[i] = (uint32_t)cur.v[i];
      rep->start_thresholds[i] = (uint32_t)start.v[i];
    }
    rep->alpha = cur.v[3];
    rep->beta = cur.v[4];
    rep->start_alpha = start.v[3];
    rep->start_beta = start.v[4];
    rep->start_ms = start_ms;
    rep->best_ms = cur_ms;
    rep->evaluations = evals;
  }
  return SSTG_OK;
}

}  // namespace sstg
