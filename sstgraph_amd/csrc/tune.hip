// Degree-threshold autotuner.
//
// The vertex-pull kernels (PageRank, the bottom-up BFS steps, the TC
// orientation, the batched BFS levels) give each vertex 1, 8 or 64 lanes or
// a whole 256-thread block by its degree.  Where the boundaries pay off
// depends on the degree distribution and on how the rows sit in HBM, so they
// are per-graph settings (defaults 8 / 64 / 2048) that the tuner picks by
// timing the graph's own algorithms:
//  * objective: PageRank iterations (SSTG_TUNE_PR) and/or one BFS from a
//    given source (SSTG_TUNE_BFS), each the minimum of three timed runs
//    after an untimed one with the candidate's bins built;
//  * search: coordinate descent over a grid of candidate boundaries, the
//    block boundary first, then the 64-lane and the 8-lane ones, at most
//    three sweeps; a candidate replaces the current one only when it is at
//    least 2% faster, so timing noise does not move the boundaries.
// Results do not depend on the boundaries beyond the FP64 summation order
// of PageRank (a vertex's lane count fixes its shuffle tree).
#include <chrono>

#include "algo.hpp"
#include "graph.hpp"

namespace sstg {

namespace {

constexpr uint32_t kCand0[] = {1, 2, 4, 8, 16, 32};
constexpr uint32_t kCand1[] = {16, 32, 64, 128, 256, 512};
constexpr uint32_t kCand2[] = {256, 512, 1024, 2048, 4096, 8192, 16384, 65536};
constexpr int kTimedRuns = 3;
constexpr int kMaxSweeps = 3;
constexpr double kMinGain = 0.98;

bool valid(const uint32_t t[3]) { return t[0] < t[1] && t[1] < t[2]; }

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

  // minimum over kTimedRuns of the objective with boundaries t
  int time(const uint32_t t[3], double *ms) {
    SSTG_TRY(graph_set_thresholds(g, t));
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
  if (!valid(t)) return SSTG_E_INVAL;
  if (t[0] == g->bin_hi[0] && t[1] == g->bin_hi[1] && t[2] == g->bin_hi[2])
    return SSTG_OK;
  for (int i = 0; i < 3; i++) g->bin_hi[i] = t[i];
  g->bins_version = ~0ull;  // rebuilt on next use
  g->tune_epoch++;
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
  uint32_t cur[3] = {g->bin_hi[0], g->bin_hi[1], g->bin_hi[2]};
  const uint32_t start[3] = {cur[0], cur[1], cur[2]};
  double cur_ms = 0;
  uint32_t evals = 0;
  int rc = obj.time(cur, &cur_ms);
  evals++;
  const double start_ms = cur_ms;
  static const uint32_t *cands[3] = {kCand0, kCand1, kCand2};
  static const int ncand[3] = {(int)(sizeof(kCand0) / 4),
                               (int)(sizeof(kCand1) / 4),
                               (int)(sizeof(kCand2) / 4)};
  for (int sweep = 0; rc == SSTG_OK && sweep < kMaxSweeps; sweep++) {
    bool moved = false;
    for (int c = 2; c >= 0 && rc == SSTG_OK; c--) {
      for (int k = 0; k < ncand[c]; k++) {
        uint32_t t[3] = {cur[0], cur[1], cur[2]};
        t[c] = cands[c][k];
        if (!valid(t) || t[c] == cur[c]) continue;
        double ms = 0;
        rc = obj.time(t, &ms);
        if (rc != SSTG_OK) break;
        evals++;
        if (ms < kMinGain * cur_ms) {
          for (int i = 0; i < 3; i++) cur[i] = t[i];
          cur_ms = ms;
          moved = true;
        }
      }
    }
    if (!moved) break;
  }
  // leave the graph on the best boundaries found (the start ones on error)
  const int rc2 = graph_set_thresholds(g, rc == SSTG_OK ? cur : start);
  if (rc != SSTG_OK) return rc;
  if (rc2 != SSTG_OK) return rc2;
  if (rep) {
    for (int i = 0; i < 3; i++) {
      rep->thresholds[i] = cur[i];
      rep->start_thresholds[i] = start[i];
    }
    rep->start_ms = start_ms;
    rep->best_ms = cur_ms;
    rep->evaluations = evals;
  }
  return SSTG_OK;
}

}  // namespace sstg
