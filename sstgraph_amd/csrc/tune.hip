// Autotuner for the per-graph launch settings.
//
// The vertex-pull kernels give each vertex 1 lane, 8 lanes (4 in the BFS
// steps and the batched store's BFS levels), 16 lanes (a wave in TC's
// orientation) or a whole
// 256-thread block by its degree, with one set of degree boundaries for the
// reductions (PageRank, BC, the TC orientation, the batched store's BFS
// levels) and one for single-source BFS (the bottom-up steps), and BFS
// switches between top-down and bottom-up steps by two ratios.  Where these
// pay off depends on the degree distribution and on how the rows sit in
// HBM, so they are per-graph settings (degree boundaries 8 / 64 / 256 for
// the reductions and 32 / 128 / 2048 for BFS, alpha 15, beta 18 by default:
// the tuner's picks on rMat at scales 22 and 24) that the tuner picks by
// timing the graph's own algorithms:
//  * the vertex-pull set is tuned on PageRank iterations (SSTG_TUNE_PR),
//    the BFS set and the direction switch on one BFS from a given source
//    (SSTG_TUNE_BFS), each timing the minimum of three runs after an
//    untimed one with the candidate's bins built;
//  * search: coordinate descent over a grid of candidates -- the block
//    boundary first, then the 64-lane and the 8-lane ones, then (for BFS)
//    beta and alpha -- at most three sweeps; a candidate replaces the
//    current setting only when it is at least 2% faster, so timing noise
//    does not move it.
// Results do not depend on the settings beyond the FP64 summation order of
// PageRank (a vertex's lane count fixes its shuffle tree) and the choice
// among valid BFS parents.
#include <chrono>

#include "algo.hpp"
#include "graph.hpp"

namespace sstg {

namespace {

// candidate values of each coordinate: the three degree boundaries of a
// threshold set, then alpha and beta
constexpr double kCand[5][8] = {
    {1, 2, 4, 8, 16, 32},                              // 1-lane boundary
    {16, 32, 64, 128, 256, 512},                       // 8-lane boundary
    {256, 512, 1024, 2048, 4096, 8192, 16384, 65536},  // wave boundary
    {4, 8, 15, 30, 60},                                // alpha
    {6, 9, 12, 18, 24, 36},                            // beta
};
constexpr int kNCand[5] = {6, 6, 8, 5, 6};
constexpr int kTimedRuns = 3;
constexpr int kMaxSweeps = 3;
constexpr double kMinGain = 0.98;

struct Setting {
  double v[5];  // t0, t1, t2 of the set being tuned, alpha, beta
  bool valid() const { return v[0] < v[1] && v[1] < v[2]; }
};

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch())
      .count();
}

// One algorithm's run time under a setting of its threshold set (and, for
// BFS, of its direction switch).
struct Objective {
  sstg_graph *g;
  int set;  // kBinsPull: PageRank, kBinsBfs: BFS
  uint32_t src, pr_iters;
  double *rank = nullptr;
  uint32_t *level = nullptr;

  Setting current() const {
    const uint32_t *t = g->bin_hi[set];
    return Setting{{(double)t[0], (double)t[1], (double)t[2], g->bfs_alpha,
                    g->bfs_beta}};
  }

  int apply(const Setting &st) {
    const uint32_t t[3] = {(uint32_t)st.v[0], (uint32_t)st.v[1],
                           (uint32_t)st.v[2]};
    SSTG_TRY(graph_set_thresholds(g, set, t));
    return graph_set_bfs_switch(g, st.v[3], st.v[4]);
  }

  int once() {
    if (set == kBinsPull)  // the degree bins drive the vertex pull
      SSTG_TRY(algo_pagerank(g, pr_iters, 0.85, rank, false));
    else
      SSTG_TRY(algo_bfs(g, src, nullptr, level));
    SSTG_HIP(hipStreamSynchronize(g->stream));
    return SSTG_OK;
  }

  // minimum over kTimedRuns under setting st
  int time(const Setting &st, double *ms) {
    SSTG_TRY(apply(st));
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

  // coordinate descent; leaves the best setting applied (the start one on
  // error)
  int search(Setting *start_out, Setting *best_out, double *start_ms,
             double *best_ms, uint32_t *evals) {
    const Setting start = current();
    Setting cur = start;
    double cur_ms = 0;
    int rc = time(cur, &cur_ms);
    (*evals)++;
    *start_ms = cur_ms;
    // alpha and beta only matter to BFS
    const int ncoord = set == kBinsBfs ? 5 : 3;
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
          rc = time(t, &ms);
          if (rc != SSTG_OK) break;
          (*evals)++;
          if (ms < kMinGain * cur_ms) {
            cur = t;
            cur_ms = ms;
            moved = true;
          }
        }
      }
      if (!moved) break;
    }
    const int rc2 = apply(rc == SSTG_OK ? cur : start);
    if (rc != SSTG_OK) return rc;
    *start_out = start;
    *best_out = cur;
    *best_ms = cur_ms;
    return rc2;
  }
};

}  // namespace

int graph_set_thresholds(sstg_graph *g, int set, const uint32_t t[3]) {
  if (set != kBinsPull && set != kBinsBfs) return SSTG_E_INVAL;
  if (!(t[0] < t[1] && t[1] < t[2])) return SSTG_E_INVAL;
  uint32_t *cur = g->bin_hi[set];
  if (t[0] == cur[0] && t[1] == cur[1] && t[2] == cur[2]) return SSTG_OK;
  for (int i = 0; i < 3; i++) cur[i] = t[i];
  g->bins_version[set] = ~0ull;  // rebuilt on next use
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
  sstg_tune_report r{};
  for (int set = kBinsPull; set <= kBinsBfs; set++) {
    Objective obj{g, set, src, pr_iters};
    const uint32_t *t = g->bin_hi[set];
    Setting start = obj.current(), best = start;
    double start_ms = 0, best_ms = 0;
    const bool on = (what & (set == kBinsPull ? SSTG_TUNE_PR : SSTG_TUNE_BFS));
    if (on) {
      SSTG_TRY(g->ws.get(kWsOut, g->n_rows ? g->n_rows : 1, &obj.rank));
      SSTG_TRY(g->ws.get(kWsOut2, g->n_rows ? g->n_rows : 1, &obj.level));
      SSTG_TRY(obj.search(&start, &best, &start_ms, &best_ms, &r.evaluations));
    }
    uint32_t *out = set == kBinsPull ? r.thresholds : r.bfs_thresholds;
    uint32_t *out0 =
        set == kBinsPull ? r.start_thresholds : r.start_bfs_thresholds;
    for (int i = 0; i < 3; i++) {
      out[i] = t[i];
      out0[i] = (uint32_t)start.v[i];
    }
    if (set == kBinsPull) {
      r.pr_start_ms = start_ms;
      r.pr_best_ms = best_ms;
    } else {
      r.bfs_start_ms = start_ms;
      r.bfs_best_ms = best_ms;
      r.start_alpha = start.v[3];
      r.start_beta = start.v[4];
    }
  }
  r.alpha = g->bfs_alpha;
  r.beta = g->bfs_beta;
  if (!(what & SSTG_TUNE_BFS)) {
    r.start_alpha = r.alpha;
    r.start_beta = r.beta;
  }
  if (rep) *rep = r;
  return SSTG_OK;
}

}  // namespace sstg
