// Serving loop over one device-resident graph: continuous batching of
// update requests, concurrent queries, admission by free HBM.
//
// No reference counterpart: the reference has no server.  Its update loop
// (include/SSTGraph/internal/test.hpp:636-690) applies one batch at a time
// and its algorithm calls (test.hpp:476-618) run one at a time.
//
// Semantics.  Requests are ordered by arrival.  A query sees every update
// that arrived before it and none that arrived after it, so its result is
// the one it would get alone on the graph at that point of the stream.
//  * Updates are writers.  All updates of the same kind that are waiting at
//    the head of the queue are applied as ONE device batch (continuous
//    batching, SSTG_SERVE_COALESCE): a run of inserts is their union, a run
//    of removes their difference, exactly as applying them one by one.
//  * Without SSTG_SERVE_SNAPSHOTS an update waits until the earlier queries
//    are done.  With it, an update that finds queries still reading the
//    store first copies the store (device to device, about 2 x 4 bytes per
//    stored edge of traffic) and applies the batch to the copy; the running
//    queries keep the old arrays as an immutable snapshot, which returns to
//    a spare-buffer cache when its last reader finishes.  The copy counts
//    against the HBM budget while the snapshot lives.
//  * Queries (BFS, PR, CC, TC) are readers: each runs on a free slot (its
//    own HIP stream, workspace and read-only view of the store), so queries
//    that arrive while others are on the GPU overlap with them.  They are
//    admitted in arrival order while a slot is free and their device memory
//    estimate fits the HBM budget; otherwise they wait.  With
//    SSTG_SERVE_ONE_HEAVY at most one PageRank or TC runs at a time, and
//    lighter queries admitted behind a waiting one pass it (readers in any
//    order see the same store).
//  * With SSTG_SERVE_SNAPSHOTS a query still waiting when an update behind
//    it could run is pinned to the current store arrays, so the update goes
//    ahead on a copy and the query later runs on the version it arrived at.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "algo.hpp"
#include "graph.hpp"

namespace sstg {
namespace {

double now_s() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// One set of store arrays that queries read.  The live set aliases the
// graph's arrays; a retired set owns the arrays it holds (the graph moved
// on to a copy) and is recycled when its last reader is done.
struct StoreSet {
  uint64_t *off = nullptr;
  uint32_t *deg = nullptr, *cap = nullptr, *pool = nullptr;
  uint64_t pool_cap = 0;  // elements
  uint32_t readers = 0;   // running or pinned queries
  bool owned = false;
  uint64_t bytes = 0;     // device bytes held while owned
  // the graph as it stood when the set was taken
  uint32_t n_rows = 0, n_cols = 0;
  uint64_t m = 0, pool_tail = 0, version = 0;
  bool sym_known = false, symmetric = false;
};

struct Request {
  uint64_t ticket = 0;
  sstg_request spec{};
  std::vector<uint32_t> src, dst;  // update payload (host copy)
  double t_arrive = 0, t_start = 0, t_done = 0;
  uint64_t est_bytes = 0;
  uint64_t coalesced = 1;  // requests applied in the same device batch
  int status = SSTG_OK;
  bool done = false;
  void *result = nullptr;  // device result (queries)
  uint64_t result_bytes = 0;
  StoreSet *set = nullptr; // store arrays the query reads
  uint64_t scalar = 0;     // TC count
};

// A read-only view of the shared store with its own stream and workspace.
struct Slot {
  sstg_graph view;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv;
  Request *req = nullptr;
  bool stop = false;
};

bool is_update(uint32_t kind) {
  return kind == SSTG_REQ_INSERT || kind == SSTG_REQ_REMOVE;
}

bool is_heavy(uint32_t kind) {
  return kind == SSTG_REQ_PR || kind == SSTG_REQ_TC;
}

}  // namespace
}  // namespace sstg

struct sstg_server {
  sstg_graph *g = nullptr;
  std::vector<sstg::Slot *> slots;
  uint64_t budget = 0, reserved = 0;
  int coalesce = 1;
  int snapshots = 0;
  int one_heavy = 0;
  uint32_t heavy_running = 0;
  uint32_t owned_live = 0;  // store copies some query still reads
  sstg::StoreSet *live = nullptr;        // set the next admitted query reads
  std::vector<sstg::StoreSet> spare;     // recycled store arrays
  std::vector<sstg::StoreSet> to_free;   // extra spares for the reaper
  std::thread reaper;
  // the update run on the writer thread (one at a time)
  std::thread writer;
  std::vector<sstg::Request *> write_run;
  std::vector<uint32_t> write_src, write_dst;
  bool writing = false;
  uint64_t write_ticket = 0;
  std::multimap<uint64_t, void *> spare_results;  // bytes -> device buffer
  uint64_t detaches = 0;
  std::deque<sstg::Request *> pending;
  std::map<uint64_t, sstg::Request *> all;
  std::mutex mu;
  std::condition_variable cv;       // scheduler wake-up
  std::condition_variable done_cv;  // request completion
  std::thread sched;
  bool stop = false;
  uint64_t next_ticket = 1;
  uint64_t running = 0;  // queries on slots
  uint64_t device_batches = 0;
  // SSTG_SERVE_MAINTAIN: while every edge has its reverse, the writer logs
  // for each store version the triangles closed minus broken since the
  // chain began (tc_chain: version -> cumulative delta); one full count of
  // any version in the chain (the anchor) then gives the count of every
  // version in it
  hipStream_t fetch_stream = nullptr;  // result copies to the host
  std::mutex fetch_mu;
  int maintain = 0;
  std::map<uint64_t, int64_t> tc_chain;
  bool tc_anchor = false;
  uint64_t tc_anchor_count = 0;
  int64_t tc_anchor_delta = 0;
  uint64_t tc_from_count = 0;   // TC queries answered from the chain
};

namespace sstg {
namespace {

// Device bytes a query needs beyond the shared store (result + scratch).
uint64_t estimate(const sstg_graph *g, const sstg_request &r) {
  const uint64_t n = g->n_rows;
  switch (r.kind) {
    case SSTG_REQ_BFS:
      return n * 4 + n * 24;            // levels + queue, flags, bitmaps
    case SSTG_REQ_PR:
      // ranks, contributions, run ranges; the column-segment layout and its
      // build (4 B per edge + at most one run per edge at 40 B, 1 B per
      // edge of counts) when it applies, else the degree bins
      return n * 32 + (pagerank_segmented_applies(g) ? g->m * 45 : n * 20);
    case SSTG_REQ_CC:
      return n * 8;
    case SSTG_REQ_TC:
      return g->m * 4 + n * 17;         // oriented lists + out-degrees, lists
    default:
      return 0;
  }
}

uint64_t store_bytes(uint32_t n, uint64_t pool_cap) {
  return (uint64_t)n * 16 + pool_cap * 4;
}

void free_set_arrays(StoreSet &t) {
  hipFree(t.off);
  hipFree(t.deg);
  hipFree(t.cap);
  hipFree(t.pool);
}

// Result buffers are recycled by size: hipFree would wait for the whole
// device, including the other slots' queries.
int result_alloc(sstg_server *S, uint64_t bytes, void **out) {
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->spare_results.find(bytes);
    if (it != S->spare_results.end()) {
      *out = it->second;
      S->spare_results.erase(it);
      return SSTG_OK;
    }
  }
  SSTG_HIP(hipMalloc(out, bytes));
  return SSTG_OK;
}

// A finished query drops its read of a store set (caller holds S->mu).  A
// retired set without readers becomes a spare for the next store copy (its
// bytes stay reserved); spares beyond one per slot plus two go to the
// reaper thread, which frees them outside the lock -- hipFree waits for the
// whole device, and the scheduler must not wait with it.
void drop_reader(sstg_server *S, StoreSet *t) {
  if (!t) return;
  t->readers--;
  if (t->owned && t->readers == 0) {
    S->owned_live--;
    S->spare.push_back(*t);
    delete t;
    while (S->spare.size() > S->slots.size() + 2) {
      S->to_free.push_back(S->spare.front());
      S->spare.erase(S->spare.begin());
      S->cv.notify_all();
    }
  }
}

void reaper_main(sstg_server *S) {
  hipSetDevice(S->g->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    S->cv.wait(lk, [&] { return S->stop || !S->to_free.empty(); });
    if (S->to_free.empty()) return;  // stopping with nothing left
    StoreSet t = S->to_free.back();
    S->to_free.pop_back();
    lk.unlock();
    free_set_arrays(t);
    lk.lock();
    S->reserved -= t.bytes;  // released only once actually freed
    S->cv.notify_all();
  }
}

// index of a spare set large enough for the graph's store, or -1
int find_spare(const sstg_server *S) {
  for (size_t i = 0; i < S->spare.size(); i++)
    if (S->spare[i].pool_cap >= S->g->pool_cap) return (int)i;
  return -1;
}

// Gives the graph a private copy of its store so that the queries reading
// the current arrays can keep them (caller holds S->mu; the copy is queued
// on the graph's stream ahead of the update that follows).
int detach_store(sstg_server *S) {
  sstg_graph *g = S->g;
  StoreSet *t = S->live;
  const uint32_t n = g->n_rows;
  StoreSet b;
  const int k = find_spare(S);
  if (k >= 0) {
    b = S->spare[k];
    S->spare.erase(S->spare.begin() + k);
    S->reserved -= b.bytes;  // the spare becomes the graph's store
  } else {
    b.pool_cap = g->pool_cap;
    if (hipMalloc(&b.off, (uint64_t)n * 8) != hipSuccess ||
        hipMalloc(&b.deg, (uint64_t)n * 4) != hipSuccess ||
        hipMalloc(&b.cap, (uint64_t)n * 4) != hipSuccess ||
        hipMalloc(&b.pool, b.pool_cap * 4) != hipSuccess) {
      (void)hipGetLastError();
      free_set_arrays(b);
      return SSTG_E_NOMEM;
    }
  }
  hipStream_t s = g->stream;
  SSTG_HIP(hipMemcpyAsync(b.off, g->off, (uint64_t)n * 8,
                          hipMemcpyDeviceToDevice, s));
  SSTG_HIP(hipMemcpyAsync(b.deg, g->deg, (uint64_t)n * 4,
                          hipMemcpyDeviceToDevice, s));
  SSTG_HIP(hipMemcpyAsync(b.cap, g->cap, (uint64_t)n * 4,
                          hipMemcpyDeviceToDevice, s));
  if (g->pool_tail)
    SSTG_HIP(hipMemcpyAsync(b.pool, g->pool, g->pool_tail * 4,
                            hipMemcpyDeviceToDevice, s));
  // the old arrays now belong to the snapshot its readers hold
  t->owned = true;
  S->owned_live++;
  t->bytes = store_bytes(n, t->pool_cap);
  S->reserved += t->bytes;
  g->off = b.off;
  g->deg = b.deg;
  g->cap = b.cap;
  g->pool = b.pool;
  g->pool_cap = b.pool_cap;
  S->live = nullptr;
  S->detaches++;
  return SSTG_OK;
}

// May the update at the head of the queue start now?  (caller holds S->mu)
bool update_may_start(sstg_server *S) {
  if (!S->live || S->live->readers == 0) return true;
  if (!S->snapshots) return false;
  // at most one live store copy per slot: beyond that the update waits for
  // readers to finish, as without snapshots
  if (S->owned_live >= S->slots.size()) return false;
  const uint64_t need = store_bytes(S->g->n_rows, S->g->pool_cap);
  int k = find_spare(S);
  if (k < 0 && S->reserved + need > S->budget) {
    // spares too small for the grown pool only hold the budget: free them
    // (rare: hipFree waits for the device)
    for (size_t i = 0; i < S->spare.size();) {
      if (S->spare[i].pool_cap < S->g->pool_cap) {
        S->reserved -= S->spare[i].bytes;
        free_set_arrays(S->spare[i]);
        S->spare.erase(S->spare.begin() + i);
      } else {
        i++;
      }
    }
    k = find_spare(S);
  }
  const uint64_t freed = k >= 0 ? S->spare[k].bytes : 0;
  return S->reserved - freed + need <= S->budget;
}

// The set the next admitted or pinned query reads (caller holds S->mu).
StoreSet *live_set(sstg_server *S) {
  if (!S->live) {
    const sstg_graph *g = S->g;
    StoreSet *t = new StoreSet();
    t->off = g->off;
    t->deg = g->deg;
    t->cap = g->cap;
    t->pool = g->pool;
    t->pool_cap = g->pool_cap;
    t->n_rows = g->n_rows;
    t->n_cols = g->n_cols;
    t->m = g->m;
    t->pool_tail = g->pool_tail;
    t->version = g->version;
    t->sym_known = g->sym_version == g->version;
    t->symmetric = g->symmetric;
    S->live = t;
  }
  return S->live;
}

// a slot's view of the store as the set captured it
void view_from_set(sstg_graph *v, const StoreSet *t) {
  v->n_rows = t->n_rows;
  v->n_cols = t->n_cols;
  v->m = t->m;
  v->off = t->off;
  v->deg = t->deg;
  v->cap = t->cap;
  v->pool = t->pool;
  v->pool_cap = t->pool_cap;
  v->pool_tail = t->pool_tail;
  v->version = t->version;  // the view's derived caches follow the version
  v->sym_version = t->sym_known ? t->version : ~0ull;
  v->symmetric = t->symmetric;
}

void refresh_view(sstg_graph *v, const sstg_graph *g) {
  v->n_rows = g->n_rows;
  v->n_cols = g->n_cols;
  v->m = g->m;
  v->off = g->off;
  v->deg = g->deg;
  v->cap = g->cap;
  v->pool = g->pool;
  v->pool_cap = g->pool_cap;
  v->pool_tail = g->pool_tail;
  if (v->version != g->version) {
    v->version = g->version;  // derived caches of the view are recomputed
  }
  v->sym_version = g->sym_version == g->version ? v->version : ~0ull;
  v->symmetric = g->symmetric;
}

int run_query(sstg_server *S, sstg_graph *v, Request *r) {
  const uint64_t n = v->n_rows;
  switch (r->spec.kind) {
    case SSTG_REQ_BFS: {
      uint32_t *lev = nullptr;
      SSTG_TRY(result_alloc(S, n * 4, (void **)&lev));
      r->result = lev;
      r->result_bytes = n * 4;
      if (r->spec.source >= n) return SSTG_E_RANGE;
      return algo_bfs(v, r->spec.source, nullptr, lev);
    }
    case SSTG_REQ_PR: {
      double *rank = nullptr;
      SSTG_TRY(result_alloc(S, n * 8, (void **)&rank));
      r->result = rank;
      r->result_bytes = n * 8;
      return algo_pagerank(v, r->spec.iters, r->spec.damping, rank);
    }
    case SSTG_REQ_CC: {
      uint32_t *lab = nullptr;
      SSTG_TRY(result_alloc(S, n * 4, (void **)&lab));
      r->result = lab;
      r->result_bytes = n * 4;
      return algo_cc(v, lab);
    }
    case SSTG_REQ_TC: {
      SSTG_TRY(algo_tc(v, &r->scalar));
      if (S->maintain) {
        // a full count of a version in the chain anchors it
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->tc_chain.find(v->version);
        if (it != S->tc_chain.end() && !S->tc_anchor) {
          S->tc_anchor = true;
          S->tc_anchor_count = r->scalar;
          S->tc_anchor_delta = it->second;
        }
      }
      return SSTG_OK;
    }
    default:
      return SSTG_E_INVAL;
  }
}

void slot_main(sstg_server *S, Slot *sl) {
  hipSetDevice(S->g->device);
  for (;;) {
    Request *r = nullptr;
    {
      std::unique_lock<std::mutex> lk(sl->mu);
      sl->cv.wait(lk, [&] { return sl->stop || sl->req; });
      if (sl->stop && !sl->req) return;
      r = sl->req;
    }
    int rc = run_query(S, &sl->view, r);
    // drain the slot's stream whatever the query returned: a query that
    // failed part-way may have kernels in flight on the snapshot and on
    // r->result, which must not be recycled before they finish
    {
      hipError_t e = hipStreamSynchronize(sl->view.stream);
      if (e != hipSuccess && rc == SSTG_OK) rc = SSTG_E_DEVICE;
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      r->status = rc;
      r->t_done = now_s();
      r->done = true;
      S->reserved -= r->est_bytes;
      S->running--;
      if (is_heavy(r->spec.kind)) S->heavy_running--;
      drop_reader(S, r->set);
      r->set = nullptr;
      {
        std::lock_guard<std::mutex> lk2(sl->mu);
        sl->req = nullptr;
      }
    }
    S->cv.notify_all();
    S->done_cv.notify_all();
  }
}

Slot *free_slot(sstg_server *S) {
  for (Slot *sl : S->slots) {
    std::lock_guard<std::mutex> lk(sl->mu);
    if (!sl->req) return sl;
  }
  return nullptr;
}

// Starts the update run at pending[at] on the writer thread (caller holds
// S->mu).  Queries ahead of it are admitted or pinned to the current store,
// so the graph can move to a copy of its arrays while they are read.
void start_updates(sstg_server *S, size_t at) {
  std::vector<Request *> run;
  const uint32_t kind = S->pending[at]->spec.kind;
  size_t e = at;
  while (e < S->pending.size() && S->pending[e]->spec.kind == kind &&
         (S->coalesce || run.empty()))
    run.push_back(S->pending[e++]);
  S->pending.erase(S->pending.begin() + at, S->pending.begin() + e);
  const double t0 = now_s();
  for (Request *r : run) r->t_start = t0;
  if (S->live && S->live->readers) {
    int rc = detach_store(S);
    if (rc != SSTG_OK) {
      // no room for a copy: fail the run rather than race the readers
      for (Request *r : run) {
        r->status = rc;
        r->t_done = now_s();
        r->done = true;
      }
      S->done_cv.notify_all();
      return;
    }
  } else if (S->live) {
    delete S->live;  // no readers: the update may move the graph's arrays
    S->live = nullptr;
  }
  uint64_t total = 0;
  for (Request *r : run) total += r->src.size();
  S->write_src.clear();
  S->write_dst.clear();
  S->write_src.reserve(total);
  S->write_dst.reserve(total);
  for (Request *r : run) {
    S->write_src.insert(S->write_src.end(), r->src.begin(), r->src.end());
    S->write_dst.insert(S->write_dst.end(), r->dst.begin(), r->dst.end());
  }
  S->write_run = run;
  S->write_ticket = run.front()->ticket;
  S->writing = true;
  S->cv.notify_all();
}

// Applies an update run to the served graph and carries the cumulative
// triangle delta count0 (every edge with its reverse before the run): the
// pairs the run changes are found first; a removal's broken triangles are
// counted on the graph before it, an insertion's closed ones on the graph
// after it.  *carried is false when the delta could not be carried (then the
// run was applied all the same, or not at all if an id was out of range).
int update_carrying_tc(sstg_graph *g, const std::vector<uint32_t> &src,
                       const std::vector<uint32_t> &dst, bool ins,
                       int64_t count0, int64_t *count1, bool *carried) {
  *carried = false;
  const uint64_t n = src.size();
  if (n >= (1ull << 31))
    return graph_update(g, src.data(), dst.data(), n, ins, ins, false);
  uint32_t *ds = nullptr, *dd = nullptr;
  SSTG_TRY(g->ws.get(kWsInSrc, n + 1, &ds));
  SSTG_TRY(g->ws.get(kWsInDst, n + 1, &dd));
  SSTG_HIP(hipMemcpyAsync(ds, src.data(), n * 4, hipMemcpyHostToDevice,
                          g->stream));
  SSTG_HIP(hipMemcpyAsync(dd, dst.data(), n * 4, hipMemcpyHostToDevice,
                          g->stream));
  const uint64_t *keys = nullptr;
  uint64_t count = 0;
  int rc = graph_effective_keys(g, ds, dd, n, ins, ins, &keys, &count);
  if (rc != SSTG_OK)  // e.g. an id out of range: the caller sorts it out
    return graph_update(g, src.data(), dst.data(), n, ins, ins, false);
  uint64_t delta = 0;
  if (!ins) SSTG_TRY(tc_delta(g, keys, count, &delta));
  SSTG_TRY(graph_update(g, ds, dd, n, ins, ins, true));
  if (ins) SSTG_TRY(tc_delta(g, keys, count, &delta));
  *count1 = ins ? count0 + (int64_t)delta : count0 - (int64_t)delta;
  *carried = true;
  return SSTG_OK;
}

void writer_main(sstg_server *S) {
  hipSetDevice(S->g->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    S->cv.wait(lk, [&] { return S->stop || !S->write_run.empty(); });
    if (S->write_run.empty()) return;  // stopping
    std::vector<Request *> run = S->write_run;
    const uint32_t kind = run.front()->spec.kind;
    const bool ins = kind == SSTG_REQ_INSERT;
    sstg_graph *g = S->g;
    auto link = S->tc_chain.find(g->version);
    const bool carry = S->maintain && link != S->tc_chain.end() &&
                       g->sym_version == g->version && g->symmetric;
    const int64_t delta0 = carry ? link->second : 0;
    lk.unlock();
    int rc = SSTG_OK;
    bool carried = false;
    int64_t delta1 = 0;
    if (carry) {
      rc = update_carrying_tc(g, S->write_src, S->write_dst, ins, delta0,
                              &delta1, &carried);
    } else {
      rc = graph_update(g, S->write_src.data(), S->write_dst.data(),
                        S->write_src.size(), ins, ins, false);
    }
    std::vector<int> rcs(run.size(), rc);
    if (rc == SSTG_E_RANGE && run.size() > 1) {
      // one request holds a bad id: apply them one by one so only it fails
      for (size_t i = 0; i < run.size(); i++)
        rcs[i] = graph_update(S->g, run[i]->src.data(), run[i]->dst.data(),
                              run[i]->src.size(), ins, ins, false);
    }
    lk.lock();
    if (S->maintain) {
      const bool sym = g->sym_version == g->version && g->symmetric;
      if (!(carried && rc == SSTG_OK && sym)) {
        // the chain breaks; a symmetric graph starts a new one, whose
        // versions a later full count anchors again
        S->tc_chain.clear();
        S->tc_anchor = false;
        if (sym) S->tc_chain[g->version] = 0;
      } else {
        S->tc_chain[g->version] = delta1;
        while (S->tc_chain.size() > 4096) S->tc_chain.erase(S->tc_chain.begin());
      }
    }
    const double t1 = now_s();
    for (size_t i = 0; i < run.size(); i++) {
      Request *r = run[i];
      r->status = rcs[i];
      r->t_done = t1;
      r->done = true;
      r->coalesced = run.size();
      r->src.clear();
      r->src.shrink_to_fit();
      r->dst.clear();
      r->dst.shrink_to_fit();
    }
    S->write_run.clear();
    S->writing = false;
    S->device_batches++;
    S->cv.notify_all();
    S->done_cv.notify_all();
  }
}

bool admissible(sstg_server *S, const Request *q) {
  if (S->reserved + q->est_bytes > S->budget) return false;
  if (S->one_heavy && is_heavy(q->spec.kind) && S->heavy_running) return false;
  return free_slot(S) != nullptr;
}

void admit(sstg_server *S, Request *q) {
  Slot *sl = free_slot(S);
  S->reserved += q->est_bytes;
  S->running++;
  if (is_heavy(q->spec.kind)) S->heavy_running++;
  q->t_start = now_s();
  if (!q->set) {  // not pinned: reads the current store
    q->set = live_set(S);
    q->set->readers++;
  }
  view_from_set(&sl->view, q->set);
  {
    std::lock_guard<std::mutex> lk2(sl->mu);
    sl->req = q;
  }
  sl->cv.notify_one();
}

// One pass over the queue in arrival order (caller holds S->mu): admits what
// it can, pins what must wait, and starts the first update run once no query
// ahead of it is left unpinned.  Returns whether anything changed.
bool sched_step(sstg_server *S) {
  bool update_waiting = false;
  for (Request *r : S->pending)
    if (is_update(r->spec.kind)) {
      update_waiting = true;
      break;
    }
  bool changed = false;
  bool unpinned_ahead = false;
  for (size_t i = 0; i < S->pending.size(); i++) {
    Request *r = S->pending[i];
    // while an update run is applied, only the queries ahead of it (all
    // pinned to their store) may start
    if (S->writing && r->ticket > S->write_ticket) return changed;
    if (is_update(r->spec.kind)) {
      if (!S->writing && !unpinned_ahead && update_may_start(S)) {
        start_updates(S, i);
        return true;
      }
      return changed;  // later queries must see this update
    }
    std::map<uint64_t, int64_t>::const_iterator link;
    if (r->spec.kind == SSTG_REQ_TC && S->tc_anchor &&
        (r->set || !S->writing) &&
        (link = S->tc_chain.find(r->set ? r->set->version : S->g->version)) !=
            S->tc_chain.end()) {
      // the count of the version this query reads, from the chain
      S->pending.erase(S->pending.begin() + i);
      i--;
      r->t_start = r->t_done = now_s();
      r->scalar = (uint64_t)((int64_t)S->tc_anchor_count + link->second -
                             S->tc_anchor_delta);
      r->status = SSTG_OK;
      r->done = true;
      if (r->set) {
        drop_reader(S, r->set);
        r->set = nullptr;
      }
      S->tc_from_count++;
      S->done_cv.notify_all();
      changed = true;
      continue;
    }
    if (admissible(S, r)) {
      S->pending.erase(S->pending.begin() + i);
      admit(S, r);
      i--;
      changed = true;
      continue;
    }
    if (S->snapshots && update_waiting && !r->set && !S->writing) {
      r->set = live_set(S);  // pinned: the update behind it need not wait
      r->set->readers++;
      changed = true;
    }
    if (!r->set) unpinned_ahead = true;
  }
  return changed;
}

void sched_main(sstg_server *S) {
  hipSetDevice(S->g->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    if (S->stop) return;
    if (!sched_step(S)) S->cv.wait(lk);
  }
}

int view_init(sstg_graph *v, const sstg_graph *g) {
  v->device = g->device;
  SSTG_HIP(hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking));
  v->ws.stream = v->stream;
  SSTG_HIP(hipMalloc(&v->dev_ctr, 16 * sizeof(uint64_t)));
  SSTG_HIP(hipHostMalloc(&v->host_rb, 64 * sizeof(uint64_t),
                         hipHostMallocDefault));
  v->version = ~0ull;
  v->no_capture = true;
  refresh_view(v, g);
  return SSTG_OK;
}

void view_free(sstg_graph *v) {
  if (v->stream) hipStreamSynchronize(v->stream);
  pagerank_release(v);
  v->ws.release_all();
  hipFree(v->dev_ctr);
  hipFree(v->deg_prefix);
  if (v->host_rb) hipHostFree(v->host_rb);
  if (v->stream) hipStreamDestroy(v->stream);
  // the store arrays belong to the served graph
  v->off = nullptr;
  v->deg = v->cap = v->pool = nullptr;
}

}  // namespace
}  // namespace sstg

using namespace sstg;

extern "C" {

int sstg_server_create(sstg_graph *g, uint32_t slots, uint64_t hbm_budget,
                       uint32_t flags, sstg_server **out) {
  if (!g || !out || slots == 0 || slots > 64) return SSTG_E_INVAL;
  if (flags & ~(SSTG_SERVE_COALESCE | SSTG_SERVE_SNAPSHOTS |
                SSTG_SERVE_ONE_HEAVY | SSTG_SERVE_MAINTAIN))
    return SSTG_E_INVAL;
  *out = nullptr;
  // store copies and update batches carry no weight table
  if (g->wt_keys) return SSTG_E_UNSUPPORTED;
  hipSetDevice(g->device);
  bool sym = false;
  SSTG_TRY(graph_is_symmetric(g, &sym));  // known once, then tracked
  sstg_server *S = new (std::nothrow) sstg_server();
  if (!S) return SSTG_E_NOMEM;
  S->g = g;
  S->coalesce = (flags & SSTG_SERVE_COALESCE) ? 1 : 0;
  S->snapshots = (flags & SSTG_SERVE_SNAPSHOTS) ? 1 : 0;
  S->one_heavy = (flags & SSTG_SERVE_ONE_HEAVY) ? 1 : 0;
  S->maintain = (flags & SSTG_SERVE_MAINTAIN) ? 1 : 0;
  if (hipStreamCreateWithFlags(&S->fetch_stream, hipStreamNonBlocking) !=
      hipSuccess) {
    delete S;
    return SSTG_E_DEVICE;
  }
  if (S->maintain && sym) S->tc_chain[g->version] = 0;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    hipStreamDestroy(S->fetch_stream);
    delete S;
    return SSTG_E_DEVICE;
  }
  // keep 1/8 of the free memory for update scratch and pool growth
  uint64_t avail = free_b - free_b / 8;
  S->budget = hbm_budget ? (hbm_budget < avail ? hbm_budget : avail) : avail;
  for (uint32_t i = 0; i < slots; i++) {
    Slot *sl = new Slot();
    int rc = view_init(&sl->view, g);
    if (rc) {
      view_free(&sl->view);
      delete sl;
      sstg_server_destroy(S);
      return rc;
    }
    S->slots.push_back(sl);
  }
  for (Slot *sl : S->slots) sl->th = std::thread(slot_main, S, sl);
  S->sched = std::thread(sched_main, S);
  S->reaper = std::thread(reaper_main, S);
  S->writer = std::thread(writer_main, S);
  *out = S;
  return SSTG_OK;
}

int sstg_server_submit(sstg_server *S, const sstg_request *req,
                       uint64_t *ticket) {
  if (!S || !req || !ticket) return SSTG_E_INVAL;
  if (req->kind < SSTG_REQ_INSERT || req->kind > SSTG_REQ_CC)
    return SSTG_E_INVAL;
  Request *r = new (std::nothrow) Request();
  if (!r) return SSTG_E_NOMEM;
  r->spec = *req;
  r->t_arrive = now_s();
  if (is_update(req->kind)) {
    if (req->n && (!req->src || !req->dst)) {
      delete r;
      return SSTG_E_INVAL;
    }
    r->src.assign(req->src, req->src + req->n);
    r->dst.assign(req->dst, req->dst + req->n);
  } else {
    r->est_bytes = estimate(S->g, *req);
  }
  std::lock_guard<std::mutex> lk(S->mu);
  if (r->est_bytes > S->budget) {
    delete r;
    return SSTG_E_NOMEM;  // could never be admitted
  }
  r->ticket = S->next_ticket++;
  S->all[r->ticket] = r;
  S->pending.push_back(r);
  *ticket = r->ticket;
  S->cv.notify_all();
  return SSTG_OK;
}

int sstg_server_wait(sstg_server *S, uint64_t ticket, sstg_request_info *info) {
  if (!S) return SSTG_E_INVAL;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->all.find(ticket);
  if (it == S->all.end()) return SSTG_E_INVAL;
  Request *r = it->second;
  S->done_cv.wait(lk, [&] { return r->done; });
  if (info) {
    info->status = r->status;
    info->t_arrive = r->t_arrive;
    info->t_start = r->t_start;
    info->t_done = r->t_done;
    info->scalar = r->scalar;
    info->coalesced = r->coalesced;
  }
  return SSTG_OK;
}

int sstg_server_fetch(sstg_server *S, uint64_t ticket, void *host_out,
                      uint64_t bytes) {
  if (!S) return SSTG_E_INVAL;
  Request *r = nullptr;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return SSTG_E_INVAL;
    r = it->second;
    S->done_cv.wait(lk, [&] { return r->done; });
  }
  if (!r->result || !host_out) return SSTG_E_INVAL;
  if (bytes < r->result_bytes) return SSTG_E_INVAL;
  hipSetDevice(S->g->device);
  // on the server's own non-blocking stream: a synchronous legacy-stream
  // copy would wait on (or be refused beside) the slots' streams
  std::lock_guard<std::mutex> lk(S->fetch_mu);
  SSTG_HIP(hipMemcpyAsync(host_out, r->result, r->result_bytes,
                          hipMemcpyDeviceToHost, S->fetch_stream));
  SSTG_HIP(hipStreamSynchronize(S->fetch_stream));
  return SSTG_OK;
}

int sstg_server_release(sstg_server *S, uint64_t ticket) {
  if (!S) return SSTG_E_INVAL;
  Request *r = nullptr;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return SSTG_E_INVAL;
    r = it->second;
    S->done_cv.wait(lk, [&] { return r->done; });
    S->all.erase(it);
  }
  if (r->result) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->spare_results.emplace(r->result_bytes, r->result);
  }
  delete r;
  return SSTG_OK;
}

int sstg_server_stats(sstg_server *S, uint64_t *device_batches,
                      uint64_t *budget, uint64_t *snapshots) {
  if (!S) return SSTG_E_INVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  if (device_batches) *device_batches = S->device_batches;
  if (budget) *budget = S->budget;
  if (snapshots) *snapshots = S->detaches;
  return SSTG_OK;
}

int sstg_server_destroy(sstg_server *S) {
  if (!S) return SSTG_E_INVAL;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    // drain: wait for every admitted or pending request
    S->done_cv.wait(lk, [&] {
      for (auto &kv : S->all)
        if (!kv.second->done) return false;
      return true;
    });
    S->stop = true;
  }
  S->cv.notify_all();
  if (S->sched.joinable()) S->sched.join();
  if (S->reaper.joinable()) S->reaper.join();
  if (S->writer.joinable()) S->writer.join();
  for (Slot *sl : S->slots) {
    {
      std::lock_guard<std::mutex> lk(sl->mu);
      sl->stop = true;
    }
    sl->cv.notify_one();
    if (sl->th.joinable()) sl->th.join();
    view_free(&sl->view);
    delete sl;
  }
  hipSetDevice(S->g->device);
  for (auto &kv : S->all) {
    if (kv.second->result) hipFree(kv.second->result);
    delete kv.second;
  }
  for (auto &kv : S->spare_results) hipFree(kv.second);
  for (StoreSet &t : S->spare) free_set_arrays(t);
  for (StoreSet &t : S->to_free) free_set_arrays(t);
  if (S->fetch_stream) hipStreamDestroy(S->fetch_stream);
  delete S->live;  // after the drain no query reads it; never owned
  delete S;
  return SSTG_OK;
}

}  // extern "C"
