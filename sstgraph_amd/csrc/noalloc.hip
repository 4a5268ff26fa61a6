// Allocation-free batch insert: sstg_insert_batch with every scratch buffer
// carved from one fixed device array instead of hipMalloc'd workspace.
//
// The array below is the scratch.  Its name is a host-side stand-in in host
// code, not a device address: the address of each GPU's copy comes from
// hipGetSymbolAddress, taken once per device with that device current, and
// only that address is handed to kernels (through the graph's workspace in
// arena mode).  The update's counters and read-back block are the graph's
// own, allocated when the graph was created.
#include <mutex>

#include "graph.hpp"

constexpr size_t kScratchBytes = 64ull << 20;

// Bound on the scratch one insert of n pairs needs outside its long-row
// staging: two 8-byte key arrays, flags and two positions per pair (25 B),
// and per touched row (at most n of them) the segment, class and tile
// arrays (about 110 B), plus the scans' partials.
static_assert(SSTG_NOALLOC_MAX_PAIRS * (25 + 110) + (4ull << 20) <=
                  kScratchBytes,
              "scratch too small for SSTG_NOALLOC_MAX_PAIRS");

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch (the current device must be `dev`).
int scratch_for(int dev, void **out) {
  if (dev < 0 || dev >= kMaxDevices) return SSTG_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    SSTG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SSTG_OK;
}

}  // namespace

extern "C" int sstg_insert_batch_noalloc(sstg_graph *g, const uint32_t *src,
                                         const uint32_t *dst, uint64_t n,
                                         uint32_t flags) {
  if (!g || (n && (!src || !dst))) return SSTG_E_INVAL;
  if (!(flags & SSTG_DEVICE_PTRS)) return SSTG_E_INVAL;
  if (n > SSTG_NOALLOC_MAX_PAIRS) return SSTG_E_TOO_LARGE;
  if (hipSetDevice(g->device) != hipSuccess) return SSTG_E_DEVICE;
  void *base = nullptr;
  SSTG_TRY(scratch_for(g->device, &base));
  g->ws.begin_arena(base, kScratchBytes);
  g->noalloc = true;
  int rc = sstg::graph_update(g, src, dst, n, true,
                              !(flags & SSTG_KEEP_SELF_LOOPS), true);
  g->noalloc = false;
  g->ws.end_arena();
  // an update that stopped early may still have kernels reading the scratch
  if (hipStreamSynchronize(g->stream) != hipSuccess && rc == SSTG_OK)
    rc = SSTG_E_DEVICE;
  if (rc != SSTG_OK) {
    // a refused batch (pool tail or scratch spent) left the graph unchanged
    g->last = sstg_update_stats{};
    g->last.batch = n;
    g->last.rebuilds = g->rebuilds;
  }
  return rc;
}
