// Many independent fills in one launch per kernel class and stage position,
// and the same captured once in a hipGraph.
//
// Every plan brings its tiles with its own job (working array, factor table,
// strides); position k runs the k-th stage group of every plan -- the plans
// share nothing, so any alignment of their stages is a valid order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "api_common.hpp"
#include "kernels.hpp"
#include "plan.hpp"

using namespace tah;

namespace tah {

namespace {
// Live plans and the serial of each one's current state. A serial is never
// reused, so a plan freed and another created at the same address cannot be
// taken for the old one.
std::mutex registry_mutex;
std::unordered_map<const tah_plan*, uint64_t> live_plans;
uint64_t next_serial = 1;
}  // namespace

void plan_register(tah_plan* p) {
  std::lock_guard<std::mutex> lock(registry_mutex);
  p->serial = next_serial++;
  live_plans[p] = p->serial;
}

void plan_unregister(tah_plan* p) {
  std::lock_guard<std::mutex> lock(registry_mutex);
  live_plans.erase(p);
}

bool plan_serial_live(const tah_plan* p, uint64_t serial) {
  std::lock_guard<std::mutex> lock(registry_mutex);
  auto it = live_plans.find(p);
  return it != live_plans.end() && it->second == serial;
}

struct Batch {
  std::vector<tah_plan*> plans;
  std::vector<uint64_t> serials;  // each plan's state when the batch was built
  std::vector<double*> works;
  std::vector<const double*> factors;
  std::vector<Job> jobs;
  Job* djobs = nullptr;
  Tile* dtiles = nullptr;
  LdsTile* dltiles = nullptr;
  std::vector<std::vector<Launch>> launches;  // per stage position
  std::vector<hipEvent_t> events;
  int device = 0;
  bool fork = false;  // forked when any of its plans is
  // The work queue (every plan's option `queue` set): per stage position
  // the count of its queued tiles (dqtiles, Tile.pad the kernel code), nine
  // share bounds (dshares) and eight counters (dcounters); the classes the
  // queue kernel does not take stay in `launches`.
  bool queue = false;
  uint32_t queue_grid = 0;
  std::vector<int64_t> qcount;
  Tile* dqtiles = nullptr;
  uint32_t* dshares = nullptr;
  uint32_t* dcounters = nullptr;
  ~Batch() {
    if (djobs) (void)hipFree(djobs);
    if (dtiles) (void)hipFree(dtiles);
    if (dltiles) (void)hipFree(dltiles);
    if (dqtiles) (void)hipFree(dqtiles);
    if (dshares) (void)hipFree(dshares);
    if (dcounters) (void)hipFree(dcounters);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
  }
};

namespace {

std::unique_ptr<Batch> make_batch(tah_ctx* ctx, int32_t n, tah_plan* const* plans, double* const* works,
                                  const double* const* factors) {
  if (!ctx || n < 1 || !plans || !works) fail("null argument");
  auto b = std::make_unique<Batch>();
  b->device = ctx->device;
  size_t npos = 0;
  for (int32_t i = 0; i < n; ++i) {
    tah_plan* p = plans[i];
    if (!p || !works[i]) fail("null plan or working array in the batch");
    if (p->regrid) fail("plan " + std::to_string(i) + " is a regrid transfer's, not a fill's");
    if (p->ctx->device != ctx->device) fail("every plan of a batch must be on the context's device");
    const double* f = factors ? factors[i] : nullptr;
    if (p->needs_factors && !f)
      fail("plan " + std::to_string(i) + " mirrors ghosts across a reflecting face, but no factor table was given");
    b->plans.push_back(p);
    b->serials.push_back(p->serial);
    b->fork = b->fork || p->fork != 0;
    b->queue = i == 0 ? p->queue != 0 : b->queue && p->queue != 0;
    b->works.push_back(works[i]);
    b->factors.push_back(f);
    Job j = plan_job(p, works[i], f);
    b->jobs.push_back(j);
    npos = std::max(npos, p->launches.size());
  }
  for (size_t i = 0; i < b->works.size(); ++i)
    for (size_t k = i + 1; k < b->works.size(); ++k)
      if (b->works[i] == b->works[k]) fail("two fills of one batch share a working array");
  std::vector<Tile> tiles;
  std::vector<LdsTile> ltiles;
  std::vector<Tile> qtiles;
  std::vector<uint32_t> shares;
  // the block a tile works on: its target block (its source block for
  // restriction, whose tiles are ordered by source block)
  auto block_of = [&](const Tile& t, int kind) {
    const tah_plan* p = b->plans[size_t(t.job)];
    const bool unit = kind >= kQueueMerged;
    const PlanGroup& g = p->groups[size_t(unit ? p->wclasses[size_t(t.item)].group : t.item)];
    return (g.kind == 2 ? g.sblocks : g.tblocks)[size_t(t.transfer)];
  };
  for (size_t pos = 0; pos < npos; ++pos) {
    std::map<std::pair<int, int32_t>, std::vector<Tile>> bycls;  // by (class, workgroup size)
    std::map<std::pair<int, int32_t>, std::vector<LdsTile>> ldscls;  // LDS-staged, by (class, workgroup size)
    std::map<std::pair<int, int32_t>, int32_t> ldsbytes;
    for (int32_t i = 0; i < n; ++i) {
      tah_plan* p = b->plans[i];
      if (pos >= p->launches.size()) continue;
      for (const Launch& L : p->launches[pos]) {
        if (L.lds > 0) {
          // the chunk index stays the plan's: its job points at its chunks
          for (int64_t t = L.first; t < L.first + L.count; ++t) {
            LdsTile tl = p->lds_tiles[size_t(t)];
            tl.job = i;
            ldscls[{L.cls, L.threads}].push_back(tl);
          }
          ldsbytes[{L.cls, L.threads}] = std::max(ldsbytes[{L.cls, L.threads}], L.lds);
          continue;
        }
        for (int64_t t = L.first; t < L.first + L.count; ++t) {
          Tile tl = p->tiles[t];
          tl.job = i;
          bycls[{L.cls, L.threads}].push_back(tl);
        }
      }
    }
    if (b->queue) {
      // the classes the queue kernel takes, in one list ordered by (plan,
      // block) so that every XCD's share holds a mix of them, each tile
      // where the tiles that read or write its blocks are
      std::vector<std::pair<uint64_t, Tile>> q;
      for (auto it = bycls.begin(); it != bycls.end();) {
        const int kind = queue_kind(it->first.first);
        if (kind < 0) {
          ++it;
          continue;
        }
        for (Tile t : it->second) {
          t.pad = kind;
          q.push_back({(uint64_t(uint32_t(t.job)) << 32) | uint32_t(block_of(t, kind)), t});
        }
        it = bycls.erase(it);
      }
      std::stable_sort(q.begin(), q.end(), [](const std::pair<uint64_t, Tile>& x, const std::pair<uint64_t, Tile>& y) {
        return x.first < y.first;
      });
      const uint32_t first = uint32_t(qtiles.size());
      for (const auto& e : q) qtiles.push_back(e.second);
      for (uint32_t x = 0; x <= 8; ++x) shares.push_back(first + uint32_t(uint64_t(q.size()) * x / 8));
      b->qcount.push_back(int64_t(q.size()));
      if (qtiles.size() > size_t(UINT32_MAX / 2)) fail("too many tiles for the batch's work queue", kUnsupported);
    }
    std::vector<Launch> ls;
    for (auto& kv : bycls) {
      ls.push_back(Launch{kv.first.first, int64_t(tiles.size()), int64_t(kv.second.size()), 0, 0, 0, kv.first.second});
      tiles.insert(tiles.end(), kv.second.begin(), kv.second.end());
    }
    for (auto& kv : ldscls) {
      ls.push_back(Launch{kv.first.first, int64_t(ltiles.size()), int64_t(kv.second.size()), 0, 0, ldsbytes[kv.first],
                          kv.first.second});
      ltiles.insert(ltiles.end(), kv.second.begin(), kv.second.end());
    }
    // The largest launch first, copies or not: on the 64-forest batch the
    // forked graph fill took 20.30-20.40 ms so, against 20.71-20.87 ms with
    // interpolations before copies (the order of a single plan's launches).
    std::stable_sort(ls.begin(), ls.end(), [](const Launch& x, const Launch& y) { return x.count > y.count; });
    b->launches.push_back(std::move(ls));
  }
  HIPCHECK(hipSetDevice(ctx->device));
  HIPCHECK(hipMalloc(&b->djobs, sizeof(Job) * b->jobs.size()));
  HIPCHECK(hipMemcpy(b->djobs, b->jobs.data(), sizeof(Job) * b->jobs.size(), hipMemcpyHostToDevice));
  if (!tiles.empty()) {
    HIPCHECK(hipMalloc(&b->dtiles, sizeof(Tile) * tiles.size()));
    HIPCHECK(hipMemcpy(b->dtiles, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice));
  }
  if (!ltiles.empty()) {
    HIPCHECK(hipMalloc(&b->dltiles, sizeof(LdsTile) * ltiles.size()));
    HIPCHECK(hipMemcpy(b->dltiles, ltiles.data(), sizeof(LdsTile) * ltiles.size(), hipMemcpyHostToDevice));
  }
  if (b->queue) {
    // a persistent grid: as many workgroups as the CUs hold at the queue
    // kernel's occupancy, a multiple of 8 (one share per XCD)
    const int32_t wgs = ctx->prop.multiProcessorCount * (queue_waves_per_cu() / (kThreads / 64));
    b->queue_grid = uint32_t(std::max(8, wgs / 8 * 8));
    if (!qtiles.empty()) {
      HIPCHECK(hipMalloc(&b->dqtiles, sizeof(Tile) * qtiles.size()));
      HIPCHECK(hipMemcpy(b->dqtiles, qtiles.data(), sizeof(Tile) * qtiles.size(), hipMemcpyHostToDevice));
    }
    HIPCHECK(hipMalloc(&b->dshares, sizeof(uint32_t) * std::max<size_t>(1, shares.size())));
    if (!shares.empty())
      HIPCHECK(hipMemcpy(b->dshares, shares.data(), sizeof(uint32_t) * shares.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMalloc(&b->dcounters, sizeof(uint32_t) * 8 * std::max<size_t>(1, npos)));
  }
  for (int i = 0; i < 8; ++i) {
    hipEvent_t e;
    HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->events.push_back(e);
  }
  return b;
}

void run_batch(tah_ctx* ctx, Batch& b, hipStream_t s) {
  if (b.queue) launch_queue_reset(s, b.dcounters, uint32_t(8 * b.launches.size()));
  for (size_t pos = 0; pos < b.launches.size(); ++pos) {
    if (b.queue && b.qcount[pos] > 0)
      launch_queue(b.queue_grid, s, b.djobs, b.dqtiles, b.dshares + 9 * pos, b.dcounters + 8 * pos);
    run_launches(ctx, b.launches[pos], s, b.djobs, b.dtiles, b.dltiles, b.events.data(), b.fork);
  }
}

// Every plan of the batch alive and in the state the batch was built from.
bool batch_current(const Batch& b) {
  for (size_t i = 0; i < b.plans.size(); ++i)
    if (!plan_serial_live(b.plans[i], b.serials[i])) return false;
  return true;
}

bool same_batch(const Batch& b, int32_t n, tah_plan* const* plans, double* const* works,
                const double* const* factors) {
  if (int32_t(b.plans.size()) != n) return false;
  for (int32_t i = 0; i < n; ++i)
    if (b.plans[i] != plans[i] || b.works[i] != works[i] || b.factors[i] != (factors ? factors[i] : nullptr))
      return false;
  return batch_current(b);
}

}  // namespace
}  // namespace tah

struct tah_graph {
  tah_ctx* ctx = nullptr;
  std::unique_ptr<Batch> batch;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~tah_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace {
// The last batch of each context, reused while the same fills (the same
// plans in the same states, arrays and factor tables) are asked for.
std::mutex batch_mutex;
std::map<const tah_ctx*, std::unique_ptr<Batch>> last_batch;
}  // namespace

namespace tah {

void batch_forget_plan(const tah_plan* p) {
  std::lock_guard<std::mutex> lock(batch_mutex);
  for (auto it = last_batch.begin(); it != last_batch.end();) {
    const auto& ps = it->second ? it->second->plans : std::vector<tah_plan*>{};
    if (std::find(ps.begin(), ps.end(), p) != ps.end()) {
      (void)hipDeviceSynchronize();  // its launches may still read the batch's tables
      it = last_batch.erase(it);
    } else {
      ++it;
    }
  }
}

void batch_forget_ctx(const tah_ctx* c) {
  std::lock_guard<std::mutex> lock(batch_mutex);
  auto it = last_batch.find(c);
  if (it == last_batch.end()) return;
  (void)hipDeviceSynchronize();
  last_batch.erase(it);
}

}  // namespace tah

extern "C" {

tah_status tah_fill_batched(tah_ctx* ctx, int32_t n, tah_plan* const* plans, double* const* works,
                            const double* const* factors, void* stream) {
  return guard([&] {
    if (!ctx) fail("null context");
    HIPCHECK(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lock(batch_mutex);
    auto& slot = last_batch[ctx];
    if (!slot || !same_batch(*slot, n, plans, works, factors)) {
      HIPCHECK(hipDeviceSynchronize());  // the previous batch's tables may still be in use
      slot = make_batch(ctx, n, plans, works, factors);
    }
    run_batch(ctx, *slot, reinterpret_cast<hipStream_t>(stream));
  });
}

tah_status tah_graph_create(tah_ctx* ctx, int32_t n, tah_plan* const* plans, double* const* works,
                            const double* const* factors, tah_graph** out) {
  return guard([&] {
    if (!out) fail("null argument");
    auto g = std::make_unique<tah_graph>();
    g->ctx = ctx;
    g->batch = make_batch(ctx, n, plans, works, factors);
    ctx_streams(ctx);  // created outside the capture
    hipStream_t cs;
    HIPCHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      (void)hipStreamDestroy(cs);
      HIPCHECK(e);
    }
    try {
      run_batch(ctx, *g->batch, cs);
    } catch (...) {
      hipGraph_t dropped = nullptr;
      (void)hipStreamEndCapture(cs, &dropped);
      if (dropped) (void)hipGraphDestroy(dropped);
      (void)hipStreamDestroy(cs);
      throw;
    }
    e = hipStreamEndCapture(cs, &g->graph);
    (void)hipStreamDestroy(cs);
    HIPCHECK(e);
    HIPCHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g.release();
  });
}

tah_status tah_graph_launch(tah_graph* g, void* stream) {
  return guard([&] {
    if (!g) fail("null graph");
    if (!batch_current(*g->batch))
      fail("a plan of this graph was destroyed, or its options changed, after the graph was created; "
           "create the graph again");
    HIPCHECK(hipSetDevice(g->ctx->device));
    HIPCHECK(hipGraphLaunch(g->exec, reinterpret_cast<hipStream_t>(stream)));
  });
}

tah_status tah_graph_destroy(tah_graph* g) {
  delete g;
  return TAH_OK;
}

}  // extern "C"
