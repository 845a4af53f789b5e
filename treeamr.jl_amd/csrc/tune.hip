// Plan options and the autotuner.
//
// A plan's launches have six knobs: stage groups forked over the
// context's streams or run in order on one; the unit classes of one
// interpolation order in one launch or a launch each; phase-1 transfers
// deferred into the last stage group or not; the copy kernel's tile (4, 8 or
// 16 values per thread); the copies from one workgroup per tile or from a
// bounded grid that leaves wave slots to the interpolations beside them;
// merged interpolations over variable pairs or single variables. tah_tune times every
// combination on the caller's working array and keeps the fastest, cached
// in a text file under a signature of the mesh, the layout and the device,
// so that a second run over the same mesh skips the timing. The reference
// has no autotuner (its launches are fixed, src/ghosts.jl:404-521).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "api_common.hpp"
#include "plan.hpp"

using namespace tah;

namespace {

// FNV-1a over what the launches depend on.
struct Hash {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
      h ^= c[i];
      h *= 1099511628211ull;
    }
  }
  template <class T>
  void add(const T& x) {
    add(&x, sizeof(T));
  }
};

std::string signature(const tah_plan* p) {
  Hash H;
  H.add(p->layout.stored);
  H.add(p->layout.nvars);
  H.add(p->layout.nblocks);
  H.add(p->layout.stride);
  H.add(p->nstages);
  for (int32_t t : p->tags) H.add(t);
  for (const auto& sg : p->stagegroups) H.add(sg);
  for (const PlanGroup& g : p->groups) {
    H.add(g.kind);
    H.add(g.stage);
    H.add(g.factorcol);
    H.add(g.n);
    H.add(g.order);
    H.add(g.tfirst);
    if (!g.tblocks.empty()) H.add(g.tblocks.data(), 4 * g.tblocks.size());
    if (!g.sblocks.empty()) H.add(g.sblocks.data(), 4 * g.sblocks.size());
    for (int d = 0; d < 3; ++d)
      if (!g.srcstart[d].empty()) H.add(g.srcstart[d].data(), 4 * g.srcstart[d].size());
  }
  char buf[64];
  std::snprintf(buf, sizeof buf, "%016llx", static_cast<unsigned long long>(H.h));
  return std::string(p->ctx->prop.gcnArchName) + "-" + std::to_string(p->ctx->prop.multiProcessorCount) +
         "-" + buf;
}

void set_options(tah_plan* p, const tah_plan_options& o) {
  if (o.fork < 0 || o.fork > 1 || o.merge < 0 || o.merge > 1 || o.defer < 0 || o.defer > 1)
    fail("plan options fork, merge and defer are 0 or 1");
  if (o.copy_per_thread != 4 && o.copy_per_thread != 8 && o.copy_per_thread != 16)
    fail("copy_per_thread is 4, 8 or 16");
  if (o.fuse != 0)
    fail("the fused copy-and-interpolation launch is switched off in this build (fuse must be 0)",
         kUnsupported);
  const int32_t wgs = o.copy_wgs;
  if (wgs < 0 || wgs % 8 != 0) fail("copy_wgs is 0 or a positive multiple of 8");
  if (o.interp_pairs < 0 || o.interp_pairs > 1) fail("plan option interp_pairs is 0 or 1");
  if (o.lds < 0 || o.lds > 3) fail("plan option lds is 0 to 3 (bit 0 restriction, bit 1 prolongation)");
  if (o.queue < 0 || o.queue > 1) fail("plan option queue is 0 or 1");
  if (o.reserved_[0] || o.reserved_[1] || o.reserved_[2]) fail("tah_plan_options.reserved_ must be zero");
  const bool rebuild = o.merge != p->merge || o.defer != p->defer || o.copy_per_thread != p->copy_per_thread ||
                       wgs != p->copy_wgs || o.interp_pairs != p->interp_pairs || o.lds != p->lds;
  p->fork = o.fork;
  p->merge = o.merge;
  p->defer = o.defer;
  p->copy_per_thread = o.copy_per_thread;
  p->copy_wgs = wgs;
  p->interp_pairs = o.interp_pairs;
  p->lds = o.lds;
  if (o.queue != p->queue) {
    // batches built from this plan are taken again with the new mode
    p->queue = o.queue;
    plan_register(p);
  }
  if (rebuild) build_launches(p);
}

bool lookup(const std::string& path, const std::string& key, tah_plan_options& o) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  bool found = false;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string tag, k;
    int f, m, d, c, w, ip, l;
    if (!(ss >> tag >> k >> f >> m >> d >> c >> w >> ip >> l) || tag != "tah-tune-4" || k != key) continue;
    o = tah_plan_options{f, m, d, c, 0, w, ip, l};
    found = true;  // the last entry wins
  }
  return found;
}

// The candidates: every combination, the bounded copy grid only beside
// forked interpolations (it has nothing to leave its wave slots to
// otherwise), variable pairs only for merged launches (the only ones that
// take them).
std::vector<tah_plan_options> candidates() {
  std::vector<tah_plan_options> out;
  for (int f = 0; f <= 1; ++f)
    for (int m = 0; m <= 1; ++m)
      for (int ip = m; ip >= 0; --ip)
        for (int d = 0; d <= 1; ++d)
          for (int c : {4, 8, 16})
            for (int w : {0, 512, 1024, 2048})
              for (int l = 3; l >= 0; --l) {
                if (w && !f) continue;
                out.push_back(tah_plan_options{f, m, d, c, 0, w, m ? ip : 1, l});
              }
  return out;
}

}  // namespace

extern "C" {

tah_status tah_plan_get_options(const tah_plan* p, tah_plan_options* out) {
  return guard([&] {
    if (!p || !out) fail("null argument");
    out->fork = p->fork;
    out->merge = p->merge;
    out->defer = p->defer;
    out->copy_per_thread = p->copy_per_thread;
    out->fuse = 0;
    out->copy_wgs = p->copy_wgs;
    out->interp_pairs = p->interp_pairs;
    out->lds = p->lds;
    out->queue = p->queue;
    out->reserved_[0] = out->reserved_[1] = out->reserved_[2] = 0;
  });
}

tah_status tah_plan_set_options(tah_plan* p, const tah_plan_options* o) {
  return guard([&] {
    if (!p || !o) fail("null argument");
    HIPCHECK(hipSetDevice(p->ctx->device));
    set_options(p, *o);
  });
}

tah_status tah_tune(tah_plan* p, double* work, const double* factors, const char* cache_path,
                    double* best_ms) {
  return guard([&] {
    if (!p || !work) fail("null argument");
    if (p->regrid) fail("tah_tune times fills; this plan is a regrid transfer's");
    if (p->needs_factors && !factors)
      fail("this schedule mirrors ghosts across a reflecting face, but no factor table was given");
    HIPCHECK(hipSetDevice(p->ctx->device));
    const std::string key = signature(p);
    tah_plan_options o{};
    if (cache_path && lookup(cache_path, key, o)) {
      o.queue = p->queue;  // a batch setting: the tuner times single fills
      set_options(p, o);
      if (best_ms) *best_ms = -1.0;
      return;
    }
    hipStream_t s;
    HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    tah_plan_options best{p->fork, p->merge, p->defer, p->copy_per_thread, 0, p->copy_wgs, p->interp_pairs, p->lds,
                          p->queue, {0, 0, 0}};
    float best_t = 1e30f;
    auto run = [&]() {
      tah_status st = tah_fill(p, work, factors, s);
      if (st != TAH_OK) fail(std::string("tah_tune: a trial fill failed: ") + tah_last_error(), st);
    };
    try {
      for (tah_plan_options cand : candidates()) {
        cand.queue = p->queue;
        set_options(p, cand);
        for (int w = 0; w < 2; ++w) run();
        std::vector<float> ts;
        for (int r = 0; r < 5; ++r) {
          HIPCHECK(hipEventRecord(e0, s));
          run();
          HIPCHECK(hipEventRecord(e1, s));
          HIPCHECK(hipEventSynchronize(e1));
          float ms = 0.f;
          HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
          ts.push_back(ms);
        }
        std::sort(ts.begin(), ts.end());
        if (ts[2] < best_t) {
          best_t = ts[2];
          best = cand;
        }
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      (void)hipStreamDestroy(s);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHECK(hipStreamSynchronize(s));
    (void)hipStreamDestroy(s);
    set_options(p, best);
    if (best_ms) *best_ms = best_t;
    if (cache_path) {
      std::ofstream out(cache_path, std::ios::app);
      if (!out) fail(std::string("tah_tune: cannot write the cache file ") + cache_path);
      out << "tah-tune-4 " << key << " " << best.fork << " " << best.merge << " " << best.defer << " "
          << best.copy_per_thread << " " << best.copy_wgs << " " << best.interp_pairs << " " << best.lds << " "
          << best_t << "\n";
    }
  });
}

}  // extern "C"
