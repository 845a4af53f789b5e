// Plan options and the autotuner.
//
// A plan's launches have plan-wide knobs -- stage groups forked over the
// context's streams or run in order on one; phase-1 transfers deferred into
// the last stage group or not; the copy kernel's tile (4, 8 or 16 values per
// thread); the copies from one workgroup per tile or from a bounded grid
// that leaves wave slots to the interpolations beside them -- and knobs per
// interpolation family (order and direction): the workgroup size, the unit
// classes in one launch or a launch each, variable pairs, LDS staging.
// tah_tune times every plan-wide combination on the caller's working array
// and keeps the fastest, then searches each family's knobs by the time of
// that family's own launches, and keeps those where the whole fill is no
// slower; cached in a text file under a signature of the mesh, the layout
// and the device, so that a second run over the same mesh skips the timing.
// The reference has no autotuner (its launches are fixed,
// src/ghosts.jl:404-521).
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
  if (o.lds < 0 || o.lds > 15)
    fail("plan option lds is 0 to 15 (bits 0 and 1: restriction and prolongation from LDS chunks; bits 2 and "
         "3: restriction and prolongation streamed along z)");
  if (o.queue < 0 || o.queue > 1) fail("plan option queue is 0 or 1");
  if (o.copy_order < 0 || o.copy_order > 1) fail("plan option copy_order is 0 or 1");
  if (o.reserved_[0] || o.reserved_[1]) fail("tah_plan_options.reserved_ must be zero");
  bool rebuild = o.merge != p->merge || o.defer != p->defer || o.copy_per_thread != p->copy_per_thread ||
                 wgs != p->copy_wgs || o.interp_pairs != p->interp_pairs || o.lds != p->lds ||
                 o.copy_order != p->copy_order;
  p->fork = o.fork;
  p->merge = o.merge;
  p->defer = o.defer;
  p->copy_per_thread = o.copy_per_thread;
  p->copy_wgs = wgs;
  p->copy_order = o.copy_order;
  p->interp_pairs = o.interp_pairs;
  p->lds = o.lds;
  if (o.queue != p->queue) {
    // batches built from this plan are taken again with the new mode
    p->queue = o.queue;
    plan_register(p);
  }
  // every family to the plan-wide options
  const auto th = p->fam_threads, me = p->fam_merge, pa = p->fam_pairs, ld = p->fam_lds, va = p->fam_vars;
  families_from_options(p);
  rebuild = rebuild || th != p->fam_threads || me != p->fam_merge || pa != p->fam_pairs || ld != p->fam_lds ||
            va != p->fam_vars;
  if (rebuild) build_launches(p);
}

// The family of a launch class (-1: copies, generic transfers).
int family_of_class(int cls) {
  if (cls >= kClsStream && cls < kClsStream + kFamilies) return cls - kClsStream;
  if (cls >= kClsLdsProlong && cls < kClsLdsProlong + kFamilies) return cls - kClsLdsProlong;
  if (cls >= kClsLds && cls < kClsLds + kFamilies) return cls - kClsLds;
  if (cls >= kClsMergedPair && cls < kClsMergedPair + kFamilies) return cls - kClsMergedPair;
  if (cls >= kClsMerged && cls < kClsMerged + kFamilies) return cls - kClsMerged;
  if (cls >= kClsInterp && cls < kClsInterp + 8 * kFamilies) return (cls - kClsInterp) / 8;
  return -1;
}

// The families the plan's interpolation groups fall in.
std::vector<int> plan_families(const tah_plan* p) {
  std::vector<int> out;
  for (const PlanWClass& c : p->wclasses) {
    const int f = family_of_class(c.cls);
    if (f >= 0 && std::find(out.begin(), out.end(), f) == out.end()) out.push_back(f);
  }
  std::sort(out.begin(), out.end());
  return out;
}

tah_family_options family_options(const tah_plan* p, int f) {
  tah_family_options o{};
  o.order = kInterpTable[size_t(f)][0];
  o.restriction = kInterpTable[size_t(f)][1] != 0;
  o.threads = p->fam_threads[size_t(f)];
  o.merge = p->fam_merge[size_t(f)];
  o.pairs = p->fam_pairs[size_t(f)];
  o.lds = p->fam_lds[size_t(f)];
  o.vars = p->fam_vars[size_t(f)];
  return o;
}

void set_family(tah_plan* p, const tah_family_options& o) {
  if (o.restriction < 0 || o.restriction > 1) fail("family option restriction is 0 or 1");
  const int f = interp_index(o.order, o.restriction ? 2 : 0);
  if (f < 0)
    fail("no interpolation family of order " + std::to_string(o.order) +
         (o.restriction ? " restricts" : " prolongates") + " here");
  if (o.threads != 64 && o.threads != 128 && o.threads != 256) fail("a family's workgroup size is 64, 128 or 256");
  if (o.merge < 0 || o.merge > 1 || o.pairs < 0 || o.pairs > 1 || o.lds < 0 || o.lds > 2)
    fail("family options merge and pairs are 0 or 1; lds is 0, 1 or 2 (streamed along z)");
  if (o.restriction && o.pairs) fail("variable pairs are for prolongation (pairs must be 0 for restriction)");
  if (o.vars < 0 || o.vars > o.threads) fail("a family's variables per tile are 0 (every one) to its threads");
  const size_t i = size_t(f);
  if (p->fam_threads[i] == o.threads && p->fam_merge[i] == o.merge && p->fam_pairs[i] == o.pairs &&
      p->fam_lds[i] == o.lds && p->fam_vars[i] == o.vars)
    return;
  p->fam_threads[i] = o.threads;
  p->fam_merge[i] = o.merge;
  p->fam_pairs[i] = o.pairs;
  p->fam_lds[i] = o.lds;
  p->fam_vars[i] = o.vars;
  build_launches(p);
}

// A cache line: "tah-tune-7 <signature> fork merge defer cpt wgs pairs lds
// copy_order <families> (order restriction threads merge pairs lds vars)...
// <ms>".
bool lookup(const std::string& path, const std::string& key, tah_plan_options& o,
            std::vector<tah_family_options>& fams) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  bool found = false;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string tag, k;
    int f, m, d, c, w, ip, l, co, nf;
    if (!(ss >> tag >> k >> f >> m >> d >> c >> w >> ip >> l >> co >> nf) || tag != "tah-tune-7" || k != key ||
        nf < 0 || nf > kFamilies)
      continue;
    std::vector<tah_family_options> fs(static_cast<size_t>(nf));
    bool ok = true;
    for (auto& e : fs)
      ok = ok && bool(ss >> e.order >> e.restriction >> e.threads >> e.merge >> e.pairs >> e.lds >> e.vars);
    if (!ok) continue;
    o = tah_plan_options{f, m, d, c, 0, w, ip, l, 0, co, {0, 0}};
    fams = std::move(fs);
    found = true;  // the last entry wins
  }
  return found;
}

// The plan-wide candidates: forked or not, deferral, the copy tile and
// the copies' bounded grid (only beside forked interpolations: it has
// nothing to leave its wave slots to otherwise). The interpolations' own
// options are the families' and searched after these.
std::vector<tah_plan_options> candidates(const tah_plan* p) {
  std::vector<tah_plan_options> out;
  for (int f = 0; f <= 1; ++f)
    for (int d = 0; d <= 1; ++d)
      for (int c : {4, 8, 16})
        for (int w : {0, 512, 1024, 2048}) {
          if (w && !f) continue;
          out.push_back(
              tah_plan_options{f, p->merge, d, c, 0, w, p->interp_pairs, p->lds, p->queue, p->copy_order, {0, 0}});
        }
  return out;
}

// One family's knobs and their values: workgroup size, variables per tile
// (where variables are innermost), one launch or one per unit class,
// variable pairs (prolongation, where the layout allows), LDS staging
// (where the layout allows and a staged kernel exists).
std::vector<std::vector<int32_t>> family_knobs(const tah_plan* p, int f) {
  const int64_t* st = p->layout.stride;
  const bool vfast = st[3] < st[0];
  const bool staged_ok = st[3] == 1 && p->layout.nvars % 2 == 0 && st[0] % 2 == 0 && st[1] % 2 == 0 &&
                         st[2] % 2 == 0 && st[4] % 2 == 0;
  const bool pairs_ok = st[3] == 1 && p->layout.nvars > 1;
  const int P = kInterpTable[size_t(f)][0];
  const bool restriction = kInterpTable[size_t(f)][1] != 0;
  const bool lds_ok = staged_ok && (restriction ? (P == 2 || P == 4 || P == 6) : (P >= 2 && P <= 6));
  std::vector<int32_t> vars{0};
  if (vfast)
    for (int v : {8, 4, 2, 1})
      if (v < p->layout.nvars) vars.push_back(v);
  return {{256, 128, 64}, vars, {1, 0}, (!restriction && pairs_ok) ? std::vector<int32_t>{1, 0} : std::vector<int32_t>{0},
          lds_ok ? std::vector<int32_t>{2, 0, 1} : std::vector<int32_t>{2, 0}};
}

int32_t& knob(tah_family_options& o, int k) {
  switch (k) {
    case 0: return o.threads;
    case 1: return o.vars;
    case 2: return o.merge;
    case 3: return o.pairs;
    default: return o.lds;
  }
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
    out->copy_order = p->copy_order;
    out->reserved_[0] = out->reserved_[1] = 0;
  });
}

tah_status tah_plan_set_options(tah_plan* p, const tah_plan_options* o) {
  return guard([&] {
    if (!p || !o) fail("null argument");
    HIPCHECK(hipSetDevice(p->ctx->device));
    set_options(p, *o);
  });
}

tah_status tah_plan_get_family_options(const tah_plan* p, int32_t* n, tah_family_options* out) {
  return guard([&] {
    if (!p || !n) fail("null argument");
    const std::vector<int> fams = plan_families(p);
    if (!out) {
      *n = int32_t(fams.size());
      return;
    }
    if (*n < int32_t(fams.size()))
      fail("the array holds " + std::to_string(*n) + " families; the plan has " + std::to_string(fams.size()));
    for (size_t k = 0; k < fams.size(); ++k) out[k] = family_options(p, fams[k]);
    *n = int32_t(fams.size());
  });
}

tah_status tah_plan_set_family_options(tah_plan* p, const tah_family_options* o) {
  return guard([&] {
    if (!p || !o) fail("null argument");
    HIPCHECK(hipSetDevice(p->ctx->device));
    set_family(p, *o);
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
    std::vector<tah_family_options> cached;
    if (cache_path && lookup(cache_path, key, o, cached)) {
      o.queue = p->queue;  // a batch setting: the tuner times single fills
      set_options(p, o);
      for (const auto& f : cached) set_family(p, f);
      if (best_ms) *best_ms = -1.0;
      return;
    }
    hipStream_t s;
    HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    auto run = [&]() {
      tah_status st = tah_fill(p, work, factors, s);
      if (st != TAH_OK) fail(std::string("tah_tune: a trial fill failed: ") + tah_last_error(), st);
    };
    // the median of five fills after two unmeasured ones
    auto time_fill = [&]() {
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
      return ts[2];
    };
    // the median over five fills of family f's own launches, timed one by one
    auto time_family = [&](int f) {
      int32_t n = 0;
      tah_status st = tah_fill_timed(p, work, factors, s, &n, nullptr, nullptr, nullptr);
      if (st != TAH_OK) fail(std::string("tah_tune: ") + tah_last_error(), st);
      std::vector<int32_t> cls(size_t(std::max(1, n)));
      std::vector<int64_t> vals(cls.size());
      std::vector<float> ms(cls.size());
      std::vector<float> ts;
      for (int r = 0; r < 6; ++r) {
        int32_t m = n;
        st = tah_fill_timed(p, work, factors, s, &m, cls.data(), vals.data(), ms.data());
        if (st != TAH_OK) fail(std::string("tah_tune: a trial fill failed: ") + tah_last_error(), st);
        if (r == 0) continue;  // unmeasured
        float t = 0.f;
        for (int32_t k = 0; k < m; ++k)
          if (family_of_class(cls[size_t(k)]) == f) t += ms[size_t(k)];
        ts.push_back(t);
      }
      std::sort(ts.begin(), ts.end());
      return ts[2];
    };
    tah_plan_options best{p->fork, p->merge, p->defer, p->copy_per_thread, 0, p->copy_wgs, p->interp_pairs, p->lds,
                          p->queue, p->copy_order, {0, 0}};
    std::vector<tah_family_options> best_fams;
    float best_t = 1e30f;
    try {
      // 1. the plan-wide options, every family at the plan's own
      for (const tah_plan_options& cand : candidates(p)) {
        set_options(p, cand);
        const float t = time_fill();
        if (t < best_t) {
          best_t = t;
          best = cand;
        }
      }
      set_options(p, best);
      // 2. from there, each family by the time of its own launches: one
      // knob at a time over its values, the others at their best so far,
      // twice round
      for (int f : plan_families(p)) {
        tah_family_options keep = family_options(p, f);
        float keep_t = time_family(f);
        const auto knobs = family_knobs(p, f);
        for (int round = 0; round < 2; ++round)
          for (int k = 0; k < int(knobs.size()); ++k)
            for (int32_t v : knobs[size_t(k)]) {
              tah_family_options c = keep;
              knob(c, k) = v;
              if (knob(c, k) == knob(keep, k) || (c.pairs && !c.merge)) continue;  // pairs are a merged launch's
              set_family(p, c);
              const float t = time_family(f);
              if (t < keep_t) {
                keep_t = t;
                keep = c;
              }
            }
        set_family(p, keep);
      }
      // 3. each family's choice where the whole fill is no slower with it
      // than with the family at the plan-wide options: the two timed in
      // turn, three times each, the better of each's medians compared
      // (the fill's time is what counts when launches overlap)
      for (int f : plan_families(p)) {
        const tah_family_options chosen = family_options(p, f);
        tah_family_options base = chosen;
        base.threads = kThreads;
        base.merge = p->merge;
        base.pairs = base.restriction || p->layout.nvars % 2 ? 0 : p->interp_pairs;
        base.lds = base.restriction ? ((p->lds & 4) ? 2 : (p->lds & 1)) : ((p->lds & 8) ? 2 : (p->lds >> 1) & 1);
        base.vars = 0;
        if (std::memcmp(&chosen, &base, sizeof base) == 0) continue;
        float tc = 1e30f, tb = 1e30f;
        for (int r = 0; r < 3; ++r) {
          set_family(p, chosen);
          tc = std::min(tc, time_fill());
          set_family(p, base);
          tb = std::min(tb, time_fill());
        }
        set_family(p, tc < tb ? chosen : base);
      }
      // 4. the copies' tile order (by source or by target block), the
      // families as chosen: the two timed in turn, three times each
      {
        std::vector<tah_family_options> fams;
        for (int f : plan_families(p)) fams.push_back(family_options(p, f));
        auto with_order = [&](int32_t order) {
          tah_plan_options o = best;
          o.copy_order = order;
          set_options(p, o);  // sets every family to the plan-wide options
          for (const auto& fo : fams) set_family(p, fo);
        };
        float t0 = 1e30f, t1 = 1e30f;
        for (int r = 0; r < 3; ++r) {
          with_order(0);
          t0 = std::min(t0, time_fill());
          with_order(1);
          t1 = std::min(t1, time_fill());
        }
        best.copy_order = t1 < t0 ? 1 : 0;
        with_order(best.copy_order);
      }
      best_t = time_fill();
      for (int f : plan_families(p)) best_fams.push_back(family_options(p, f));
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
    for (const auto& f : best_fams) set_family(p, f);
    if (best_ms) *best_ms = best_t;
    if (cache_path) {
      std::ofstream out(cache_path, std::ios::app);
      if (!out) fail(std::string("tah_tune: cannot write the cache file ") + cache_path);
      out << "tah-tune-7 " << key << " " << best.fork << " " << best.merge << " " << best.defer << " "
          << best.copy_per_thread << " " << best.copy_wgs << " " << best.interp_pairs << " " << best.lds << " "
          << best.copy_order << " " << best_fams.size();
      for (const auto& f : best_fams)
        out << " " << f.order << " " << f.restriction << " " << f.threads << " " << f.merge << " " << f.pairs << " "
            << f.lds << " " << f.vars;
      out << " " << best_t << "\n";
    }
  });
}

tah_status tah_tune_batched(tah_ctx* ctx, int32_t n, tah_plan* const* plans, double* const* works,
                            const double* const* factors, const char* cache_path, double* best_ms) {
  return guard([&] {
    if (!ctx || n < 1 || !plans || !works) fail("null argument or empty batch");
    for (int32_t i = 0; i < n; ++i) {
      if (!plans[i] || !works[i]) fail("null plan or array in the batch");
      if (plans[i]->ctx != ctx) fail("a plan of the batch belongs to another context");
      if (plans[i]->regrid) fail("tah_tune_batched times fills; a plan of the batch is a regrid transfer's");
    }
    HIPCHECK(hipSetDevice(ctx->device));
    Hash H;
    H.add(n);
    for (int32_t i = 0; i < n; ++i) {
      const std::string k = signature(plans[i]);
      H.add(k.data(), k.size());
    }
    char buf[64];
    std::snprintf(buf, sizeof buf, "%016llx", static_cast<unsigned long long>(H.h));
    const std::string key = std::string(ctx->prop.gcnArchName) + "-" +
                            std::to_string(ctx->prop.multiProcessorCount) + "-batch-" + buf;
    auto apply = [&](int32_t fork, int32_t defer, int32_t cpt, int32_t order) {
      for (int32_t i = 0; i < n; ++i) {
        tah_plan* p = plans[i];
        // set_options sets every family to the plan-wide options: keep the
        // families' own
        std::vector<tah_family_options> fams;
        for (int f : plan_families(p)) fams.push_back(family_options(p, f));
        tah_plan_options o{fork, p->merge, defer, cpt, 0, p->copy_wgs, p->interp_pairs, p->lds, p->queue, order, {0, 0}};
        set_options(p, o);
        for (const auto& f : fams) set_family(p, f);
      }
    };
    if (cache_path) {
      std::ifstream in(cache_path);
      std::string line, best_line;
      while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string tag, k;
        if (ss >> tag >> k && tag == "tah-tune-batch-2" && k == key) best_line = line;
      }
      if (!best_line.empty()) {
        std::istringstream ss(best_line);
        std::string tag, k;
        int32_t fork = 0, defer = 0, cpt = 8, order = 0;
        if (ss >> tag >> k >> fork >> defer >> cpt >> order && fork >= 0 && fork <= 1 && defer >= 0 && defer <= 1 &&
            (cpt == 4 || cpt == 8 || cpt == 16) && order >= 0 && order <= 1) {
          apply(fork, defer, cpt, order);
          if (best_ms) *best_ms = -1.0;
          return;
        }
      }
    }
    hipStream_t s;
    HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    auto run = [&]() {
      tah_status st = tah_fill_batched(ctx, n, plans, works, factors, s);
      if (st != TAH_OK) fail(std::string("tah_tune_batched: a trial fill failed: ") + tah_last_error(), st);
    };
    // the median of five batched fills after two unmeasured ones (the
    // first builds the batch for these options)
    auto time_fill = [&]() {
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
      return ts[2];
    };
    int32_t bf = plans[0]->fork, bd = plans[0]->defer, bc = plans[0]->copy_per_thread, bo = plans[0]->copy_order;
    float best_t = 1e30f;
    try {
      for (int32_t fork : {0, 1})
        for (int32_t defer : {0, 1})
          for (int32_t cpt : {4, 8, 16}) {
            apply(fork, defer, cpt, bo);
            const float t = time_fill();
            if (t < best_t) {
              best_t = t;
              bf = fork;
              bd = defer;
              bc = cpt;
            }
          }
      // then the copies' tile order at those options: the two timed in
      // turn, twice each
      float t0 = 1e30f, t1 = 1e30f;
      for (int r = 0; r < 2; ++r) {
        apply(bf, bd, bc, 0);
        t0 = std::min(t0, time_fill());
        apply(bf, bd, bc, 1);
        t1 = std::min(t1, time_fill());
      }
      bo = t1 < t0 ? 1 : 0;
      best_t = std::min(t0, t1);
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
    apply(bf, bd, bc, bo);
    if (best_ms) *best_ms = best_t;
    if (cache_path) {
      std::ofstream out(cache_path, std::ios::app);
      if (!out) fail(std::string("tah_tune_batched: cannot write the cache file ") + cache_path);
      out << "tah-tune-batch-2 " << key << " " << bf << " " << bd << " " << bc << " " << bo << " " << best_t << "\n";
    }
  });
}

}  // extern "C"
