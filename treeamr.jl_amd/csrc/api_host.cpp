// C ABI, host side: forests, operators, schedules and their flattened
// descriptors. Nothing here touches a GPU.
#include <cmath>
#include <cstring>
#include <memory>

#include "api_common.hpp"
#include "schedule.hpp"

namespace tah {
namespace {
thread_local std::string last_error;
}
void set_last_error(const std::string& m) { last_error = m; }
}  // namespace tah

using namespace tah;

struct tah_forest {
  std::unique_ptr<Forest> f;
};

// A flattened, 0-based descriptor and the arrays it points into.
struct FlatDesc {
  std::vector<std::array<std::vector<int32_t>, kMaxDim>> srcstart0;
  std::vector<tah_group_desc> gdesc;
  std::vector<int32_t> tags, sgfirst, sglast;
  tah_schedule_desc desc{};
};

struct tah_schedule {
  std::unique_ptr<GhostSchedule> s;
  // The flattened, 0-based form handed out through tah_schedule_desc_get.
  std::vector<std::array<std::vector<int32_t>, kMaxDim>> srcstart0;
  std::vector<tah_group_desc> gdesc;
  std::vector<int32_t> tags, sgfirst, sglast;
  std::vector<tah_boundary_region> bregions;
  tah_schedule_desc desc{};
  // Stages P and 0 alone, in the working array's indices (0) or a state
  // vector's (1), built on first request (tah_unify_desc_get).
  std::unique_ptr<FlatDesc> unify[2];
};

namespace {

// The flattened, 0-based descriptor of a built schedule.
void flatten(tah_schedule* S, int nblocks) {
  GhostSchedule& gs = *S->s;
  S->srcstart0.resize(gs.groups.size());
  S->gdesc.resize(gs.groups.size());
  for (const Stage& st : gs.stages) S->tags.push_back(st.tag);
  for (size_t i = 0; i < gs.groups.size(); ++i) {
    const TransferGroup& tg = gs.groups[i];
    tah_group_desc& gd = S->gdesc[i];
    gd.kind = tg.kind;
    gd.stage = -1;
    for (size_t k = 0; k < gs.stages.size(); ++k)
      if (gs.stages[k].tag == tg.stage) gd.stage = int32_t(k);
    gd.factorcol = tg.factorcol;
    gd.ntransfers = int32_t(tg.targetblocks.size());
    gd.targetblocks = tg.targetblocks.data();
    gd.sourceblocks = tg.sourceblocks.data();
    for (int d = 0; d < kMaxDim; ++d) {
      const Stencil1D& s = tg.stencils[d];
      auto& v = S->srcstart0[i][d];
      v.resize(s.srcstart.size());
      for (size_t r = 0; r < v.size(); ++r) v[r] = s.srcstart[r] - 1;
      gd.stencils[d].targetfirst = s.targetfirst - 1;
      gd.stencils[d].ntarget = s.ntarget();
      gd.stencils[d].order = s.order;
      gd.stencils[d].unit = s.unit ? 1 : 0;
      gd.stencils[d].srcstart = v.data();
      gd.stencils[d].weights = s.weights.data();
    }
  }
  for (auto& sg : gs.stagegroups) {
    S->sgfirst.push_back(sg[0]);
    S->sglast.push_back(sg[1]);
  }
  tah_schedule_desc& d = S->desc;
  d.dim = gs.D;
  d.nblocks = nblocks;
  for (int k = 0; k < kMaxDim; ++k) d.stored[k] = gs.stored[k];
  d.nstages = int32_t(S->tags.size());
  d.stage_tags = S->tags.data();
  d.nstagegroups = int32_t(S->sgfirst.size());
  d.stagegroup_first = S->sgfirst.data();
  d.stagegroup_last = S->sglast.data();
  d.ngroups = int32_t(S->gdesc.size());
  d.groups = S->gdesc.data();
  d.generation = gs.generation;
  d.nsources = gs.nsources;
  int64_t first = 0;
  for (const BoundaryRegion& r : gs.boundaries) {
    tah_boundary_region e{};
    e.block = r.block;
    int64_t n = 1;
    for (int k = 0; k < kMaxDim; ++k) {
      e.direction[k] = r.direction[k];
      e.lo[k] = r.lo[k] - 1;
      e.hi[k] = r.hi[k];
      e.clamps[k] = r.clamps[k] - 1;
      n *= std::max(0, e.hi[k] - e.lo[k]);
    }
    e.slot = r.slot;
    e.first = first;
    first += n;
    S->bregions.push_back(e);
  }
  d.nboundaries = int32_t(S->bregions.size());
  d.boundaries = S->bregions.empty() ? nullptr : S->bregions.data();
  d.boundary_points = first;
}

// Stages P and 0 of a schedule (unify_arrays!, src/ghosts.jl:1874-1906): on
// the working array, or with `state` on a state-shaped array, every index
// shifted by -G and the stored size the own extent (state_unify_stages,
// src/unify.jl).
std::unique_ptr<FlatDesc> unify_desc(const tah_schedule* S, bool state) {
  const GhostSchedule& gs = *S->s;
  auto U = std::make_unique<FlatDesc>();
  std::vector<int> map(gs.stages.size(), -1);
  for (size_t k = 0; k < gs.stages.size(); ++k)
    if (gs.stages[k].tag == kStageP || gs.stages[k].tag == kUnifyTag) {
      map[k] = int(U->tags.size());
      U->tags.push_back(gs.stages[k].tag);
    }
  if (U->tags.empty() || U->tags.back() != kUnifyTag)
    fail("this schedule has no stage 0: its set has no vertex-like dimension, so nothing to unify");
  int own[kMaxDim];
  for (int d = 0; d < kMaxDim; ++d) own[d] = d < gs.D ? gs.N + gs.c[d] : 1;
  for (size_t i = 0; i < gs.groups.size(); ++i) {
    const TransferGroup& tg = gs.groups[i];
    int k = -1;
    for (size_t q = 0; q < gs.stages.size(); ++q)
      if (gs.stages[q].tag == tg.stage) k = map[q];
    if (k < 0) continue;
    tah_group_desc gd{};
    gd.kind = tg.kind;
    gd.stage = k;
    gd.factorcol = tg.factorcol;
    gd.ntransfers = int32_t(tg.targetblocks.size());
    gd.targetblocks = tg.targetblocks.data();
    gd.sourceblocks = tg.sourceblocks.data();
    U->srcstart0.emplace_back();
    auto& ss = U->srcstart0.back();
    for (int d = 0; d < kMaxDim; ++d) {
      const Stencil1D& s = tg.stencils[d];
      const int shift = state && d < gs.D ? gs.G[d] : 0;
      ss[d].resize(s.srcstart.size());
      for (size_t r = 0; r < ss[d].size(); ++r) {
        ss[d][r] = s.srcstart[r] - 1 - shift;
        if (state && (ss[d][r] < 0 || ss[d][r] + s.order > own[d]))
          fail("a stage-0 transfer reads outside the own range; this is a bug");
      }
      gd.stencils[d].targetfirst = s.targetfirst - 1 - shift;
      if (state && (gd.stencils[d].targetfirst < 0 || gd.stencils[d].targetfirst + s.ntarget() > own[d]))
        fail("a stage-0 transfer writes outside the own range; this is a bug");
      gd.stencils[d].ntarget = s.ntarget();
      gd.stencils[d].order = s.order;
      gd.stencils[d].unit = s.unit ? 1 : 0;
      gd.stencils[d].weights = s.weights.data();
    }
    U->gdesc.push_back(gd);
  }
  for (size_t i = 0; i < U->gdesc.size(); ++i)
    for (int d = 0; d < kMaxDim; ++d) U->gdesc[i].stencils[d].srcstart = U->srcstart0[i][d].data();
  for (size_t k = 0; k < U->tags.size(); ++k) {
    U->sgfirst.push_back(int32_t(k));
    U->sglast.push_back(int32_t(k));
  }
  tah_schedule_desc& d = U->desc;
  d = S->desc;
  for (int k = 0; k < kMaxDim; ++k) d.stored[k] = state ? own[k] : gs.stored[k];
  d.nstages = int32_t(U->tags.size());
  d.stage_tags = U->tags.data();
  d.nstagegroups = int32_t(U->sgfirst.size());
  d.stagegroup_first = U->sgfirst.data();
  d.stagegroup_last = U->sglast.data();
  d.ngroups = int32_t(U->gdesc.size());
  d.groups = U->gdesc.data();
  d.nboundaries = 0;
  d.boundaries = nullptr;
  d.boundary_points = 0;
  return U;
}

}  // namespace

extern "C" {

const char* tah_last_error(void) { return tah::last_error.c_str(); }

tah_status tah_forest_create(int32_t dim, const int32_t* roots, int32_t N, const int32_t* periodic,
                             const int32_t* reflecting, const int32_t* outflow,
                             const double* extents, tah_forest** out) {
  return guard([&] {
    if (!roots || !out) fail("null argument");
    if (dim < 1 || dim > kMaxDim) fail("dimension must be 1, 2 or 3");
    int r[kMaxDim];
    bool per[kMaxDim], refl[2 * kMaxDim], outf[2 * kMaxDim];
    for (int d = 0; d < dim; ++d) {
      r[d] = roots[d];
      per[d] = periodic ? periodic[d] != 0 : false;
      for (int s = 0; s < 2; ++s) {
        refl[2 * d + s] = reflecting ? reflecting[2 * d + s] != 0 : false;
        outf[2 * d + s] = outflow ? outflow[2 * d + s] != 0 : false;
      }
    }
    auto f = std::make_unique<tah_forest>();
    f->f = std::make_unique<Forest>(dim, r, N, per, refl, outf, extents);
    *out = f.release();
  });
}

tah_status tah_forest_destroy(tah_forest* f) {
  delete f;
  return TAH_OK;
}

tah_status tah_forest_info(const tah_forest* f, int32_t* nleaves, int64_t* generation) {
  return guard([&] {
    if (!f) fail("null forest");
    if (nleaves) *nleaves = f->f->nleaves();
    if (generation) *generation = f->f->generation;
  });
}

tah_status tah_forest_leaves(const tah_forest* f, int32_t* root, int32_t* level, uint32_t* coords) {
  return guard([&] {
    if (!f) fail("null forest");
    const auto& L = f->f->leaves;
    for (size_t i = 0; i < L.size(); ++i) {
      if (root) root[i] = L[i].root;
      if (level) level[i] = L[i].level;
      if (coords)
        for (int d = 0; d < kMaxDim; ++d) coords[3 * i + d] = L[i].c[d];
    }
  });
}

namespace {

// A double as an exact integer times a power of two.
void exact_parts(double x, __int128& m, int& e) {
  int ex = 0;
  const double fr = std::frexp(x, &ex);  // x = fr * 2^ex, 0.5 <= |fr| < 1
  m = (__int128)std::ldexp(fr, 53);
  e = ex - 53;
}

// The anchor of dimension d (src/geometry.jl:76-91): m = -lo*n/(hi-lo), n
// = roots*N level-0 spacings, when that is an integer in 0..n (zero is a
// node of the level-0 grid), else 0. Exact: lo and hi are brought to one
// power of two; a difference of exponents beyond 60 bits leaves m tiny or
// far outside 0..n (or lo zero: m = 0 either way).
int64_t position_anchor(double lo, double hi, int64_t n) {
  if (lo == 0.0 || !(hi > lo)) return 0;
  __int128 L, H;
  int el, eh;
  exact_parts(lo, L, el);
  exact_parts(hi, H, eh);
  const int e = std::min(el, eh);
  if (el - e > 60 || eh - e > 60) return 0;
  L <<= (el - e);
  H <<= (eh - e);
  const __int128 num = -L * (__int128)n, den = H - L;  // den > 0
  if (num % den != 0) return 0;
  const __int128 m = num / den;
  return (m >= 0 && m <= n) ? int64_t(m) : 0;
}

}  // namespace

tah_status tah_forest_positions(const tah_forest* f, int32_t d, int32_t G, int32_t c, int32_t S, double* out) {
  return guard([&] {
    if (!f || !out) fail("null argument");
    const Forest& F = *f->f;
    if (d < 0 || d >= F.D) fail("dimension out of range");
    if (G < 0 || S < 1 || (c != 0 && c != 1)) fail("bad ghost width, stored size or centering");
    const int64_t n = int64_t(F.roots[d]) * F.N;
    const int64_t m = position_anchor(F.extents[d][0], F.extents[d][1], n);
    const double lo = m > 0 ? 0.0 : F.extents[d][0];
    // root_spacing (src/geometry.jl:19-20): from dimension 1's extent and roots
    const double h0 = (F.extents[0][1] - F.extents[0][0]) / F.roots[0] / F.N;
    for (size_t b = 0; b < F.leaves.size(); ++b) {
      const Key& k = F.leaves[b];
      const int lvl = k.level;
      const double h = std::ldexp(h0, -lvl);
      const int64_t rootpos = F.root_position(k.root)[d];
      const int64_t B = ((rootpos << lvl) + int64_t(k.c[d])) * F.N - (m << lvl);
      for (int32_t j = 0; j < S; ++j) {
        const int64_t kk = B + j - G;
        out[b * S + j] = c == 1 ? std::fma(double(kk), h, lo) : std::fma(double(2 * kk + 1), h / 2, lo);
      }
    }
  });
}

tah_status tah_forest_refine(tah_forest* f, int32_t n, const int32_t* leaves) {
  return guard([&] {
    if (!f) fail("null forest");
    std::vector<Key> keys;
    for (int32_t i = 0; i < n; ++i) {
      if (leaves[i] < 0 || leaves[i] >= f->f->nleaves()) fail("leaf index out of range");
      keys.push_back(f->f->leaves[leaves[i]]);
    }
    f->f->refine(keys);
  });
}

tah_status tah_forest_balance(tah_forest* f) {
  return guard([&] {
    if (!f) fail("null forest");
    f->f->balance();
  });
}

tah_status tah_forest_isbalanced(const tah_forest* f, int32_t* balanced) {
  return guard([&] {
    if (!f || !balanced) fail("null argument");
    *balanced = f->f->isbalanced() ? 1 : 0;
  });
}

tah_status tah_lagrange_weights(int32_t p, int64_t num, int64_t den, double* out) {
  return guard([&] {
    if (!out || p < 1 || den == 0) fail("bad argument");
    const auto& w = unit_lagrange_weights(p, Rational(num, den));
    for (int i = 0; i < p; ++i) out[i] = w[i].to_double();
  });
}

tah_status tah_conservative_weights(int32_t p, double* low, double* high) {
  return guard([&] {
    if (!low || !high || p < 1 || p % 2 == 0) fail("conservative order must be odd and positive");
    std::vector<Rational> lo, hi;
    conservative_prolong_weights(p, lo, hi);
    for (int i = 0; i < p; ++i) {
      low[i] = lo[i].to_double();
      high[i] = hi[i].to_double();
    }
  });
}

tah_status tah_operators_check(const tah_operators* ops) {
  return guard([&] {
    if (!ops) fail("null operators");
    Operators o(ops->family, ops->prolongation, ops->restriction, ops->bounded != 0,
                ops->vertex_prolongation);
    (void)o;
  });
}

static std::vector<std::array<int8_t, kMaxDim>> parities_of(int dim, int nvars, const int8_t* parity) {
  std::vector<std::array<int8_t, kMaxDim>> par;
  if (!parity) return par;
  for (int v = 0; v < nvars; ++v) {
    std::array<int8_t, kMaxDim> p{0, 0, 0};
    for (int d = 0; d < dim; ++d) {
      p[d] = parity[v * dim + d];
      if (p[d] < 0 || p[d] > 2) fail("parity must be none (0), even (1) or odd (2)");
    }
    par.push_back(p);
  }
  return par;
}

tah_status tah_schedule_create(const tah_forest* f, const tah_operators* ops, const int32_t* G,
                               const int32_t* centering, int32_t nvars, const int8_t* parity,
                               tah_schedule** out) {
  return guard([&] {
    if (!f || !ops || !G || !out) fail("null argument");
    const Forest& F = *f->f;
    if (nvars <= 0) fail("nvars must be positive");
    Operators o(ops->family, ops->prolongation, ops->restriction, ops->bounded != 0,
                ops->vertex_prolongation);
    int g[kMaxDim] = {0, 0, 0}, c[kMaxDim] = {0, 0, 0};
    for (int d = 0; d < F.D; ++d) {
      g[d] = G[d];
      c[d] = centering ? centering[d] : 0;
    }
    auto par = parities_of(F.D, nvars, parity);
    auto S = std::make_unique<tah_schedule>();
    S->s = std::make_unique<GhostSchedule>(F, o, g, c, par);
    flatten(S.get(), F.nleaves());
    *out = S.release();
  });
}

tah_status tah_regrid_schedule_create(const tah_forest* f, int32_t nold, const int32_t* old_root,
                                      const int32_t* old_level, const uint32_t* old_coords,
                                      const tah_operators* ops, const int32_t* G, const int32_t* centering,
                                      int32_t nvars, const int8_t* parity, tah_schedule** out) {
  return guard([&] {
    if (!f || !ops || !G || !out || !old_root || !old_level || !old_coords) fail("null argument");
    const Forest& F = *f->f;
    if (nvars <= 0) fail("nvars must be positive");
    if (nold <= 0) fail("a regrid needs the leaves before it");
    std::vector<Key> old(nold);
    for (int32_t i = 0; i < nold; ++i) {
      old[i].root = old_root[i];
      old[i].level = old_level[i];
      for (int d = 0; d < kMaxDim; ++d) old[i].c[d] = d < F.D ? old_coords[3 * i + d] : 0;
    }
    Operators o(ops->family, ops->prolongation, ops->restriction, ops->bounded != 0,
                ops->vertex_prolongation);
    int g[kMaxDim] = {0, 0, 0}, c[kMaxDim] = {0, 0, 0};
    for (int d = 0; d < F.D; ++d) {
      g[d] = G[d];
      c[d] = centering ? centering[d] : 0;
    }
    auto par = parities_of(F.D, nvars, parity);
    auto S = std::make_unique<tah_schedule>();
    S->s = std::make_unique<GhostSchedule>(RegridFrom{old}, F, o, g, c, par);
    flatten(S.get(), F.nleaves());
    *out = S.release();
  });
}

tah_status tah_unify_schedule_create(const tah_forest* f, const int32_t* G, const int32_t* centering,
                                     int32_t nvars, const int8_t* parity, tah_schedule** out) {
  return guard([&] {
    if (!f || !G || !centering || !out) fail("null argument");
    const Forest& F = *f->f;
    if (nvars <= 0) fail("nvars must be positive");
    int g[kMaxDim] = {0, 0, 0}, c[kMaxDim] = {0, 0, 0};
    for (int d = 0; d < F.D; ++d) {
      g[d] = G[d];
      c[d] = centering[d];
    }
    auto par = parities_of(F.D, nvars, parity);
    auto S = std::make_unique<tah_schedule>();
    S->s = std::make_unique<GhostSchedule>(UnifyOnly{}, F, g, c, par);
    flatten(S.get(), F.nleaves());
    *out = S.release();
  });
}

tah_status tah_unify_desc_get(tah_schedule* s, int32_t state, const tah_schedule_desc** out) {
  return guard([&] {
    if (!s || !out) fail("null argument");
    const int k = state ? 1 : 0;
    if (!s->unify[k]) s->unify[k] = unify_desc(s, state != 0);
    *out = &s->unify[k]->desc;
  });
}

tah_status tah_forest_coarsen(tah_forest* f, int32_t n, const int32_t* leaves, int32_t* coarsened) {
  return guard([&] {
    if (!f || (n > 0 && !leaves)) fail("null argument");
    Forest& F = *f->f;
    std::vector<Key> keys;
    for (int32_t i = 0; i < n; ++i) {
      if (leaves[i] < 0 || leaves[i] >= F.nleaves()) fail("leaf index out of range");
      keys.push_back(F.leaves[leaves[i]]);
    }
    const int m = F.coarsen(keys);
    if (coarsened) *coarsened = m;
  });
}

tah_status tah_schedule_destroy(tah_schedule* s) {
  delete s;
  return TAH_OK;
}

tah_status tah_schedule_check(const tah_schedule* s) {
  return guard([&] {
    if (!s) fail("null schedule");
    if (s->s->forest->generation != s->s->generation)
      fail("the schedule is stale: the forest has changed since it was built; rebuild it", kStale);
  });
}

tah_status tah_parity_factors(int32_t dim, int32_t nvars, const int8_t* parity, double* out) {
  return guard([&] {
    if (!parity || !out) fail("null argument");
    if (dim < 1 || dim > kMaxDim) fail("dimension must be 1, 2 or 3");
    auto f = parity_factors(dim, parities_of(dim, nvars, parity));
    std::memcpy(out, f.data(), f.size() * sizeof(double));
  });
}

tah_status tah_schedule_desc_get(const tah_schedule* s, const tah_schedule_desc** out) {
  return guard([&] {
    if (!s || !out) fail("null argument");
    *out = &s->desc;
  });
}

tah_status tah_schedule_boundaries(const tah_schedule* s, int32_t* n, int32_t* block, int32_t* direction,
                                   int32_t* lo, int32_t* hi) {
  return guard([&] {
    if (!s || !n) fail("null argument");
    const auto& B = s->s->boundaries;
    *n = int32_t(B.size());
    for (size_t i = 0; i < B.size(); ++i) {
      if (block) block[i] = B[i].block;
      for (int d = 0; d < kMaxDim; ++d) {
        if (direction) direction[3 * i + d] = B[i].direction[d];
        if (lo) lo[3 * i + d] = B[i].lo[d] - 1;
        if (hi) hi[3 * i + d] = B[i].hi[d];
      }
    }
  });
}

tah_status tah_schedule_points(const tah_schedule* s, int64_t* points) {
  return guard([&] {
    if (!s || !points) fail("null argument");
    int64_t n = 0;
    for (const auto& g : s->s->groups) n += g.points() * int64_t(g.targetblocks.size());
    *points = n;
  });
}

}  // extern "C"
