// The ghost-fill schedule builder; see schedule.hpp.
#include "schedule.hpp"

#include <algorithm>
#include <string>
#include <unordered_map>

namespace tah {

namespace {

inline long fld(long a, long b) { long q = a / b; return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q; }
inline long cld(long a, long b) { return -fld(-a, b); }
inline long clampl(long x, long lo, long hi) { return x < lo ? lo : x > hi ? hi : x; }

struct Range {
  int lo, hi;  // 1-based inclusive; empty when hi < lo
  int size() const { return hi >= lo ? hi - lo + 1 : 0; }
};

// src/schedule.jl:929-940
Range target_range(int N, int G, int c, int dd, int od, bool halved) {
  if (dd == 1) return {G + N + c + 1, N + 2 * G + c};
  if (dd == -1) return {1, G};
  if (halved) {
    int h = N / 2;
    if (c == 1) return od == 0 ? Range{G + 1, G + 1 + h} : Range{G + 2 + h, G + N + 1};
    return {G + 1 + od * h, G + (od + 1) * h};
  }
  return {G + 1, G + N + c};
}

// src/schedule.jl:944-950 (span codes 8lo+hi of piece classes)
Range piece_range(int N, int G, int c, int code) {
  int lo = code >> 3, hi = code & 7, h = N / 2;
  if (c == 0) return {lo == 1 ? G + 1 : G + h + 1, hi == 1 ? G + h : G + N};
  const int first[5] = {G + 1, G + 2, G + h + 1, G + h + 2, G + N + 1};
  const int last[5] = {G + 1, G + h, G + h + 1, G + N, G + N + 1};
  return {first[lo], last[hi]};
}

void set_weights(Stencil1D& s, int i, const std::vector<Rational>& w) {
  for (int m = 0; m < s.order; ++m) s.weights[size_t(s.order) * i + m] = w[m].to_double();
}

Stencil1D make(int first, int n, int order) {
  Stencil1D s;
  s.targetfirst = first;
  s.order = order;
  s.srcstart.assign(n, 0);
  s.weights.assign(size_t(order) * n, 0.0);
  return s;
}

// src/schedule.jl:960-965
Stencil1D copy_stencil(int N, int G, int c, int dd, int span) {
  Range r = span >= 0 ? piece_range(N, G, c, span) : target_range(N, G, c, dd, 0, false);
  Stencil1D s = make(r.lo, r.size(), 1);
  for (int j = r.lo, i = 0; j <= r.hi; ++j, ++i) {
    s.srcstart[i] = j - N * dd;
    s.weights[i] = 1.0;
  }
  s.finish();
  return s;
}

// src/schedule.jl:1002-1025
Stencil1D restrict_stencil_over(Range r, int N, int G, int c, int dd, int od, int p) {
  int shift = -N * dd;
  if (c == 1) {
    Stencil1D s = make(r.lo, r.size(), 1);
    for (int j = r.lo, i = 0; j <= r.hi; ++j, ++i) {
      s.srcstart[i] = G + 1 + 2 * (j + shift - G - 1 - od * (N / 2));
      s.weights[i] = 1.0;
    }
    s.finish();
    return s;
  }
  Stencil1D s = make(r.lo, r.size(), p);
  for (int j = r.lo, i = 0; j <= r.hi; ++j, ++i) {
    long q = j + shift - G - 1;
    long f0 = G + 1 + 2 * (q - od * (N / 2));
    long lo = clampl(f0 - p / 2 + 1, G + 1, G + N - p + 1);
    s.srcstart[i] = int32_t(lo);
    set_weights(s, i, interpolation_weights(lo, p, Rational(2 * f0 + 1, 2), "restriction"));
  }
  s.finish();
  return s;
}

Stencil1D restrict_stencil(int N, int G, int c, int dd, int od, int p) {
  return restrict_stencil_over(target_range(N, G, c, dd, od, true), N, G, c, dd, od, p);
}

// src/schedule.jl:1035-1036
int source_direction(int dd, int od) {
  return (dd == 1 && od == 1) ? 1 : (dd == -1 && od == 0) ? -1 : 0;
}

// src/schedule.jl:1046-1079
Stencil1D prolong_stencil_over(Range r, int N, int G, int c, int dd, int od, int p) {
  long stored = N + 2 * G + c;
  Stencil1D s = make(r.lo, r.size(), p);
  for (int j = r.lo, i = 0; j <= r.hi; ++j, ++i) {
    long phi = long(od) * N + j - G - 1;
    long origin = -long(N) * source_direction(dd, od) + G + 1;
    long lo;
    Rational x;
    if (c == 1) {
      lo = clampl(fld(phi, 2) - p / 2 + 1 + origin, 1, stored - p + 1);
      x = Rational(phi + 2 * origin, 2);
    } else {
      lo = clampl(cld(phi, 2) - p / 2 + origin, 1, stored - p + 1);
      x = Rational(2 * phi - 1 + 4 * origin, 4);
    }
    s.srcstart[i] = int32_t(lo);
    set_weights(s, i, interpolation_weights(lo, p, x, "prolongation"));
  }
  s.finish();
  return s;
}

// src/schedule.jl:1122-1143
Stencil1D conservative_prolong_stencil(int N, int G, int c, int dd, int od, int p) {
  Range r = target_range(N, G, c, dd, od, false);
  std::vector<Rational> low, high;
  conservative_prolong_weights(p, low, high);
  int rr = (p - 1) / 2;
  Stencil1D s = make(r.lo, r.size(), p);
  for (int j = r.lo, i = 0; j <= r.hi; ++j, ++i) {
    long phi = long(od) * N + j - G - 1;
    long cc = fld(phi, 2);
    long subcell = phi - 2 * cc;
    long origin = -long(N) * source_direction(dd, od) + G + 1;
    s.srcstart[i] = int32_t(cc - rr + origin);
    set_weights(s, i, subcell == 0 ? low : high);
  }
  s.finish();
  return s;
}

// src/schedule.jl:1151-1162
Stencil1D prolongation_stencil(int N, int G, int c, int dd, int od, const Operators& ops, int span) {
  if (span >= 0) {
    if (c != 1) fail("a prolongation span along a cell-like dimension; this is a bug");
    return prolong_stencil_over(piece_range(N, G, c, span), N, G, c, dd, od, vertex_order(ops));
  }
  if (ops.family == Conservative && c == 0)
    return conservative_prolong_stencil(N, G, c, dd, od, ops.prolongation);
  return prolong_stencil_over(target_range(N, G, c, dd, od, false), N, G, c, dd, od,
                              vertex_order(ops));
}

// src/schedule.jl:1183-1195 (P = 0: no element sets)
Stencil1D kind_stencil(int N, int G, int c, int kind, int dd, int od, const Operators& ops, int span) {
  if (kind == KindCopy) return copy_stencil(N, G, c, dd, span);
  if (kind == KindRestrict) return restrict_stencil(N, G, c, dd, od, ops.restriction);
  return prolongation_stencil(N, G, c, dd, od, ops, span);
}

Stencil1D remap_rows(const Stencil1D& s, int first, const std::vector<int>& rows) {
  Stencil1D out = make(first, (int)rows.size(), s.order);
  for (size_t i = 0; i < rows.size(); ++i) {
    int r = rows[i];
    out.srcstart[i] = s.srcstart[r];
    for (int m = 0; m < s.order; ++m)
      out.weights[size_t(s.order) * i + m] = s.weights[size_t(s.order) * r + m];
  }
  out.finish();
  return out;
}

// src/schedule.jl:1309-1324 (outflow: every ghost row takes the nearest own row)
Stencil1D clamp_rows(const Stencil1D& s, int N, int G, int c, int dd) {
  Range r = dd < 0 ? Range{1, G} : Range{G + N + c + 1, N + 2 * G + c};
  int row = (dd < 0 ? G + 1 : G + N + c) - s.targetfirst;  // 0-based
  if (row < 0 || row >= s.ntarget()) fail("clamped row falls outside the tangential stencil; this is a bug");
  return remap_rows(s, r.lo, std::vector<int>(r.size(), row));
}

// src/schedule.jl:1332-1353 (reflecting: ghost j holds mirror point j*)
Stencil1D mirror_rows(const Stencil1D& s, int N, int G, int c, int dd) {
  Range r = dd < 0 ? Range{1, G} : Range{G + N + 1 + c, N + 2 * G + c};
  std::vector<int> rows;
  for (int j = r.lo; j <= r.hi; ++j) {
    int js = dd < 0 ? 2 * G + 1 + c - j : 2 * (G + N) + 1 + c - j;
    int row = js - s.targetfirst;
    if (row < 0 || row >= s.ntarget()) fail("mirror rows fall outside the tangential stencil; this is a bug");
    rows.push_back(row);
  }
  return remap_rows(s, r.lo, rows);
}

// src/schedule.jl:1387-1395 (vertex-like wall dimension)
Stencil1D mirror_restriction(int N, int G, int od, int p) {
  int h = N / 2;
  Range r = od == 0 ? Range{G + 1, G + 1 + h} : Range{G + 1 + h, G + N + 1};
  return restrict_stencil_over(r, N, G, 1, 0, od, p);
}

Stencil1D unit_point() {
  Stencil1D s = make(1, 1, 1);
  s.srcstart[0] = 1;
  s.weights[0] = 1.0;
  s.finish();
  return s;
}

using Box = std::array<Range, kMaxDim>;
bool overlaps(const Box& a, const Box& b, int D) {
  for (int d = 0; d < D; ++d)
    if (std::max(a[d].lo, b[d].lo) > std::min(a[d].hi, b[d].hi)) return false;
  return true;
}

}  // namespace

void Stencil1D::finish() {
  int n = ntarget();
  unit = order == 1;
  for (double w : weights) unit = unit && w == 1.0;
  single.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    int nz = 0, row = 0;
    for (int m = 0; m < order; ++m)
      if (weights[size_t(order) * i + m] != 0.0) { ++nz; row = m + 1; }
    single[i] = nz == 1 ? row : 0;
  }
}

static int kindrank(int k) { return k; }

bool key_less(const GroupKey& a, const GroupKey& b) {
  if (a.kind != b.kind) return kindrank(a.kind) < kindrank(b.kind);
  if (a.direction != b.direction) return a.direction < b.direction;
  if (a.offset != b.offset) return a.offset < b.offset;
  if (a.level != b.level) return a.level < b.level;
  if (a.mirror != b.mirror) return a.mirror < b.mirror;
  if (a.orientation != b.orientation) return a.orientation < b.orientation;
  if (a.axes != b.axes) return a.axes < b.axes;
  return a.span < b.span;
}

std::vector<double> parity_factors(int D, const std::vector<std::array<int8_t, kMaxDim>>& par) {
  int nstates = 1;
  for (int d = 0; d < D; ++d) nstates *= 3;
  int nv = (int)par.size();
  std::vector<double> f(size_t(nv) * nstates, 1.0);
  for (int v = 0; v < nv; ++v)
    for (int col = 1; col <= nstates; ++col) {
      int x = 1, r = col - 1;
      for (int d = 0; d < D; ++d) {
        int state = r % 3;
        r /= 3;
        int sigma = par[v][d] == OddParity ? -1 : 1;
        x *= state == 0 ? 1 : state == 1 ? sigma : (1 + sigma) / 2;
      }
      f[size_t(v) + size_t(nv) * (col - 1)] = double(x);
    }
  return f;
}

namespace {

struct Builder {
  const Forest& F;
  int D, N;
  std::array<int, kMaxDim> G, c;
  std::map<GroupKey, std::pair<std::vector<int32_t>, std::vector<int32_t>>, KeyLess> pairs;
  std::vector<BoundaryRegion> boundaries;

  Builder(const Forest& f, const std::array<int, kMaxDim>& g, const std::array<int, kMaxDim>& cc)
      : F(f), D(f.D), N(f.N), G(g), c(cc) {}

  Dir childoffset(const Key& k) const {
    Dir o{0, 0, 0};
    for (int d = 0; d < D; ++d) o[d] = int(k.c[d] & 1u);
    return o;
  }

  void record(int kind, const Dir& dir, const Dir& offset, int lvl,
              const std::array<int8_t, kMaxDim>& mirror, int b, int s) {
    GroupKey key;
    key.kind = kind;
    key.direction = dir;
    key.offset = offset;
    key.level = lvl;
    key.mirror = mirror;
    auto& e = pairs[key];
    e.first.push_back(b);
    e.second.push_back(s);
  }

  void boundary(int b, const Dir& delta) {
    BoundaryRegion r;
    r.block = b;
    r.direction = delta;
    for (int d = 0; d < kMaxDim; ++d) {
      Range rg = d < D ? target_range(N, G[d], c[d], delta[d], 0, false) : Range{1, 1};
      r.lo[d] = rg.lo;
      r.hi[d] = rg.hi;
    }
    boundaries.push_back(r);
  }

  // src/schedule.jl:1525-1610, cell-like dimensions
  void block_sources(int b, const std::vector<Dir>& dirs) {
    const Key& k = F.leaves[b];
    const Dir zero{0, 0, 0};
    const std::array<int8_t, kMaxDim> nomirror{0, 0, 0};
    for (const Dir& delta : dirs) {
      bool empty = false;
      for (int d = 0; d < D; ++d)
        if (target_range(N, G[d], c[d], delta[d], 0, false).size() == 0) empty = true;
      if (empty) continue;
      std::array<int8_t, kMaxDim> walls;
      Dir dprime = F.wall_direction(k, delta, walls);
      bool anywall = false;
      for (int d = 0; d < D; ++d) anywall = anywall || walls[d] != OuterFace;
      if (anywall) {
        mirror_sources(b, delta, dprime, walls);
        continue;
      }
      auto nbrs = F.neighbors(k, delta);
      if (nbrs.empty()) {
        boundary(b, delta);
        continue;
      }
      int nblevel = F.leaves[nbrs[0]].level;
      if (nblevel == k.level) {
        record(KindCopy, delta, zero, 0, nomirror, b, nbrs[0]);
      } else if (nblevel < k.level) {
        record(KindProlong, delta, childoffset(k), k.level, nomirror, b, nbrs[0]);
      } else {
        for (int nb : nbrs)
          record(KindRestrict, delta, childoffset(F.leaves[nb]), 0, nomirror, b, nb);
      }
    }
  }

  // src/schedule.jl:2237-2297
  void mirror_sources(int b, const Dir& delta, const Dir& dprime,
                      const std::array<int8_t, kMaxDim>& walls) {
    const Key& k = F.leaves[b];
    const Dir zero{0, 0, 0};
    std::array<int8_t, kMaxDim> state{0, 0, 0};
    for (int d = 0; d < D; ++d)
      state[d] = walls[d] == OuterFace ? 0 : walls[d] == OutflowFace ? 3 : 1;
    if (dprime == zero) {
      record(KindCopy, delta, zero, 0, state, b, b);
      return;
    }
    auto nbrs = F.neighbors(k, dprime);
    if (nbrs.empty()) {
      boundary(b, delta);
      return;
    }
    int nblevel = F.leaves[nbrs[0]].level;
    if (nblevel == k.level) {
      record(KindCopy, delta, zero, 0, state, b, nbrs[0]);
    } else if (nblevel < k.level) {
      record(KindProlong, delta, childoffset(k), k.level, state, b, nbrs[0]);
    } else {
      for (int nb : nbrs) {
        Dir o = childoffset(F.leaves[nb]);
        bool ok = true;
        for (int d = 0; d < D; ++d)
          if (state[d] != 0 && o[d] != (delta[d] < 0 ? 0 : 1)) ok = false;
        if (ok) record(KindRestrict, delta, o, 0, state, b, nb);
      }
    }
  }
};

}  // namespace

GhostSchedule::GhostSchedule(const Forest& f, const Operators& ops_, const int* G_, const int* c_,
                             const std::vector<std::array<int8_t, kMaxDim>>& par)
    : forest(&f), generation(f.generation), D(f.D), N(f.N), ops(ops_), parities(par) {
  for (int d = 0; d < kMaxDim; ++d) {
    G[d] = d < D ? G_[d] : 0;
    c[d] = d < D ? c_[d] : 0;
    if (d < D && G[d] < 0) fail("G must be nonnegative");
    if (d < D && c[d] != 0 && c[d] != 1) fail("centering must be cell (0) or vertex (1) per dimension");
    // src/storage.jl:865-872
    if (d < D && N < 2 * G[d])
      fail("N must be >= 2G in every dimension, got N=" + std::to_string(N) + ", G=" + std::to_string(G[d]));
    stored[d] = d < D ? N + 2 * G[d] + c[d] : 1;
  }
  check_operators(N, G.data(), c.data(), D, ops);
  bool closed = false;
  for (int d = 0; d < D; ++d) closed = closed || c[d] == 1;
  if (closed)
    fail("vertex-like dimensions (closed blocks, stage 0 and stage P) are not implemented yet",
         kUnsupported);
  // A variable needs a parity in every dimension that has a reflecting face
  // (src/storage.jl:512-526).
  bool reflecting = false;
  for (int d = 0; d < D; ++d)
    for (int s = 0; s < 2; ++s) reflecting = reflecting || f.faces[d][s] == ReflectingFace;
  for (size_t v = 0; v < par.size(); ++v)
    for (int d = 0; d < D; ++d)
      if ((f.faces[d][0] == ReflectingFace || f.faces[d][1] == ReflectingFace) && par[v][d] == NoParity)
        fail("variable " + std::to_string(v + 1) + " has NoParity in dimension " + std::to_string(d + 1) +
             ", which has a reflecting face");

  Builder B(f, G, c);
  auto dirs = alldirections(D);
  for (int b = 0; b < f.nleaves(); ++b) B.block_sources(b, dirs);
  boundaries = std::move(B.boundaries);

  // One group per key (src/schedule.jl:3193-3232, 2970-2992).
  for (auto& kv : B.pairs) {
    const GroupKey& key = kv.first;
    TransferGroup g;
    g.key = key;
    g.kind = key.kind;
    g.targetblocks = kv.second.first;
    g.sourceblocks = kv.second.second;
    for (int d = 0; d < kMaxDim; ++d) {
      if (d >= D) {
        g.stencils[d] = unit_point();
        continue;
      }
      int dd = key.direction[d], od = key.offset[d], st = key.mirror[d];
      int ds = st == 0 ? dd : 0;
      Stencil1D s = (st != 0 && key.kind == KindRestrict && c[d] == 1)
                        ? mirror_restriction(N, G[d], od, ops.restriction)
                        : kind_stencil(N, G[d], c[d], key.kind, ds, od, ops, key.span[d]);
      if (st == 3) s = clamp_rows(s, N, G[d], c[d], dd);
      else if (st != 0) s = mirror_rows(s, N, G[d], c[d], dd);
      g.stencils[d] = std::move(s);
    }
    // src/schedule.jl:3218-3222: clamped dimensions carry no sign.
    bool signed_ = false;
    int col = 1, p3 = 1;
    for (int d = 0; d < D; ++d) {
      int s = key.mirror[d] == 3 ? 0 : key.mirror[d];
      signed_ = signed_ || s != 0;
      col += s * p3;
      p3 *= 3;
    }
    g.factorcol = signed_ ? col : 0;
    // src/schedule.jl:653-656
    g.stage = key.isunify(D) ? kUnifyTag
              : key.kind == KindProlong ? prolongation_tag(key.level) : kPhase1Tag;
    groups.push_back(std::move(g));
  }
  std::stable_sort(groups.begin(), groups.end(), [](const TransferGroup& a, const TransferGroup& b) {
    if (a.stage != b.stage) return a.stage < b.stage;
    return key_less(a.key, b.key);
  });

  // Stages in fill order; phase 1 always present (src/schedule.jl:3259-3266).
  std::vector<int> tags;
  if (closed) tags.push_back(kUnifyTag);
  tags.push_back(kPhase1Tag);
  for (auto& g : groups)
    if (g.stage >= 2 && (tags.empty() || tags.back() != g.stage)) tags.push_back(g.stage);
  for (int t : tags) stages.push_back(Stage{t, {}});
  for (int i = 0; i < (int)groups.size(); ++i)
    for (auto& st : stages)
      if (st.tag == groups[i].stage) st.groups.push_back(i);

  // src/schedule.jl:2876-2931: which prolongation stages run together.
  int p1 = 0;
  for (int i = 0; i < (int)tags.size(); ++i)
    if (tags[i] == kPhase1Tag) p1 = i;
  for (int i = 0; i <= p1; ++i) stagegroups.push_back({i, i});
  if ((int)tags.size() > p1 + 1) {
    struct Access {
      std::vector<std::pair<int, Box>> writes, reads;
    };
    std::vector<Access> access(tags.size());
    for (int si = p1 + 1; si < (int)tags.size(); ++si)
      for (int gi : stages[si].groups) {
        const TransferGroup& g = groups[gi];
        Box wb, rb;
        bool empty = false;
        for (int d = 0; d < kMaxDim; ++d) {
          const Stencil1D& s = g.stencils[d];
          wb[d] = Range{s.targetfirst, s.targetfirst + s.ntarget() - 1};
          if (wb[d].size() == 0) empty = true;
          int mn = *std::min_element(s.srcstart.begin(), s.srcstart.end());
          int mx = *std::max_element(s.srcstart.begin(), s.srcstart.end());
          rb[d] = Range{mn, mx + s.order - 1};
        }
        if (empty) continue;
        for (size_t t = 0; t < g.targetblocks.size(); ++t) {
          access[si].reads.push_back({g.sourceblocks[t], rb});
          access[si].writes.push_back({g.targetblocks[t], wb});
        }
      }
    std::unordered_map<int, std::vector<Box>> written, read;
    auto meets = [&](std::unordered_map<int, std::vector<Box>>& boxes,
                     const std::vector<std::pair<int, Box>>& list) {
      for (auto& e : list) {
        auto it = boxes.find(e.first);
        if (it == boxes.end()) continue;
        for (auto& w : it->second)
          if (overlaps(w, e.second, D)) return true;
      }
      return false;
    };
    for (int i = p1 + 1; i < (int)tags.size(); ++i) {
      Access& a = access[i];
      bool joins = i > p1 + 1 && tags[i] == tags[i - 1] + 1 && !meets(written, a.reads) &&
                   !meets(written, a.writes) && !meets(read, a.writes);
      if (joins) {
        stagegroups.back()[1] = i;
      } else {
        stagegroups.push_back({i, i});
        written.clear();
        read.clear();
      }
      for (auto& e : a.writes) written[e.first].push_back(e.second);
      for (auto& e : a.reads) read[e.first].push_back(e.second);
    }
  }
}

}  // namespace tah
