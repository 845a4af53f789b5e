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
              const std::array<int8_t, kMaxDim>& mirror, int b, int s,
              const std::array<int8_t, kMaxDim>& span = {-1, -1, -1}) {
    GroupKey key;
    key.kind = kind;
    key.direction = dir;
    key.offset = offset;
    key.level = lvl;
    key.mirror = mirror;
    key.span = span;
    auto& e = pairs[key];
    e.first.push_back(b);
    e.second.push_back(s);
  }

  bool closed() const {
    for (int d = 0; d < D; ++d)
      if (c[d] == 1) return true;
    return false;
  }

  // A coarser neighbour's ghosts, recorded whole or split by
  // ghost_plane_parts into prolongation spans and copies.
  void record_prolong(int b, const Dir& delta, const Dir& look, int src,
                      const std::array<int8_t, kMaxDim>& mirror) {
    const Key& k = F.leaves[b];
    const Dir zero{0, 0, 0};
    std::vector<PlanePart> parts;
    if (!closed() || !ghost_plane_parts(b, delta, look, src, parts)) {
      record(KindProlong, delta, childoffset(k), k.level, mirror, b, src);
      return;
    }
    for (const PlanePart& pp : parts) {
      if (pp.kind == KindProlong)
        record(KindProlong, delta, childoffset(k), k.level, mirror, b, pp.source, pp.span);
      else
        record(KindCopy, pp.dir, zero, 0, mirror, b, pp.source, pp.span);
    }
  }

  // --- Closed blocks: ghost-plane parts, stage 0 and stage P ----------

  struct PlanePart {
    int kind;
    Dir dir;
    std::array<int8_t, kMaxDim> span;
    int source;
  };

  // src/schedule.jl:1434-1507. The ghosts of leaf b in direction delta,
  // whose neighbour there (global leaf `coarse`) is coarser: along each
  // tangential vertex-like dimension the face-plane rows that a block of
  // b's level holds (found across `look` plus steps along the plane
  // dimensions) are copies from it; the rest stays a prolongation. False
  // where nothing splits.
  bool ghost_plane_parts(int b, const Dir& delta, const Dir& look, int coarse,
                         std::vector<PlanePart>& parts) const {
    const Key& k = F.leaves[b];
    std::vector<int> tdims;
    for (int d = 0; d < D; ++d)
      if (delta[d] == 0 && c[d] == 1) tdims.push_back(d);
    if (tdims.empty()) return false;
    const int l = k.level;
    std::map<Dir, std::pair<int, int>> found;  // dir -> (level, leaf); level -1: none
    auto holder = [&](const Dir& dir) {
      auto it = found.find(dir);
      if (it != found.end()) return it->second;
      std::pair<int, int> h{-1, -1};
      auto nb = F.neighbors(k, dir);
      if (!nb.empty()) {
        int lv = F.leaves[nb[0]].level;
        if (lv > l)
          fail("a ghost face-plane row is held by a finer block than the block itself next to a "
               "coarser neighbour; 2:1 balance forbids that, so this is a bug");
        h = {lv, nb[0]};
      }
      found[dir] = h;
      return h;
    };
    const int nt = int(tdims.size());
    int npieces = 1;
    for (int i = 0; i < nt; ++i) npieces *= 3;
    std::vector<std::array<int, kMaxDim>> pieces(npieces);
    for (int q = 0; q < npieces; ++q) {
      int r = q;
      for (int i = 0; i < nt; ++i) {  // the first plane dimension fastest
        pieces[q][i] = r % 3;
        r /= 3;
      }
    }
    struct Src {
      bool any = false;
      Dir dir;
      int leaf = -1;
    };
    std::vector<Src> sources(npieces);
    bool anysrc = false;
    for (int q = 0; q < npieces; ++q) {
      std::vector<std::pair<int, int>> plane;  // (dimension, step)
      for (int i = 0; i < nt; ++i)
        if (pieces[q][i] != 1) plane.push_back({tdims[i], pieces[q][i] == 0 ? -1 : 1});
      const int np = int(plane.size());
      std::vector<int> subsets;
      for (int m = 1; m < (1 << np); ++m) subsets.push_back(m);
      std::stable_sort(subsets.begin(), subsets.end(), [](int x, int y) {
        int cx = __builtin_popcount(unsigned(x)), cy = __builtin_popcount(unsigned(y));
        return cx != cy ? cx < cy : x < y;
      });
      int bestm = -1, bestleaf = -1;
      for (int m : subsets) {
        Dir dir = look;
        for (int i = 0; i < np; ++i)
          if ((m >> i) & 1) dir[plane[i].first] = plane[i].second;
        auto h = holder(dir);
        if (h.first < 0) continue;
        if (h.second == coarse) continue;
        if (h.first == l && bestm < 0) {
          bestm = m;
          bestleaf = h.second;
        }
      }
      if (bestm < 0) continue;
      Dir dir = delta;
      for (int i = 0; i < np; ++i)
        if ((bestm >> i) & 1) dir[plane[i].first] = plane[i].second;
      sources[q] = Src{true, dir, bestleaf};
      anysrc = true;
    }
    if (!anysrc) return false;
    auto code = [](int cls) { return int8_t(cls == 0 ? 0 : cls == 1 ? 8 * 1 + 3 : 8 * 4 + 4); };
    auto spanof = [&](const std::array<int, kMaxDim>& cls) {
      std::array<int8_t, kMaxDim> sp{-1, -1, -1};
      for (int i = 0; i < nt; ++i) sp[tdims[i]] = code(cls[i]);
      return sp;
    };
    std::vector<int> rest;
    for (int q = 0; q < npieces; ++q)
      if (!sources[q].any) rest.push_back(q);
    if (!rest.empty()) {
      std::array<int, kMaxDim> lo{9, 9, 9}, hi{-1, -1, -1};
      for (int q : rest)
        for (int i = 0; i < nt; ++i) {
          lo[i] = std::min(lo[i], pieces[q][i]);
          hi[i] = std::max(hi[i], pieces[q][i]);
        }
      int box = 1;
      for (int i = 0; i < nt; ++i) box *= hi[i] - lo[i] + 1;
      if (box == int(rest.size())) {
        auto first_sub = [](int x) { return x == 0 ? 0 : x == 1 ? 1 : 4; };
        auto last_sub = [](int x) { return x == 0 ? 0 : x == 1 ? 3 : 4; };
        std::array<int8_t, kMaxDim> sp{-1, -1, -1};
        for (int i = 0; i < nt; ++i) sp[tdims[i]] = int8_t(8 * first_sub(lo[i]) + last_sub(hi[i]));
        parts.push_back(PlanePart{KindProlong, delta, sp, coarse});
      } else {
        for (int q : rest) parts.push_back(PlanePart{KindProlong, delta, spanof(pieces[q]), coarse});
      }
    }
    for (int q = 0; q < npieces; ++q)
      if (sources[q].any) parts.push_back(PlanePart{KindCopy, sources[q].dir, spanof(pieces[q]), sources[q].leaf});
    return true;
  }

  // The neighbours of one leaf across each direction, looked up once
  // (NeighbourTable, src/schedule.jl:1743-1788): level (-1: none to copy
  // from) and each neighbour with its child offset.
  struct Table {
    struct Entry {
      bool known = false;
      int level = -1;
      std::vector<std::pair<int, Dir>> nbrs;
    };
    const Forest& F;
    int j, level;
    std::vector<Entry> e;
    Table(const Forest& f, int jj) : F(f), j(jj), level(f.leaves[jj].level), e(27) {}
    const Entry& at(const Dir& delta) {
      int i = 0, p = 1;
      for (int d = 0; d < F.D; ++d) {
        i += (delta[d] + 1) * p;
        p *= 3;
      }
      Entry& x = e[i];
      if (!x.known) {
        x.known = true;
        auto nb = F.neighbors(F.leaves[j], delta);
        if (!nb.empty()) {
          x.level = F.leaves[nb[0]].level;
          for (int leaf : nb) {
            Dir o{0, 0, 0};
            for (int d = 0; d < F.D; ++d) o[d] = int(F.leaves[leaf].c[d] & 1u);
            x.nbrs.push_back({leaf, o});
          }
        }
      }
      return x;
    }
  };

  struct PieceCopy {
    int level, leaf;
    Dir cls, delta, offset;
  };

  static int piece_class(int sub) { return sub == 0 ? 0 : sub == 4 ? 2 : 1; }
  Dir piece_of(const Dir& sub) const {
    Dir p{1, 1, 1};
    for (int d = 0; d < D; ++d) p[d] = piece_class(sub[d]);
    return p;
  }
  int piece_index(const Dir& cls) const {
    int i = 0, p = 1;
    for (int d = 0; d < D; ++d) {
      i += cls[d] * p;
      p *= 3;
    }
    return i;
  }
  // The sub-classes each dimension has (closed_subpieces, src/schedule.jl:1687-1703).
  std::vector<int> allowed(int d) const {
    if (c[d] == 0) return {1, 3};
    if (N == 2) return {0, 2, 4};
    return {0, 1, 2, 3, 4};
  }
  // Every sub-piece (or piece), dimension 1 fastest.
  std::vector<Dir> product(const std::array<std::vector<int>, kMaxDim>& vals) const {
    std::vector<Dir> out;
    int n = 1;
    for (int d = 0; d < D; ++d) n *= int(vals[d].size());
    for (int q = 0; q < n; ++q) {
      Dir x{0, 0, 0};
      int r = q;
      for (int d = 0; d < D; ++d) {
        x[d] = vals[d][r % vals[d].size()];
        r /= int(vals[d].size());
      }
      out.push_back(x);
    }
    return out;
  }
  std::vector<Dir> subpieces() const {
    std::array<std::vector<int>, kMaxDim> v;
    for (int d = 0; d < D; ++d) v[d] = allowed(d);
    return product(v);
  }
  std::vector<Dir> pieces() const {
    std::array<std::vector<int>, kMaxDim> v;
    for (int d = 0; d < D; ++d) v[d] = c[d] == 1 ? std::vector<int>{0, 1, 2} : std::vector<int>{1};
    return product(v);
  }
  Dir representative(const Dir& cls) const {
    Dir r{0, 0, 0};
    for (int d = 0; d < D; ++d) r[d] = c[d] == 0 ? 1 : 2 * cls[d];
    return r;
  }

  // Every copy of sub-piece `sub` of t's leaf on its level or a finer one,
  // its own first (piece_copies, src/schedule.jl:1802-1847).
  std::vector<PieceCopy> piece_copies(Table& t, const Dir& sub) const {
    const Dir zero{0, 0, 0};
    Dir cls = piece_of(sub);
    std::vector<PieceCopy> copies{{t.level, t.j, cls, zero, zero}};
    Dir side{0, 0, 0};
    for (int d = 0; d < D; ++d) side[d] = c[d] == 1 && cls[d] != 1 ? (cls[d] == 0 ? -1 : 1) : 0;
    for (int m = 1; m < (1 << D); ++m) {
      bool skip = false;
      for (int d = 0; d < D; ++d)
        if (((m >> d) & 1) && side[d] == 0) skip = true;
      if (skip) continue;
      Dir delta{0, 0, 0};
      for (int d = 0; d < D; ++d) delta[d] = ((m >> d) & 1) ? side[d] : 0;
      const Table::Entry& e = t.at(delta);
      if (e.level < 0) continue;
      Dir across{1, 1, 1};
      for (int d = 0; d < D; ++d) across[d] = delta[d] == 0 ? cls[d] : 2 - cls[d];
      if (e.level < t.level) continue;
      if (e.level == t.level) {
        copies.push_back({e.level, e.nbrs[0].first, across, delta, zero});
        continue;
      }
      for (const auto& no : e.nbrs) {
        const Dir& o = no.second;
        bool holds = true;
        for (int d = 0; d < D; ++d)
          if (delta[d] == 0)
            holds = holds && (c[d] == 0 ? o[d] == (sub[d] == 3 ? 1 : 0)
                                        : (sub[d] == 2 || o[d] == (sub[d] >= 3 ? 1 : 0)));
        if (!holds) continue;
        Dir fcls{1, 1, 1};
        for (int d = 0; d < D; ++d)
          fcls[d] = delta[d] != 0 ? across[d]
                    : c[d] == 0   ? 1
                    : sub[d] == 2 ? (o[d] == 0 ? 2 : 0)
                    : (sub[d] == 0 || sub[d] == 4) ? cls[d]
                                                   : 1;
        copies.push_back({e.level, no.first, fcls, delta, o});
      }
    }
    return copies;
  }

  // The primary among the copies (piece_primary, src/schedule.jl:1921):
  // the finest level, then the greatest class tuple, dimension 1 first.
  PieceCopy primary(const std::vector<PieceCopy>& copies) const {
    size_t best = 0;
    for (size_t i = 1; i < copies.size(); ++i) {
      const PieceCopy &a = copies[i], &b = copies[best];
      bool better = false;
      if (a.level != b.level) {
        better = a.level > b.level;
      } else {
        for (int d = 0; d < D; ++d)
          if (a.cls[d] != b.cls[d]) {
            better = a.cls[d] > b.cls[d];
            break;
          }
      }
      if (better) best = i;
    }
    return copies[best];
  }

  // The span codes of the bounding box of `subs` where they fill it
  // (box_spans, src/schedule.jl:2125-2131).
  bool box_spans(const std::vector<Dir>& subs, std::array<int8_t, kMaxDim>& spans) const {
    Dir lo{9, 9, 9}, hi{-1, -1, -1};
    for (const Dir& p : subs)
      for (int d = 0; d < D; ++d) {
        lo[d] = std::min(lo[d], p[d]);
        hi[d] = std::max(hi[d], p[d]);
      }
    long box = 1;
    for (int d = 0; d < D; ++d) {
      long n = 0;
      for (int x : allowed(d)) n += lo[d] <= x && x <= hi[d];
      box *= n;
    }
    if (box != long(subs.size())) return false;
    spans = {-1, -1, -1};
    for (int d = 0; d < D; ++d) spans[d] = int8_t(8 * lo[d] + hi[d]);
    return true;
  }

  // One sub-piece of leaf b and the copy it takes its value from.
  struct SubEntry {
    Dir sub;
    int kind, source;
    Dir delta, offset;
  };

  // merged_transfers (src/schedule.jl:2077-2122): b's entries by (kind,
  // source, direction, offset), each set recorded as one box of spans
  // where it fills one, else per piece, else per sub-piece.
  void record_merged(int b, const std::vector<SubEntry>& entries) {
    std::map<std::tuple<int, int, Dir, Dir>, std::vector<Dir>> bysource;
    for (const SubEntry& e : entries) {
      auto& v = bysource[std::make_tuple(e.kind, e.source, e.delta, e.offset)];
      if (std::find(v.begin(), v.end(), e.sub) == v.end()) v.push_back(e.sub);
    }
    const std::array<int8_t, kMaxDim> nomirror{0, 0, 0};
    for (const auto& kv : bysource) {
      const int kind = std::get<0>(kv.first), s = std::get<1>(kv.first);
      const Dir &delta = std::get<2>(kv.first), &offset = std::get<3>(kv.first);
      const std::vector<Dir>& subs = kv.second;
      std::array<int8_t, kMaxDim> spans;
      if (kind == KindCopy && box_spans(subs, spans)) {
        record(kind, delta, offset, 0, nomirror, b, s, spans);
        continue;
      }
      std::vector<Dir> ps;
      for (const Dir& x : subs) ps.push_back(piece_of(x));
      std::sort(ps.begin(), ps.end());
      ps.erase(std::unique(ps.begin(), ps.end()), ps.end());
      for (const Dir& piece : ps) {
        std::vector<Dir> part;
        for (const Dir& x : subs)
          if (piece_of(x) == piece) part.push_back(x);
        if (box_spans(part, spans)) {
          record(kind, delta, offset, 0, nomirror, b, s, spans);
        } else {
          for (const Dir& x : part) {
            std::array<int8_t, kMaxDim> one{-1, -1, -1};
            for (int d = 0; d < D; ++d) one[d] = int8_t(8 * x[d] + x[d]);
            record(kind, delta, offset, 0, nomirror, b, s, one);
          }
        }
      }
    }
  }

  // The kept-over-created pass into created leaf j (regrid_kept_pairs,
  // src/regrid.jl:826-877): every sub-piece of j's closed box off its open
  // interior that j holds at the finest level holding it takes the value
  // of the primary among the kept blocks holding it there. kept[x]: leaf
  // x was a leaf before the regrid.
  void kept_sources(int j, const std::vector<char>& kept) {
    Table t(F, j);
    // Most created blocks have no kept neighbour on their level across a
    // face, edge or corner of a vertex-like dimension: skip them whole.
    bool any = false;
    int n = 1;
    for (int d = 0; d < D; ++d) n *= 3;
    for (int q = 0; q < n && !any; ++q) {
      Dir delta{0, 0, 0};
      bool zero = true, ok = true;
      for (int d = 0, r = q; d < D; ++d, r /= 3) {
        delta[d] = r % 3 - 1;
        zero = zero && delta[d] == 0;
        ok = ok && (c[d] == 1 || delta[d] == 0);
      }
      if (zero || !ok) continue;
      const Table::Entry& e = t.at(delta);
      if (e.level != t.level) continue;
      for (const auto& nb : e.nbrs) any = any || kept[nb.first];
    }
    if (!any) return;
    std::vector<SubEntry> entries;
    for (const Dir& sub : subpieces()) {
      Dir cls = piece_of(sub);
      bool interior = true;
      for (int d = 0; d < D; ++d) interior = interior && cls[d] == 1;
      if (interior) continue;
      auto copies = piece_copies(t, sub);
      int finest = -1;
      for (const auto& x : copies) finest = std::max(finest, x.level);
      if (t.level != finest) continue;
      std::vector<PieceCopy> holders;
      for (const auto& x : copies)
        if (x.level == finest && kept[x.leaf]) holders.push_back(x);
      if (holders.empty()) continue;
      PieceCopy p = primary(holders);
      entries.push_back({sub, KindCopy, p.leaf, p.delta, p.offset});
    }
    record_merged(j, entries);
  }

  // Stage 0's transfers into leaf b (unify_transfers and merged_transfers,
  // src/schedule.jl:1990-2021, 2077-2122), recorded; returns b's primary
  // bits, one per piece of its closed box.
  uint32_t unify_sources(int b) {
    Table t(F, b);
    const int npieces = [&] {
      int n = 1;
      for (int d = 0; d < D; ++d) n *= 3;
      return n;
    }();
    std::vector<int> shared(npieces, 0);
    std::vector<PieceCopy> prim(npieces);
    uint32_t bits = 0;
    std::vector<SubEntry> entries;
    for (const Dir& sub : subpieces()) {
      Dir cls = piece_of(sub);
      int i = piece_index(cls);
      if (shared[i] == 0) {
        auto copies = piece_copies(t, representative(cls));
        prim[i] = primary(copies);
        bool onelevel = true;
        for (const auto& x : copies) onelevel = onelevel && x.level == t.level;
        shared[i] = onelevel ? 1 : 2;
        if (prim[i].leaf == b && prim[i].cls == cls) bits |= 1u << i;
      }
      bool interior = true;
      for (int d = 0; d < D; ++d) interior = interior && cls[d] == 1;
      if (interior) continue;
      PieceCopy p = shared[i] == 1 ? prim[i] : primary(piece_copies(t, sub));
      if (p.leaf == b && p.cls == cls) continue;
      entries.push_back({sub, p.level == t.level ? KindCopy : KindRestrict, p.leaf, p.delta, p.offset});
    }
    record_merged(b, entries);
    return bits;
  }

  // Stage P of leaf b (selfimage_sources!, src/schedule.jl:2606-2636):
  // every primary piece on a reflecting wall plane, projected in place.
  void selfimage_sources(int b, uint32_t bits, const std::vector<std::array<int8_t, kMaxDim>>& parities,
                         std::map<GroupKey, std::pair<std::vector<int32_t>, std::vector<int32_t>>, KeyLess>& ppairs) const {
    const Key& k = F.leaves[b];
    std::array<std::array<bool, 2>, kMaxDim> refl{};
    bool any = false;
    for (int d = 0; d < D; ++d)
      for (int s = 0; s < 2; ++s) {
        refl[d][s] = false;
        if (c[d] != 1) continue;
        Dir delta{0, 0, 0};
        delta[d] = s == 0 ? -1 : 1;
        std::array<int8_t, kMaxDim> walls;
        F.wall_direction(k, delta, walls);
        refl[d][s] = walls[d] == ReflectingFace;
        any = any || refl[d][s];
      }
    if (!any) return;
    std::map<std::array<int8_t, kMaxDim>, std::vector<Dir>> bystate;
    for (const Dir& cls : pieces()) {
      if (!((bits >> piece_index(cls)) & 1u)) continue;
      std::array<int8_t, kMaxDim> state{0, 0, 0};
      bool anystate = false;
      for (int d = 0; d < D; ++d) {
        state[d] = (cls[d] == 0 && refl[d][0]) || (cls[d] == 2 && refl[d][1]) ? 2 : 0;
        anystate = anystate || state[d] != 0;
      }
      if (!anystate) continue;
      if (!parities.empty()) {
        bool odd = false;
        for (const auto& p : parities)
          for (int d = 0; d < D; ++d) odd = odd || (state[d] == 2 && p[d] == OddParity);
        if (!odd) continue;
      }
      bystate[state].push_back(cls);
    }
    const Dir zero{0, 0, 0};
    auto first_sub = [](int x) { return x == 0 ? 0 : x == 1 ? 1 : 4; };
    auto last_sub = [](int x) { return x == 0 ? 0 : x == 1 ? 3 : 4; };
    for (const auto& kv : bystate) {
      std::vector<Dir> ps = kv.second;
      std::vector<Dir> uniq;
      for (const Dir& x : ps)
        if (std::find(uniq.begin(), uniq.end(), x) == uniq.end()) uniq.push_back(x);
      Dir lo{9, 9, 9}, hi{-1, -1, -1};
      for (const Dir& x : uniq)
        for (int d = 0; d < D; ++d) {
          lo[d] = std::min(lo[d], x[d]);
          hi[d] = std::max(hi[d], x[d]);
        }
      long box = 1;
      for (int d = 0; d < D; ++d) box *= hi[d] - lo[d] + 1;
      std::vector<std::array<int8_t, kMaxDim>> spanlist;
      if (box == long(uniq.size())) {
        std::array<int8_t, kMaxDim> sp{-1, -1, -1};
        for (int d = 0; d < D; ++d) sp[d] = int8_t(8 * first_sub(lo[d]) + last_sub(hi[d]));
        spanlist.push_back(sp);
      } else {
        for (const Dir& x : uniq) {
          std::array<int8_t, kMaxDim> sp{-1, -1, -1};
          for (int d = 0; d < D; ++d) sp[d] = int8_t(8 * first_sub(x[d]) + last_sub(x[d]));
          spanlist.push_back(sp);
        }
      }
      for (const auto& sp : spanlist) {
        GroupKey key;
        key.kind = KindCopy;
        key.direction = zero;
        key.offset = zero;
        key.level = 0;
        key.mirror = kv.first;
        key.span = sp;
        auto& e = ppairs[key];
        e.first.push_back(b);
        e.second.push_back(b);
      }
    }
  }

  // A hook region with its relative-hook fields (src/schedule.jl:2188-2217).
  // A component of delta leaves through an outer face when the single step
  // along it finds no block and no reflecting or outflow face. The rest of
  // the direction, mirrored at walls, names the block's own ghosts the
  // nearest evolved point lies in: filled at phase 1, or with the
  // prolongations onto the block's level where they come from a coarser
  // block; where it finds no block at all the point is clamped along every
  // component of delta instead.
  void boundary(int b, const Dir& delta) {
    const Key& k = F.leaves[b];
    BoundaryRegion r;
    r.block = b;
    r.direction = delta;
    for (int d = 0; d < kMaxDim; ++d) {
      Range rg = d < D ? target_range(N, G[d], c[d], delta[d], 0, false) : Range{1, 1};
      r.lo[d] = rg.lo;
      r.hi[d] = rg.hi;
    }
    std::array<bool, kMaxDim> exits{false, false, false};
    std::array<int8_t, kMaxDim> walls;
    for (int d = 0; d < D; ++d) {
      if (delta[d] == 0) continue;
      Dir step{0, 0, 0};
      step[d] = delta[d];
      F.wall_direction(k, step, walls);
      exits[d] = walls[d] == OuterFace && F.neighbors(k, step).empty();
    }
    Dir rest{0, 0, 0};
    bool anyrest = false;
    for (int d = 0; d < D; ++d) {
      rest[d] = exits[d] ? 0 : delta[d];
      anyrest = anyrest || rest[d] != 0;
    }
    if (anyrest) {
      const Dir mirrored = F.wall_direction(k, rest, walls);
      bool anym = false;
      for (int d = 0; d < D; ++d) anym = anym || mirrored[d] != 0;
      if (anym) {
        const std::vector<int> nbrs = F.neighbors(k, mirrored);
        if (nbrs.empty()) {
          for (int d = 0; d < D; ++d) exits[d] = delta[d] != 0;
        } else if (F.leaves[nbrs[0]].level < k.level) {
          r.slot = prolongation_tag(k.level);
        }
      }
    }
    for (int d = 0; d < D; ++d)
      r.clamps[d] = !exits[d] ? 0 : delta[d] > 0 ? G[d] + N + c[d] : G[d] + 1;
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
        record_prolong(b, delta, delta, nbrs[0], nomirror);
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
      record_prolong(b, delta, dprime, nbrs[0], state);
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

void GhostSchedule::check_layout(const int* G_, const int* c_, bool operators) {
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
  if (operators) check_operators(N, G.data(), c.data(), D, ops);
}

// One group per key (src/schedule.jl:3193-3232, 2970-2992).
void GhostSchedule::add_group(const GroupKey& key, const std::vector<int32_t>& tb, const std::vector<int32_t>& sb,
                              int stage) {
  TransferGroup g;
  g.key = key;
  g.kind = key.kind;
  g.targetblocks = tb;
  g.sourceblocks = sb;
  for (int d = 0; d < kMaxDim; ++d) {
    if (d >= D) {
      g.stencils[d] = unit_point();
      continue;
    }
    int dd = key.direction[d], od = key.offset[d], st = key.mirror[d];
    // A span of stage 0 or P: a copy or a restriction into the target's
    // closed box (unify_stencil, src/schedule.jl:973-981).
    if (key.isunify(D)) {
      g.stencils[d] = key.kind == KindCopy
                          ? copy_stencil(N, G[d], c[d], dd, key.span[d])
                          : restrict_stencil_over(piece_range(N, G[d], c[d], key.span[d]), N, G[d], c[d], dd,
                                                  od, ops.restriction);
      continue;
    }
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
  g.stage = stage;
  groups.push_back(std::move(g));
}

GhostSchedule::GhostSchedule(UnifyOnly, const Forest& f, const int* G_, const int* c_,
                             const std::vector<std::array<int8_t, kMaxDim>>& par)
    : forest(&f), generation(f.generation), D(f.D), N(f.N), ops(PointValue, 2, 2, false, 0), parities(par) {
  check_layout(G_, c_, false);
  bool closed = false;
  for (int d = 0; d < D; ++d) closed = closed || c[d] == 1;
  if (!closed)
    fail("a cell-centered field set has no shared point to unify: its values sit half a cell in from every "
         "face. Build it over the set that holds the face quantity, vertex-like in the face dimension.");
  bool reflecting = false;
  for (int d = 0; d < D; ++d) reflecting = reflecting || f.faces[d][0] == ReflectingFace || f.faces[d][1] == ReflectingFace;
  for (size_t v = 0; v < par.size(); ++v)
    for (int d = 0; d < D; ++d)
      if ((f.faces[d][0] == ReflectingFace || f.faces[d][1] == ReflectingFace) && par[v][d] == NoParity)
        fail("variable " + std::to_string(v + 1) + " has NoParity in dimension " + std::to_string(d + 1) +
             ", which has a reflecting face");
  // Stage 0 (unify_sources!) and stage P (selfimage_sources!) of every
  // block, as the fill builds them (src/unify.jl:105-160).
  Builder B(f, G, c);
  std::map<GroupKey, std::pair<std::vector<int32_t>, std::vector<int32_t>>, KeyLess> ppairs;
  for (int b = 0; b < f.nleaves(); ++b) {
    const uint32_t bits = B.unify_sources(b);
    B.selfimage_sources(b, bits, reflecting ? par : std::vector<std::array<int8_t, kMaxDim>>{}, ppairs);
  }
  for (auto& kv : B.pairs) add_group(kv.first, kv.second.first, kv.second.second, kUnifyTag);
  for (auto& kv : ppairs) add_group(kv.first, kv.second.first, kv.second.second, kStageP);
  std::stable_sort(groups.begin(), groups.end(), [](const TransferGroup& a, const TransferGroup& b) {
    if (a.stage != b.stage) return a.stage < b.stage;
    return key_less(a.key, b.key);
  });
  std::vector<int> tags;
  if (!ppairs.empty()) tags.push_back(kStageP);
  tags.push_back(kUnifyTag);
  for (int t : tags) stages.push_back(Stage{t, {}});
  for (int i = 0; i < (int)groups.size(); ++i)
    for (auto& st : stages)
      if (st.tag == groups[i].stage) st.groups.push_back(i);
  for (int i = 0; i < (int)tags.size(); ++i) stagegroups.push_back({i, i});
}

GhostSchedule::GhostSchedule(const RegridFrom& from, const Forest& f, const Operators& ops_, const int* G_,
                             const int* c_, const std::vector<std::array<int8_t, kMaxDim>>& par)
    : forest(&f), generation(f.generation), D(f.D), N(f.N), ops(ops_), parities(par) {
  check_layout(G_, c_);
  const std::vector<Key>& old = from.oldleaves;
  if (old.empty()) fail("a regrid needs the leaves before it");
  nsources = int32_t(old.size());
  auto less = [&](const Key& a, const Key& b) { return curve_less(a, b, D); };
  for (size_t i = 1; i < old.size(); ++i)
    if (!less(old[i - 1], old[i])) fail("the old leaves must be sorted along the curve, without repeats");
  auto find_old = [&](const Key& k) {
    auto it = std::lower_bound(old.begin(), old.end(), k, less);
    return it != old.end() && *it == k ? int(it - old.begin()) : -1;
  };
  // Where each new leaf's data come from (classify_regrid!,
  // src/regrid.jl:620-652): the old leaf itself, its parent, or each of its
  // children, batched by (kind, child offset) with direction zero.
  Builder B(f, G, c);
  const Dir zero{0, 0, 0};
  const std::array<int8_t, kMaxDim> nomirror{0, 0, 0};
  std::vector<char> kept(f.nleaves(), 0);
  for (int j = 0; j < f.nleaves(); ++j) {
    const Key& k = f.leaves[j];
    const int same = find_old(k);
    if (same >= 0) {
      B.record(KindCopy, zero, zero, 0, nomirror, j, same);
      kept[j] = 1;
      continue;
    }
    if (k.level > 0) {
      Key parent = k;
      parent.level -= 1;
      for (int d = 0; d < D; ++d) parent.c[d] = k.c[d] >> 1;
      const int p = find_old(parent);
      if (p >= 0) {
        B.record(KindProlong, zero, B.childoffset(k), 0, nomirror, j, p);
        continue;
      }
    }
    if (k.level >= kMaxLevel) fail("cannot rebuild a new leaf: it is neither an old leaf, a child of one, nor a parent");
    for (int o = 0; o < (1 << D); ++o) {
      Key ch = k;
      ch.level += 1;
      for (int d = 0; d < D; ++d) ch.c[d] = 2 * k.c[d] + ((o >> d) & 1u);
      const int x = find_old(ch);
      if (x < 0)
        fail("cannot rebuild a new leaf: neither it, its parent, nor its child was a leaf before the regrid. "
             "A single regrid may move a block by at most one level, which holds when the previous tree was "
             "2:1 balanced.");
      B.record(KindRestrict, zero, B.childoffset(ch), 0, nomirror, j, x);
    }
  }
  // Every target is the new block's own range: the ghost builders at
  // direction zero (regrid_stage, src/regrid.jl:755-803).
  for (auto& kv : B.pairs) add_group(kv.first, kv.second.first, kv.second.second, kRegridTag);
  std::vector<int> tags{kRegridTag};
  bool closed = false;
  for (int d = 0; d < D; ++d) closed = closed || c[d] == 1;
  if (closed) {
    // The kept blocks' copies of shared points over the created blocks'
    // (regrid_kept_pairs, regrid_kept_stage, src/regrid.jl:826-933).
    Builder K(f, G, c);
    for (int j = 0; j < f.nleaves(); ++j)
      if (!kept[j]) K.kept_sources(j, kept);
    for (auto& kv : K.pairs) add_group(kv.first, kv.second.first, kv.second.second, kRegridKeptTag);
    tags.push_back(kRegridKeptTag);
  }
  std::stable_sort(groups.begin(), groups.end(), [](const TransferGroup& a, const TransferGroup& b) {
    if (a.stage != b.stage) return a.stage < b.stage;
    return key_less(a.key, b.key);
  });
  for (int t : tags) stages.push_back(Stage{t, {}});
  for (int i = 0; i < (int)groups.size(); ++i)
    for (auto& st : stages)
      if (st.tag == groups[i].stage) st.groups.push_back(i);
  for (int i = 0; i < (int)tags.size(); ++i) stagegroups.push_back({i, i});
}

GhostSchedule::GhostSchedule(const Forest& f, const Operators& ops_, const int* G_, const int* c_,
                             const std::vector<std::array<int8_t, kMaxDim>>& par)
    : forest(&f), generation(f.generation), D(f.D), N(f.N), ops(ops_), parities(par) {
  check_layout(G_, c_);
  bool closed = false;
  for (int d = 0; d < D; ++d) closed = closed || c[d] == 1;
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
  // Stage P's transfers, apart from the others (src/schedule.jl:3268-3276).
  std::map<GroupKey, std::pair<std::vector<int32_t>, std::vector<int32_t>>, KeyLess> ppairs;
  for (int b = 0; b < f.nleaves(); ++b) {
    B.block_sources(b, dirs);
    // The secondaries of the block's closed box from their primaries, and
    // its primaries on reflecting wall planes (src/schedule.jl:1599-1607).
    if (closed) {
      uint32_t bits = B.unify_sources(b);
      bool reflecting_set = false;
      for (int d = 0; d < D; ++d)
        reflecting_set = reflecting_set || f.faces[d][0] == ReflectingFace || f.faces[d][1] == ReflectingFace;
      B.selfimage_sources(b, bits, reflecting_set ? par : std::vector<std::array<int8_t, kMaxDim>>{}, ppairs);
    }
  }
  boundaries = std::move(B.boundaries);

  for (auto& kv : B.pairs) {
    const GroupKey& key = kv.first;
    // src/schedule.jl:653-656
    int stage = key.isunify(D) ? kUnifyTag : key.kind == KindProlong ? prolongation_tag(key.level) : kPhase1Tag;
    add_group(key, kv.second.first, kv.second.second, stage);
  }
  for (auto& kv : ppairs) add_group(kv.first, kv.second.first, kv.second.second, kStageP);
  std::stable_sort(groups.begin(), groups.end(), [](const TransferGroup& a, const TransferGroup& b) {
    if (a.stage != b.stage) return a.stage < b.stage;
    return key_less(a.key, b.key);
  });

  // Stages in fill order; phase 1 always present (src/schedule.jl:3259-3266).
  std::vector<int> tags;
  if (!ppairs.empty()) tags.push_back(kStageP);
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
    // A hook region whose relative form runs after a prolongation stage
    // writes its box with that stage and reads its nearest evolved points
    // (src/schedule.jl:2907-2911, hook_read_box :2795-2797).
    for (const BoundaryRegion& r : boundaries) {
      int si = -1;
      for (int i = 0; i < (int)tags.size(); ++i)
        if (tags[i] == r.slot) si = i;
      if (si <= p1) continue;
      Box wb, rb;
      for (int d = 0; d < kMaxDim; ++d) {
        wb[d] = Range{r.lo[d], r.hi[d]};
        rb[d] = r.clamps[d] == 0 ? wb[d] : Range{r.clamps[d], r.clamps[d]};
      }
      access[si].writes.push_back({r.block, wb});
      access[si].reads.push_back({r.block, rb});
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
