// Forest, Morton keys and the neighbour search; see forest.hpp.
#include "forest.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace tah {

namespace {
inline uint64_t padded(uint32_t c, int lvl) { return uint64_t(c) << (kMaxLevel - lvl); }
// Whether the most significant set bit of y is strictly above that of x
// (src/morton.jl:106).
inline bool less_msb(uint64_t x, uint64_t y) { return x < y && x < (x ^ y); }
}  // namespace

bool curve_less(const Key& a, const Key& b, int D) {
  // src/morton.jl:141-157
  if (a.root != b.root) return a.root < b.root;
  int msd = 0;
  uint64_t xmax = padded(a.c[0], a.level) ^ padded(b.c[0], b.level);
  for (int d = 1; d < D; ++d) {
    uint64_t x = padded(a.c[d], a.level) ^ padded(b.c[d], b.level);
    if (less_msb(xmax, x)) { msd = d; xmax = x; }
  }
  if (xmax == 0) return a.level < b.level;
  return padded(a.c[msd], a.level) < padded(b.c[msd], b.level);
}

std::vector<Dir> alldirections(int D) {
  std::vector<Dir> out;
  int n = 1;
  for (int d = 0; d < D; ++d) n *= 3;
  for (int i = 0; i < n; ++i) {
    Dir delta{0, 0, 0};
    int r = i;
    bool zero = true;
    for (int d = 0; d < D; ++d) {
      delta[d] = (r % 3) - 1;
      r /= 3;
      zero = zero && delta[d] == 0;
    }
    if (!zero) out.push_back(delta);
  }
  return out;
}

Forest::Forest(int D_, const int* roots_, int N_, const bool* periodic_,
               const bool* reflecting, const bool* outflow, const double* ext) {
  if (D_ < 1 || D_ > kMaxDim) fail("dimension must be 1, 2 or 3, got " + std::to_string(D_));
  D = D_;
  for (int d = 0; d < D; ++d) {
    if (roots_[d] <= 0) fail("roots must all be positive");
    roots[d] = roots_[d];
    periodic[d] = periodic_ ? periodic_[d] : false;
  }
  // src/forest.jl:343-345
  if (N_ <= 0) fail("N must be positive, got " + std::to_string(N_));
  if (N_ % 2 != 0) fail("N must be even, got " + std::to_string(N_));
  N = N_;
  // src/forest.jl:349-371
  for (int d = 0; d < D; ++d) {
    for (int s = 0; s < 2; ++s) {
      bool r = reflecting ? reflecting[2 * d + s] : false;
      bool o = outflow ? outflow[2 * d + s] : false;
      if (periodic[d] && r)
        fail("dimension " + std::to_string(d + 1) + " is both periodic and reflecting");
      if (periodic[d] && o)
        fail("dimension " + std::to_string(d + 1) + " is both periodic and outflow");
      if (r && o)
        fail("a face of dimension " + std::to_string(d + 1) + " is both reflecting and outflow");
      faces[d][s] = r ? ReflectingFace : o ? OutflowFace : OuterFace;
    }
  }
  // src/forest.jl:376-398: nonempty increasing extents, cubic blocks.
  for (int d = 0; d < D; ++d) {
    extents[d][0] = ext ? ext[2 * d] : 0.0;
    extents[d][1] = ext ? ext[2 * d + 1] : double(roots[d]);
    if (!(extents[d][1] > extents[d][0])) fail("each extent must be nonempty and increasing");
  }
  double h0 = (extents[0][1] - extents[0][0]) / roots[0];
  for (int d = 1; d < D; ++d) {
    double h = (extents[d][1] - extents[d][0]) / roots[d];
    if (std::fabs(h - h0) > 4096 * std::numeric_limits<double>::epsilon() * std::fabs(h0))
      fail("blocks must be cubes: the root spacings differ between dimensions");
  }
  for (int r = 0; r < nroots(); ++r) {
    Key k;
    k.root = r;
    leaves.push_back(k);
  }
}

int Forest::nroots() const {
  int n = 1;
  for (int d = 0; d < D; ++d) n *= roots[d];
  return n;
}

std::array<int, kMaxDim> Forest::root_position(int root) const {
  // src/forest.jl:1266-1274, dimension 1 fastest
  std::array<int, kMaxDim> pos{0, 0, 0};
  int r = root;
  for (int d = 0; d < D; ++d) {
    pos[d] = r % roots[d];
    r /= roots[d];
  }
  return pos;
}

int Forest::root_index(const std::array<int, kMaxDim>& pos) const {
  int idx = 0, stride = 1;
  for (int d = 0; d < D; ++d) {
    idx += pos[d] * stride;
    stride *= roots[d];
  }
  return idx;
}

int Forest::find_leaf(const Key& k) const {
  auto it = std::lower_bound(leaves.begin(), leaves.end(), k,
                             [this](const Key& a, const Key& b) { return curve_less(a, b, D); });
  if (it != leaves.end() && *it == k) return int(it - leaves.begin());
  return -1;
}

int Forest::find_covering_leaf(int root, int lvl, std::array<int64_t, kMaxDim> c) const {
  for (int l = lvl; l >= 0; --l) {
    Key k;
    k.root = root;
    k.level = l;
    for (int d = 0; d < D; ++d) k.c[d] = uint32_t(c[d]);
    int i = find_leaf(k);
    if (i >= 0) return i;
    for (int d = 0; d < D; ++d) c[d] >>= 1;
  }
  return -1;
}

bool Forest::neighbor_anchor(const Key& k, const Dir& delta, int& root,
                             std::array<int64_t, kMaxDim>& coords) const {
  int lvl = k.level;
  auto rp = root_position(k.root);
  std::array<int64_t, kMaxDim> g{0, 0, 0};
  for (int d = 0; d < D; ++d) {
    g[d] = (int64_t(rp[d]) << lvl) + int64_t(k.c[d]) + delta[d];
    int64_t n = int64_t(roots[d]) << lvl;
    bool outside = g[d] < 0 || g[d] >= n;
    if (outside && !periodic[d]) return false;
    g[d] = ((g[d] % n) + n) % n;
  }
  std::array<int, kMaxDim> pos{0, 0, 0};
  int64_t mask = (int64_t(1) << lvl) - 1;
  for (int d = 0; d < D; ++d) {
    pos[d] = int(g[d] >> lvl);
    coords[d] = g[d] & mask;
  }
  for (int d = D; d < kMaxDim; ++d) coords[d] = 0;
  root = root_index(pos);
  return true;
}

void Forest::collect_touching(std::vector<int>& out, int root, int lvl,
                              std::array<int64_t, kMaxDim> coords, const Dir& delta) const {
  // src/forest.jl:1655-1675: tangential dimensions span both children,
  // normal ones only the one facing the asking block; dimension 1 fastest.
  int child = lvl + 1;
  std::array<int64_t, kMaxDim> base{0, 0, 0};
  std::array<int, kMaxDim> span{1, 1, 1};
  for (int d = 0; d < D; ++d) {
    base[d] = 2 * coords[d] + (delta[d] == -1 ? 1 : 0);
    span[d] = delta[d] == 0 ? 2 : 1;
  }
  for (int o2 = 0; o2 < span[2]; ++o2)
    for (int o1 = 0; o1 < span[1]; ++o1)
      for (int o0 = 0; o0 < span[0]; ++o0) {
        std::array<int64_t, kMaxDim> cc{base[0] + o0, base[1] + o1, base[2] + o2};
        for (int d = D; d < kMaxDim; ++d) cc[d] = 0;
        Key ck;
        ck.root = root;
        ck.level = child;
        for (int d = 0; d < D; ++d) ck.c[d] = uint32_t(cc[d]);
        int j = find_leaf(ck);
        if (j < 0)
          collect_touching(out, root, child, cc, delta);
        else
          out.push_back(j);
      }
}

std::vector<int> Forest::neighbors(const Key& k, const Dir& delta) const {
  std::vector<int> out;
  int root;
  std::array<int64_t, kMaxDim> coords;
  if (!neighbor_anchor(k, delta, root, coords)) return out;
  int i = find_covering_leaf(root, k.level, coords);
  if (i >= 0) {
    out.push_back(i);
    return out;
  }
  collect_touching(out, root, k.level, coords, delta);
  return out;
}

Dir Forest::wall_direction(const Key& k, const Dir& delta,
                           std::array<int8_t, kMaxDim>& walls) const {
  int64_t n = int64_t(1) << k.level;
  auto rp = root_position(k.root);
  Dir out{0, 0, 0};
  for (int d = 0; d < kMaxDim; ++d) walls[d] = OuterFace;
  for (int d = 0; d < D; ++d) {
    if (delta[d] != 0 && !periodic[d]) {
      int64_t stepped = int64_t(k.c[d]) + delta[d];
      bool exits = delta[d] < 0 ? (stepped < 0 && rp[d] == 0)
                                : (stepped >= n && rp[d] == roots[d] - 1);
      if (exits) walls[d] = faces[d][delta[d] < 0 ? 0 : 1];
    }
    out[d] = walls[d] == OuterFace ? delta[d] : 0;
  }
  return out;
}

void Forest::refine(const std::vector<Key>& keys) {
  std::vector<char> mark(leaves.size(), 0);
  bool any = false;
  for (const Key& k : keys) {
    int i = find_leaf(k);
    if (i < 0) fail("cannot refine a key that is not a leaf of the forest");
    if (k.level >= kMaxLevel) fail("cannot refine a key at MAX_LEVEL");
    mark[i] = 1;
    any = true;
  }
  if (!any) return;
  std::vector<Key> out;
  out.reserve(leaves.size() + keys.size() * ((1 << D) - 1));
  for (size_t i = 0; i < leaves.size(); ++i) {
    if (!mark[i]) {
      out.push_back(leaves[i]);
      continue;
    }
    std::vector<Key> kids;
    for (int o = 0; o < (1 << D); ++o) {
      Key c = leaves[i];
      c.level += 1;
      for (int d = 0; d < D; ++d) c.c[d] = 2 * c.c[d] + ((o >> d) & 1);
      kids.push_back(c);
    }
    std::sort(kids.begin(), kids.end(),
              [this](const Key& a, const Key& b) { return curve_less(a, b, D); });
    out.insert(out.end(), kids.begin(), kids.end());
  }
  leaves.swap(out);
  generation += 1;
}

int Forest::coarsen(const std::vector<Key>& keys) {
  std::vector<char> flag(leaves.size(), 0);
  for (const Key& k : keys) {
    const int i = find_leaf(k);
    if (i < 0) fail("cannot coarsen a key that is not a leaf of the forest");
    flag[i] = 1;
  }
  auto parent_of = [this](const Key& k) {
    Key p = k;
    p.level -= 1;
    for (int d = 0; d < D; ++d) p.c[d] = k.c[d] >> 1;
    return p;
  };
  // Candidate families: all children leaves and flagged. Children of one
  // parent are consecutive along the curve.
  const int nkids = 1 << D;
  std::vector<Key> parents;
  for (size_t i = 0; i + nkids <= leaves.size(); ++i) {
    const Key& k = leaves[i];
    if (k.level == 0 || !flag[i]) continue;
    const Key p = parent_of(k);
    bool family = true;
    for (int o = 0; o < nkids && family; ++o) {
      const Key& x = leaves[i + o];
      family = x.level == k.level && parent_of(x) == p && flag[i + o];
    }
    if (!family) continue;
    parents.push_back(p);
    i += nkids - 1;
  }
  const std::vector<Key> before = leaves;
  // Drop families whose parent would break 2:1 balance, until none does.
  while (true) {
    leaves.clear();
    size_t q = 0;
    for (size_t i = 0; i < before.size();) {
      if (q < parents.size() && before[i].level > 0 && parent_of(before[i]) == parents[q]) {
        leaves.push_back(parents[q++]);
        i += nkids;
      } else {
        leaves.push_back(before[i++]);
      }
    }
    std::vector<Key> keep;
    bool dropped = false;
    auto dirs = alldirections(D);
    for (const Key& p : parents) {
      bool ok = true;
      for (const Dir& delta : dirs) {
        if (!ok) break;
        for (int j : neighbors(p, delta))
          if (leaves[j].level > p.level + 1) ok = false;
      }
      if (ok) keep.push_back(p);
      else dropped = true;
    }
    if (!dropped) break;
    parents.swap(keep);
  }
  if (!parents.empty()) generation += 1;
  return int(parents.size());
}

void Forest::balance() {
  auto dirs = alldirections(D);
  while (true) {
    std::vector<Key> found;
    for (const Key& k : leaves) {
      if (k.level < 2) continue;
      for (const Dir& delta : dirs) {
        int root;
        std::array<int64_t, kMaxDim> coords;
        if (!neighbor_anchor(k, delta, root, coords)) continue;
        int i = find_covering_leaf(root, k.level, coords);
        if (i < 0) continue;
        if (leaves[i].level < k.level - 1) found.push_back(leaves[i]);
      }
    }
    if (found.empty()) break;
    std::sort(found.begin(), found.end(),
              [this](const Key& a, const Key& b) { return curve_less(a, b, D); });
    found.erase(std::unique(found.begin(), found.end()), found.end());
    refine(found);
  }
}

bool Forest::isbalanced() const {
  auto dirs = alldirections(D);
  for (const Key& k : leaves)
    for (const Dir& delta : dirs)
      for (int j : neighbors(k, delta))
        if (std::abs(leaves[j].level - k.level) > 1) return false;
  return true;
}

}  // namespace tah
