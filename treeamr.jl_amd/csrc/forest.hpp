// Host-side forest: Morton keys, a brick of roots, neighbour search,
// refinement and 2:1 balance.
//
// Restates the parts of TreeAMR.jl's tree that the ghost-fill schedule
// reads: src/morton.jl (keys and curve order) and src/forest.jl (brick,
// face kinds, neighbour search, refine!/balance!). Rotating seams, glued
// patches and MPI distribution are not part of this library.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace tah {

constexpr int kMaxDim = 3;
constexpr int kMaxLevel = 32;

// What fills the ghosts beyond a face of the brick (src/forest.jl:26).
enum FaceKind : int8_t { OuterFace = 0, ReflectingFace = 1, OutflowFace = 2 };

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
// Status codes shared with the C ABI (include/tah.h).
enum : int { kOk = 0, kInvalid = 1, kStale = 2, kLayout = 3, kHip = 4,
             kNoMem = 5, kUnsupported = 6 };

[[noreturn]] inline void fail(const std::string& m, int code = kInvalid) {
  throw Error(code, m);
}

// One node of the forest (src/morton.jl:25-39): root index, level, and
// per-dimension integer coordinates at that level. Unused dimensions
// (d >= D) hold 0.
struct Key {
  int32_t root = 0;
  int32_t level = 0;
  std::array<uint32_t, kMaxDim> c{0, 0, 0};
  bool operator==(const Key& o) const {
    return root == o.root && level == o.level && c == o.c;
  }
  bool operator!=(const Key& o) const { return !(*this == o); }
};

// Curve order (src/morton.jl:129-157): root first, then Z-order with
// dimension 1 most significant, an ancestor before its descendants.
bool curve_less(const Key& a, const Key& b, int D);

using Dir = std::array<int, kMaxDim>;

// All 3^D - 1 nonzero directions, dimension 1 varying fastest
// (src/forest.jl:1319-1322, column-major Iterators.product).
std::vector<Dir> alldirections(int D);

struct Forest {
  int D = 3;
  std::array<int, kMaxDim> roots{1, 1, 1};
  std::array<bool, kMaxDim> periodic{false, false, false};
  std::array<std::array<int8_t, 2>, kMaxDim> faces{};  // FaceKind per (d, side)
  std::array<std::array<double, 2>, kMaxDim> extents{};
  int N = 8;
  std::vector<Key> leaves;  // sorted by curve_less
  int64_t generation = 0;

  // src/forest.jl:314-341 (the brick constructor) and :349-371.
  Forest(int D, const int* roots, int N, const bool* periodic,
         const bool* reflecting /*[D][2]*/, const bool* outflow /*[D][2]*/,
         const double* extents /*[D][2] or null*/);

  int nleaves() const { return (int)leaves.size(); }
  int nroots() const;
  std::array<int, kMaxDim> root_position(int root) const;
  int root_index(const std::array<int, kMaxDim>& pos) const;

  // Index into leaves of key, or -1 (src/forest.jl:1332-1335).
  int find_leaf(const Key& k) const;
  // The leaf covering (root, lvl, coords) at lvl or coarser, or -1
  // (src/forest.jl:1636-1646).
  int find_covering_leaf(int root, int lvl, std::array<int64_t, kMaxDim> c) const;
  // Same-level node reached by stepping one cell from k by delta, wrapping
  // periodic dimensions; false where the step leaves a non-periodic face
  // (src/forest.jl:1501-1514).
  bool neighbor_anchor(const Key& k, const Dir& delta, int& root,
                       std::array<int64_t, kMaxDim>& coords) const;
  // The leaves across direction delta (src/forest.jl:1718-1739 and the
  // recursive collect_touching_leaves!, :1655-1675), as leaf indices in
  // the order the reference visits them.
  std::vector<int> neighbors(const Key& k, const Dir& delta) const;
  // Split delta at reflecting/outflow faces (src/forest.jl:1617-1631).
  Dir wall_direction(const Key& k, const Dir& delta, std::array<int8_t, kMaxDim>& walls) const;

  // src/forest.jl:1761-1781: replace each key by its 2^D children.
  void refine(const std::vector<Key>& keys);
  // Replace every family whose 2^D children are all listed leaves by its
  // parent (the coarsening flags of regrid!, src/regrid.jl), keeping 2:1
  // balance: a family whose parent would meet a leaf two levels finer
  // stays, as the reference's completed leaves keep it. Other keys are
  // ignored. Returns the number of families coarsened.
  int coarsen(const std::vector<Key>& keys);
  // src/forest.jl:1837-1870: 2:1 across faces, edges and corners.
  void balance();
  bool isbalanced() const;

 private:
  void collect_touching(std::vector<int>& out, int root, int lvl,
                        std::array<int64_t, kMaxDim> coords, const Dir& delta) const;
};

}  // namespace tah
