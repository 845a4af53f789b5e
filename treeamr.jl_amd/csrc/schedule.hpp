// The ghost-fill schedule: the neighbour search turned into transfer
// groups of tensor-product 1-D stencils, split into stages.
//
// Restates src/schedule.jl: Stencil1D (:64-93), TransferGroup (:150-197),
// GroupKey and keyorder (:400-506), the stencil builders (:876-1395),
// block_sources! (:1525-1610), mirror_sources! (:2237-2297),
// transfer_stage (:640-658), stage_groups (:2876-2931) and the
// GhostSchedule builder (:3075-3292), for brick forests without seams,
// glue or element centering, with unbounded operators.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <vector>

#include "forest.hpp"
#include "operators.hpp"

namespace tah {

enum Kind : int { KindCopy = 1, KindRestrict = 2, KindProlong = 3 };

// Stage tags in the order a fill runs them (src/schedule.jl:640-644):
// stage P (the wall-plane projection) first, then stage 0, phase 1, and
// one stage per prolongation target level, coarsest first.
constexpr int kStageP = -1;
constexpr int kUnifyTag = 0;
constexpr int kPhase1Tag = 1;
inline int prolongation_tag(int level) { return 2 + level; }
// The regrid transfer (REGRID_TAG, src/regrid.jl:755-803): old array -> new
// array; and its kept-over-created pass within the new array
// (REGRID_KEPT_TAG, src/regrid.jl:810-933).
constexpr int kRegridTag = 100;
constexpr int kRegridKeptTag = 101;

// Parities per variable and dimension (src/storage.jl, NoParity /
// EvenParity / OddParity).
enum Parity : int8_t { NoParity = 0, EvenParity = 1, OddParity = 2 };

// One dimension of a transfer (src/schedule.jl:64-93), in 1-based stored
// indices as the reference builds it.
struct Stencil1D {
  int targetfirst = 1;
  int order = 1;
  std::vector<int32_t> srcstart;  // per target row, 1-based
  std::vector<double> weights;    // order x ntarget, column-major
  bool unit = true;
  std::vector<int32_t> single;    // row of the only nonzero weight, or 0
  int ntarget() const { return (int)srcstart.size(); }
  // src/schedule.jl:78-87
  void finish();
};

struct GroupKey {
  int kind = KindCopy;
  Dir direction{0, 0, 0};
  Dir offset{0, 0, 0};
  int level = 0;
  std::array<int8_t, kMaxDim> mirror{0, 0, 0};
  int8_t orientation = 0;
  int8_t axes = 0;
  std::array<int8_t, kMaxDim> span{-1, -1, -1};
  bool isunify(int D) const {
    for (int d = 0; d < D; ++d)
      if (span[d] < 0) return false;
    return true;
  }
};
// keyorder (src/schedule.jl:493-497).
bool key_less(const GroupKey& a, const GroupKey& b);
struct KeyLess {
  bool operator()(const GroupKey& a, const GroupKey& b) const { return key_less(a, b); }
};

struct TransferGroup {
  GroupKey key;
  int kind = KindCopy;
  std::array<Stencil1D, kMaxDim> stencils;  // dims >= D: one unit point
  std::vector<int32_t> targetblocks;        // 0-based, non-decreasing
  std::vector<int32_t> sourceblocks;        // 0-based
  int factorcol = 0;                        // 0 or 1-based factor column
  int stage = kPhase1Tag;
  int64_t points() const {
    int64_t n = 1;
    for (const auto& s : stencils) n *= s.ntarget();
    return n;
  }
};

// A region of an outer face, filled by the boundary hook, never by a
// transfer (BoundaryRegion, src/schedule.jl:200-229). For the relative hook:
// clamps, per dimension, the 1-based stored index of the nearest evolved
// point along a component of the direction that leaves the domain through
// an outer face (0 along every other, where the point keeps the ghost's own
// index), and slot, the tag of the stage after which the ghosts it reads
// are filled.
struct BoundaryRegion {
  int32_t block;
  Dir direction;
  std::array<int, kMaxDim> lo, hi;  // 1-based inclusive stored ranges
  std::array<int, kMaxDim> clamps{0, 0, 0};
  int slot = kPhase1Tag;
};

struct Stage {
  int tag;
  std::vector<int> groups;  // indices into GhostSchedule::groups
};

struct RegridFrom {
  const std::vector<Key>& oldleaves;  // the leaves before the regrid, sorted
};

// Stages P and 0 alone, with the restriction the geometry fixes (a
// UnifySchedule, src/unify.jl:76-170).
struct UnifyOnly {};

struct GhostSchedule {
  const Forest* forest = nullptr;
  int64_t generation = 0;
  int32_t nsources = 0;  // a regrid's old leaves (sources); 0: the forest's leaves
  int D = 3, N = 8;
  std::array<int, kMaxDim> G{0, 0, 0}, c{0, 0, 0}, stored{1, 1, 1};
  Operators ops;
  std::vector<std::array<int8_t, kMaxDim>> parities;  // per variable; may be empty
  std::vector<TransferGroup> groups;  // in fill order: by stage, then keyorder
  std::vector<Stage> stages;          // in fill order
  std::vector<std::array<int, 2>> stagegroups;  // [first, last] stage indices
  std::vector<BoundaryRegion> boundaries;

  GhostSchedule(const Forest& f, const Operators& ops, const int* G, const int* c,
                const std::vector<std::array<int8_t, kMaxDim>>& parities);
  // The regrid transfer onto f's leaves from `from.oldleaves`
  // (transfer_groups and regrid_stage, src/regrid.jl:755-803, 935-947; the
  // kept-over-created pass, regrid_kept_pairs and regrid_kept_stage,
  // :826-933): stage kRegridTag, and kRegridKeptTag where the set has a
  // vertex-like dimension. Targets index f's leaves, sources of the first
  // stage the old leaves, of the second f's.
  GhostSchedule(const RegridFrom& from, const Forest& f, const Operators& ops, const int* G, const int* c,
                const std::vector<std::array<int8_t, kMaxDim>>& parities);
  // UnifySchedule(fs) (src/unify.jl:76-170): stage P and stage 0 of a set
  // with a vertex-like dimension, its restriction injection along a
  // vertex-like dimension and the exact two-cell average (order 2) along a
  // cell-like one, for any ghost width (none needed).
  GhostSchedule(UnifyOnly, const Forest& f, const int* G, const int* c,
                const std::vector<std::array<int8_t, kMaxDim>>& parities);

 private:
  void check_layout(const int* G, const int* c, bool operators = true);
  void add_group(const GroupKey& key, const std::vector<int32_t>& tb, const std::vector<int32_t>& sb, int stage);
};

// src/storage.jl:536-554: factors[v + nvars*(col-1)] for col in 1..3^D.
std::vector<double> parity_factors(int D, const std::vector<std::array<int8_t, kMaxDim>>& parities);

}  // namespace tah
