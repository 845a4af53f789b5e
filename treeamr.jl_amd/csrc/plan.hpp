// Plans: a schedule's tables in HBM for one layout, its groups split into
// kernel classes, and the tiles each stage group launches.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "device.hpp"
#include "forest.hpp"
#include "tah.h"

#define HIPCHECK(expr)                                                                    \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::tah::Error(::tah::kHip, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

struct tah_ctx {
  int device = 0;
  hipDeviceProp_t prop{};
};

namespace tah {

// Kernel classes.
constexpr int kClsGeneric = 0;
constexpr int kClsCopy = 1;
constexpr int kClsInterp = 2;  // + 8*(index into kInterpTable) + rbits

// (P, SHIFT) of each specialised interpolation kernel family.
extern const std::array<std::array<int, 2>, 8> kInterpTable;
int interp_index(int P, int shift);
void launch_class(int cls, uint32_t ntiles, hipStream_t s, const Job* jobs, const Tile* tiles,
                  const TileBox* boxes, uint32_t lds_bytes);
int class_tile_elems(int cls);

struct PlanGroup {
  int32_t kind = 1, factorcol = 0, stage = 0, points = 0;
  bool unit = false;
  int32_t n[3] = {0, 0, 0}, order[3] = {1, 1, 1}, tfirst[3] = {0, 0, 0};
  std::vector<int32_t> tblocks, sblocks;
  std::array<std::vector<int32_t>, 3> srcstart;
  std::array<std::vector<double>, 3> weights;
  int cls = kClsGeneric;
  std::vector<int32_t> subs;  // indices into tah_plan::subs (interp classes)
};

struct PlanSub {
  int32_t group = 0;
  int cls = 0;
  int32_t cnt[3] = {0, 0, 0};
  std::array<std::vector<int32_t>, 3> rows;
};

// One launch: a class over a contiguous range of the plan's tiles.
struct Launch {
  int cls;
  int64_t first, count;
  uint32_t lds;  // dynamic LDS bytes per workgroup
};

// Largest LDS box an interpolation tile may stage, in doubles (64 KiB).
constexpr int64_t kMaxBox = 8192;

}  // namespace tah

struct tah_plan {
  tah_ctx* ctx = nullptr;
  tah_layout layout{};
  int32_t dim = 3;
  int32_t nstages = 0;
  std::vector<int32_t> tags;
  std::vector<std::array<int32_t, 2>> stagegroups;
  std::vector<tah::PlanGroup> groups;
  std::vector<tah::PlanSub> subs;
  bool needs_factors = false;
  int64_t values = 0;
  // Device memory: tables, DevGroup and DevSub arrays, tiles, one Job.
  char* dtables = nullptr;
  std::vector<tah::DevGroup> devgroups;  // host copies (pointers into dtables)
  tah::DevGroup* dgroups = nullptr;
  tah::DevSub* dsubs = nullptr;
  std::vector<tah::Tile> tiles;          // host copy, job field 0
  std::vector<tah::TileBox> boxes;       // parallel to tiles (interpolation)
  tah::Tile* dtiles = nullptr;
  tah::TileBox* dboxes = nullptr;
  tah::Job* djob = nullptr;
  tah::Job last_job{};
  std::vector<std::vector<tah::Launch>> launches;  // per stage group
  ~tah_plan();
};

namespace tah {
// Split groups into classes, upload tables, build tiles and launches.
void build_plan_device(tah_plan* p);
// The Job of one fill of `work` by plan p.
Job plan_job(const tah_plan* p, double* work, const double* factors);
}  // namespace tah
