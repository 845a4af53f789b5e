// Plans: a schedule's tables in HBM for one layout, its groups split into
// kernel classes, and the tiles each stage group launches.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
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
  // Streams the launches of one stage group are spread over, so that
  // memory-bound copies and compute-bound interpolation overlap. Created
  // on first use, default priority.
  std::vector<hipStream_t> streams;
  ~tah_ctx();
};

namespace tah {

// Kernel classes.
constexpr int kClsGeneric = 0;
constexpr int kClsCopy = 1;
// The launch class of copies at CPT values per thread.
constexpr int kClsCopyLaunch = 900;
inline int copy_launch(int cpt) { return kClsCopyLaunch + cpt; }
constexpr int kClsInterp = 2;  // + 8*(index into kInterpTable) + rbits
constexpr int kClsMerged = 1000;  // + index into kInterpTable: every rbits in one launch
constexpr int kClsMergedPair = 1100;  // + index: the same over tiles of variable pairs
constexpr int kClsLds = 1200;  // + index: LDS-staged chunks (interp_lds_kernel)
constexpr int kClsLdsProlong = 1300;  // + index: LDS-staged prolongation by units
constexpr int kClsStream = 1400;  // + index: restriction streamed along z (interp_stream_kernel)

// (P, SHIFT) of each specialised interpolation kernel family.
extern const std::array<std::array<int, 2>, 8> kInterpTable;
int interp_index(int P, int shift);
int class_tile_elems(int cls);

struct PlanGroup {
  int32_t kind = 1, factorcol = 0, stage = 0, points = 0;
  bool unit = false;
  int32_t n[3] = {0, 0, 0}, order[3] = {1, 1, 1}, tfirst[3] = {0, 0, 0};
  std::vector<int32_t> tblocks, sblocks;
  std::array<std::vector<int32_t>, 3> srcstart;
  std::array<std::vector<double>, 3> weights;
  int cls = kClsGeneric;
  std::vector<int32_t> wclasses;  // indices into tah_plan::wclasses
};

// Interpolation units of one group sharing one weight-product table.
struct PlanWClass {
  int32_t group = 0;
  int cls = 0;                  // kernel class (P, SHIFT, rbits)
  std::vector<int32_t> units;   // 8 ints per unit (see DevWClass)
  int64_t wtable = 0;           // offset of its W in tah_plan::wblob (doubles)
  int32_t nunits() const { return int32_t(units.size() / 8); }
};

// One launch: a class over a contiguous range of the plan's tiles.
struct Launch {
  int cls;
  int64_t first, count;
  int64_t values = 0;  // values the launch writes
  int32_t grid = 0;    // copies: workgroups of a bounded grid (0: one per tile)
  int32_t lds = 0;     // LDS-staged: bytes of LDS per workgroup; first/count index LdsTiles
  int32_t threads = kThreads;  // workgroup size (interpolation families and streamed chunks: 64, 128 or 256)
};

// The launches of one stage group in the order they are issued:
// interpolations before copies, then the largest first. Forked, the
// interpolations' workgroups are then dispatched first and the copies fill
// the room they leave; one after another, the copies read the sources
// restriction has just read. The large forest's tuned fill: 11.16-11.22
// against 12.07-12.11 ms with the largest launch (the copies) first; its
// copies alone 5.65 against 6.60 ms. (A batch keeps the largest launch
// first: its forked fill was faster so.)
inline void order_launches(std::vector<Launch>& ls) {
  auto is_copy = [](const Launch& L) { return L.cls > kClsCopyLaunch && L.cls <= kClsCopyLaunch + 16; };
  std::stable_sort(ls.begin(), ls.end(), [&](const Launch& x, const Launch& y) {
    if (is_copy(x) != is_copy(y)) return !is_copy(x);
    return x.count > y.count;
  });
}

// The interpolation families: the index into kInterpTable of order P and
// shift (restriction 2, prolongation 0).
constexpr int kFamilies = 8;

struct AdjointTables;

}  // namespace tah

struct tah_plan {
  tah_ctx* ctx = nullptr;
  tah_layout layout{};
  int32_t dim = 3;
  int32_t nstages = 0;
  std::vector<int32_t> tags;
  std::vector<std::array<int32_t, 2>> stagegroups;
  std::vector<tah::PlanGroup> groups;
  std::vector<tah::PlanWClass> wclasses;
  std::vector<double> wblob;  // every distinct weight-product table
  bool needs_factors = false;
  int64_t values = 0;
  // Device memory: tables, DevGroup and DevWClass arrays, tiles, one Job.
  char* dtables = nullptr;
  std::vector<tah::DevGroup> devgroups;  // host copies (pointers into dtables)
  tah::DevGroup* dgroups = nullptr;
  tah::DevWClass* dwclasses = nullptr;
  std::vector<tah::Tile> tiles;          // host copy, job field 0
  tah::Tile* dtiles = nullptr;
  // LDS-staged launches: the chunks of the groups they cover and their
  // tiles (host copies and device arrays, rebuilt with the launches).
  std::vector<tah::LdsChunk> lds_chunks;
  std::vector<tah::LdsTile> lds_tiles;
  tah::LdsChunk* dchunks = nullptr;
  tah::LdsTile* dlds_tiles = nullptr;
  // ... of prolongation: chunks by units, their class waves and packed units
  std::vector<tah::LdsUnitChunk> lds_uchunks;
  std::vector<tah::LdsWave> lds_waves;
  std::vector<uint32_t> lds_units;
  tah::LdsUnitChunk* duchunks = nullptr;
  tah::LdsWave* dwaves = nullptr;
  uint32_t* dunits = nullptr;
  // ... of restriction streamed along z: chunks and their groups' planes
  std::vector<tah::StreamChunk> stream_chunks;
  std::vector<tah::StreamPlane> stream_planes;
  tah::StreamChunk* dschunks = nullptr;
  tah::StreamPlane* dplanes = nullptr;
  tah::Job* djob = nullptr;
  tah::Job last_job{};
  // A regrid transfer (stage tags TAH_STAGE_REGRID, then
  // TAH_STAGE_REGRID_KEPT): its first stage reads the old array, whose
  // nsources blocks its sources index, through the second Job.
  bool regrid = false;
  int32_t nsources = 0;
  tah::Job* djob2 = nullptr;
  tah::Job last_job2{};
  std::vector<std::vector<tah::Launch>> launches;  // per stage group
  int64_t deferred = 0;            // transfers moved to run with the last stage group
  // Run a stage group's launches forked over the context's streams (1) or
  // one after another on the caller's stream (0, the default; tah_tune
  // picks per plan).
  int32_t fork = 0;
  // Run every unit class of one order and shift in one launch (1, the
  // default) or each class in a launch of its own (0).
  int32_t merge = 1;
  // Move phase-1 transfers that no later stage reads or overwrites into the
  // last stage group (1, the default) or keep them in phase 1 (0).
  int32_t defer = 1;
  // Values per thread of the copy kernel's tiles: 4, 8 (default) or 16.
  int32_t copy_per_thread = 8;
  // Workgroups of the copies' bounded grid (a multiple of 8), or 0 (the
  // default) for one workgroup per tile.
  int32_t copy_wgs = 0;
  int32_t copy_order = 0;  // copy tiles by source block (0) or target block (1)
  // Merged interpolation launches over tiles of two neighbouring variables
  // per thread where the layout has variables innermost and an even count
  // (1, the default), or of one variable per thread (0).
  int32_t interp_pairs = 1;
  // Interpolation from LDS copies of its source boxes. Bit 2 (the default):
  // restriction streamed along z through LDS (interp_stream_kernel), in any
  // layout; bit 3 (the default where variables are not innermost):
  // prolongation streamed the same way. Bits 0 and 1, where the layout holds
  // variables innermost in even number: restriction (without bit 2) and the
  // prolongation of the unit classes of at least 64 units (faces) from
  // chunks of their source boxes. The rest through the unit kernels.
  int32_t lds = 4;
  // In a batched fill of plans that all have it set: one persistent launch
  // per stage position over a work queue of every plan's tiles (1), or a
  // launch per kernel class (0, the default: on the 64-forest batch the
  // queue took 26.4 ms against 21.5 ms, its one register budget holding the
  // copies to four waves per SIMD).
  int32_t queue = 0;
  // Per interpolation family (kInterpTable index), what the plan-wide
  // options above set for every family until tah_plan_set_class_options
  // (or the tuner) sets one apart: its workgroup size (the units and
  // variables of a tile fill that many threads), one launch for all its
  // unit classes, variable pairs (prolongation), LDS staging, and the
  // variables of a tile (0: every one) where variables are innermost.
  std::array<int32_t, tah::kFamilies> fam_threads{}, fam_merge{}, fam_pairs{}, fam_lds{}, fam_vars{};
  // The outer-face hook (tah_fill_boundary): the regions in HBM and their
  // tiles -- every region's first (the values form, after phase 1's stage
  // group), then per stage group those of the regions whose slot stage it
  // holds (the relative form, after that group) -- and, per stage group,
  // the boxes the relative form reads after it (block, lo[3], hi[3]
  // inclusive), which no transfer of that group or before may be deferred
  // past.
  std::vector<tah::HookRegion> hooks;
  std::vector<int32_t> hook_slots;
  int64_t hook_points = 0;
  tah::HookRegion* dhooks = nullptr;
  tah::HookTile* dhook_tiles = nullptr;
  int64_t hook_all[2] = {0, 0};
  std::vector<std::array<int64_t, 2>> hook_after;
  std::vector<std::vector<std::array<int32_t, 7>>> hook_reads;
  int32_t phase1_sg = -1;
  std::vector<hipEvent_t> events;  // fork/join events, created on first use
  // The adjoint's inverted schedule, built on first use (adjoint.hip).
  tah::AdjointTables* adj = nullptr;
  // A number no other plan, and no earlier state of this one, ever had:
  // taken anew whenever the tiles and launches are (re)built, so that a
  // batch built from this plan can tell that it is stale.
  uint64_t serial = 0;
  // Calls on one plan rewrite its device Job (and the adjoint its staging
  // buffer) on the caller's stream. A call on another stream than the
  // previous one first waits for `done`, recorded after the previous call's
  // last launch on `last_stream`.
  hipStream_t last_stream = nullptr;
  bool pending = false;
  hipEvent_t done = nullptr;
  ~tah_plan();
};

namespace tah {
// A unit copy whose source row is affine in the target row along every
// dimension: ca + cs*r with |cs| <= max_step.
bool affine_copy(const PlanGroup& g, int32_t ca[3], int32_t cs[3], int max_step);
// A layout's strides: positive, the five axes nested without overlap,
// blocks outermost (TAH_ELAYOUT otherwise).
void validate_strides(const tah_layout* l);
// Split groups into classes, upload tables, build tiles and launches.
void build_plan_device(tah_plan* p);
// (Re)build the tiles and launches for the plan's options.
void build_launches(tah_plan* p);
// The Job of one fill of `work` by plan p; sources from `src` (a regrid's
// old array) where given, else from work.
Job plan_job(const tah_plan* p, double* work, const double* factors, const double* src = nullptr);
// Streams of the context (created on first use).
const std::vector<hipStream_t>& ctx_streams(tah_ctx* c);
// Run a stage group's launches forked over the context's streams and
// joined back into `s`. ev must hold launches.size() + 1 events.
void run_launches(tah_ctx* c, const std::vector<Launch>& ls, hipStream_t s, const Job* jobs,
                  const Tile* tiles, const LdsTile* ltiles, hipEvent_t* ev, bool fork);
// The batched fill's work queue (fill.hip): the queue_kernel code of a
// launch class (-1: not in the queue), its launch over shares of tiles, the
// reset of its counters, and the waves a CU holds of it.
int queue_kind(int cls);
void launch_queue(uint32_t grid, hipStream_t s, const Job* jobs, const Tile* tiles, const uint32_t* share,
                  uint32_t* counters);
void launch_queue_reset(hipStream_t s, uint32_t* counters, uint32_t n);
int queue_waves_per_cu();
// An LDS-staged class over ntiles tiles with `bytes` of LDS per workgroup
// (fill.hip); threads: the workgroup size of a streamed class.
void launch_lds(int cls, uint32_t ntiles, hipStream_t s, const Job* jobs, const LdsTile* tiles, int32_t bytes,
                int32_t threads = kThreads);
// Launch one Launch (copy, interpolation or LDS-staged) on stream s.
void launch_one(const Launch& L, hipStream_t s, const Job* jobs, const Tile* tiles, const LdsTile* ltiles);
// Launch one class over ntiles tiles (fill.hip); grid > 0: a bounded grid.
void launch_class(int cls, uint32_t ntiles, hipStream_t s, const Job* jobs, const Tile* tiles, int32_t grid = 0,
                  int32_t threads = kThreads);
// Every family's options from the plan-wide ones.
void families_from_options(tah_plan* p);
// Order a call on plan p and stream s after the plan's previous call (on
// whatever stream it ran), before the call rewrites the plan's device Job
// or staging buffer; plan_end marks the call's last launch. A stream under
// graph capture is ordered by the graph's own launch instead: no event is
// waited on or recorded there.
void plan_begin(tah_plan* p, hipStream_t s);
void plan_end(tah_plan* p, hipStream_t s);
// Upload p's Job for this call where it differs from the last one uploaded.
void plan_upload_job(tah_plan* p, const Job& job, hipStream_t s);
// Whether the serial is that of a live plan in its current state, and the
// registry behind it (batch.hip).
bool plan_serial_live(const tah_plan* p, uint64_t serial);
void plan_register(tah_plan* p);
void plan_unregister(tah_plan* p);
// The no-allocation adjoint's staging buffer: at most this many doubles
// (64 MiB) per GPU, held by every GPU the library runs on until the process
// ends. noalloc_scratch_init takes its device address, once per device, with
// that device current (adjoint.hip; called when a context is created).
constexpr int64_t kNoallocScratchValues = int64_t(8) << 20;
void noalloc_scratch_init(int device);
// Drop every cached batch that uses plan p (or that belongs to context c),
// after the device has finished with it (batch.hip).
void batch_forget_plan(const tah_plan* p);
void batch_forget_ctx(const tah_ctx* c);
}  // namespace tah
