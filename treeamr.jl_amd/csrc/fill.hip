// Device side of the ghost fill: plans (the schedule's tables in HBM, its
// groups split into kernel classes and tiles), the fills, and their C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "api_common.hpp"
#include "device.hpp"
#include "forest.hpp"
#define TAH_DEFINE_KERNELS
#include "kernels.hpp"
#include "plan.hpp"

using namespace tah;

namespace tah {

// (P, SHIFT) pairs with specialised interpolation kernels; the class id of
// one is kClsInterp + 8*index + rbits, rbits having bit d set where R_d = 2.
const std::array<std::array<int, 2>, 8> kInterpTable = {
    {{2, 0}, {3, 0}, {4, 0}, {5, 0}, {6, 0}, {2, 2}, {4, 2}, {6, 2}}};

int interp_index(int P, int shift) {
  for (size_t i = 0; i < kInterpTable.size(); ++i)
    if (kInterpTable[i][0] == P && kInterpTable[i][1] == shift) return int(i);
  return -1;
}

namespace {

template <int P, int S>
void launch_interp_r(int rbits, dim3 grid, dim3 blk, hipStream_t s, const Job* j, const Tile* t) {
  switch (rbits) {
    case 0: hipLaunchKernelGGL((interp_kernel<P, S, 1, 1, 1>), grid, blk, 0, s, j, t); break;
    case 1: hipLaunchKernelGGL((interp_kernel<P, S, 2, 1, 1>), grid, blk, 0, s, j, t); break;
    case 2: hipLaunchKernelGGL((interp_kernel<P, S, 1, 2, 1>), grid, blk, 0, s, j, t); break;
    case 3: hipLaunchKernelGGL((interp_kernel<P, S, 2, 2, 1>), grid, blk, 0, s, j, t); break;
    case 4: hipLaunchKernelGGL((interp_kernel<P, S, 1, 1, 2>), grid, blk, 0, s, j, t); break;
    case 5: hipLaunchKernelGGL((interp_kernel<P, S, 2, 1, 2>), grid, blk, 0, s, j, t); break;
    case 6: hipLaunchKernelGGL((interp_kernel<P, S, 1, 2, 2>), grid, blk, 0, s, j, t); break;
    default: hipLaunchKernelGGL((interp_kernel<P, S, 2, 2, 2>), grid, blk, 0, s, j, t); break;
  }
}

}  // namespace

void launch_class(int cls, uint32_t ntiles, hipStream_t s, const Job* jobs, const Tile* tiles, int32_t bounded,
                  int32_t threads) {
  dim3 grid(ntiles), blk(kThreads);
  // an interpolation family's workgroup size: its tiles fill that many
  // threads (the kernels are bounded for 256 and let threads past a tile's
  // work go)
  const bool interp = cls >= kClsInterp && cls != copy_launch(4) && cls != copy_launch(8) && cls != copy_launch(16);
  if (threads != kThreads) {
    if (!interp || (threads != 64 && threads != 128)) fail("bad workgroup size for class " + std::to_string(cls));
    blk = dim3(uint32_t(threads));
  }
  if (bounded > 0) {
    // a copy class from a bounded grid (a multiple of 8, at most ntiles)
    if (bounded % 8 != 0 || uint32_t(bounded) > ntiles) fail("bad bounded grid for a copy launch");
    dim3 g(static_cast<uint32_t>(bounded));
    if (cls == copy_launch(4)) hipLaunchKernelGGL(copy_loop_kernel<4>, g, blk, 0, s, jobs, tiles, ntiles);
    else if (cls == copy_launch(8)) hipLaunchKernelGGL(copy_loop_kernel<8>, g, blk, 0, s, jobs, tiles, ntiles);
    else if (cls == copy_launch(16)) hipLaunchKernelGGL(copy_loop_kernel<16>, g, blk, 0, s, jobs, tiles, ntiles);
    else fail("a bounded grid is only for copies");
    HIPCHECK(hipGetLastError());
    return;
  }
  if (cls == kClsGeneric) {
    hipLaunchKernelGGL(generic_kernel, grid, blk, 0, s, jobs, tiles);
  } else if (cls == copy_launch(4)) {
    hipLaunchKernelGGL(copy_kernel<4>, grid, blk, 0, s, jobs, tiles);
  } else if (cls == copy_launch(8)) {
    hipLaunchKernelGGL(copy_kernel<8>, grid, blk, 0, s, jobs, tiles);
  } else if (cls == copy_launch(16)) {
    hipLaunchKernelGGL(copy_kernel<16>, grid, blk, 0, s, jobs, tiles);
  } else if (cls >= kClsMergedPair) {
    switch (cls - kClsMergedPair) {
      case 0: hipLaunchKernelGGL((interp_pair_kernel<2, 0>), grid, blk, 0, s, jobs, tiles); break;
      case 1: hipLaunchKernelGGL((interp_pair_kernel<3, 0>), grid, blk, 0, s, jobs, tiles); break;
      case 2: hipLaunchKernelGGL((interp_pair_kernel<4, 0>), grid, blk, 0, s, jobs, tiles); break;
      case 3: hipLaunchKernelGGL((interp_pair_kernel<5, 0>), grid, blk, 0, s, jobs, tiles); break;
      case 4: hipLaunchKernelGGL((interp_pair_kernel<6, 0>), grid, blk, 0, s, jobs, tiles); break;
      default: fail("no variable-pair kernel for restriction (class " + std::to_string(cls) + ")");
    }
  } else if (cls >= kClsMerged) {
    switch (cls - kClsMerged) {
      case 0: hipLaunchKernelGGL((interp_all_kernel<2, 0>), grid, blk, 0, s, jobs, tiles); break;
      case 1: hipLaunchKernelGGL((interp_all_kernel<3, 0>), grid, blk, 0, s, jobs, tiles); break;
      case 2: hipLaunchKernelGGL((interp_all_kernel<4, 0>), grid, blk, 0, s, jobs, tiles); break;
      case 3: hipLaunchKernelGGL((interp_all_kernel<5, 0>), grid, blk, 0, s, jobs, tiles); break;
      case 4: hipLaunchKernelGGL((interp_all_kernel<6, 0>), grid, blk, 0, s, jobs, tiles); break;
      case 5: hipLaunchKernelGGL((interp_all_kernel<2, 2>), grid, blk, 0, s, jobs, tiles); break;
      case 6: hipLaunchKernelGGL((interp_all_kernel<4, 2>), grid, blk, 0, s, jobs, tiles); break;
      default: hipLaunchKernelGGL((interp_all_kernel<6, 2>), grid, blk, 0, s, jobs, tiles); break;
    }
  } else {
    int id = cls - kClsInterp, ps = id / 8, rbits = id % 8;
    switch (ps) {
      case 0: launch_interp_r<2, 0>(rbits, grid, blk, s, jobs, tiles); break;
      case 1: launch_interp_r<3, 0>(rbits, grid, blk, s, jobs, tiles); break;
      case 2: launch_interp_r<4, 0>(rbits, grid, blk, s, jobs, tiles); break;
      case 3: launch_interp_r<5, 0>(rbits, grid, blk, s, jobs, tiles); break;
      case 4: launch_interp_r<6, 0>(rbits, grid, blk, s, jobs, tiles); break;
      case 5: launch_interp_r<2, 2>(rbits, grid, blk, s, jobs, tiles); break;
      case 6: launch_interp_r<4, 2>(rbits, grid, blk, s, jobs, tiles); break;
      default: launch_interp_r<6, 2>(rbits, grid, blk, s, jobs, tiles); break;
    }
  }
  HIPCHECK(hipGetLastError());
}

void launch_lds(int cls, uint32_t ntiles, hipStream_t s, const Job* jobs, const LdsTile* tiles, int32_t bytes,
                int32_t threads) {
  if (bytes <= 0 || bytes > kLdsMaxBytes) fail("bad LDS size for an LDS-staged launch");
  const dim3 grid(ntiles), blk(kThreads);
  const size_t sh = size_t(bytes);
  if (cls >= kClsStream) {
    if (threads != 64 && threads != 128 && threads != kThreads) fail("bad workgroup size for a streamed launch");
    const dim3 blk(threads);
    switch (cls - kClsStream) {
      case 0: hipLaunchKernelGGL((interp_stream_kernel<2, stream_ka(2, 0)>), grid, blk, sh, s, jobs, tiles); break;
      case 1: hipLaunchKernelGGL((interp_stream_kernel<3, stream_ka(3, 0)>), grid, blk, sh, s, jobs, tiles); break;
      case 2: hipLaunchKernelGGL((interp_stream_kernel<4, stream_ka(4, 0)>), grid, blk, sh, s, jobs, tiles); break;
      case 3: hipLaunchKernelGGL((interp_stream_kernel<5, stream_ka(5, 0)>), grid, blk, sh, s, jobs, tiles); break;
      case 4: hipLaunchKernelGGL((interp_stream_kernel<6, stream_ka(6, 0)>), grid, blk, sh, s, jobs, tiles); break;
      case 5: hipLaunchKernelGGL((interp_stream_kernel<2, stream_ka(2, 2)>), grid, blk, sh, s, jobs, tiles); break;
      case 6: hipLaunchKernelGGL((interp_stream_kernel<4, stream_ka(4, 2)>), grid, blk, sh, s, jobs, tiles); break;
      case 7: hipLaunchKernelGGL((interp_stream_kernel<6, stream_ka(6, 2)>), grid, blk, sh, s, jobs, tiles); break;
      default: fail("no streamed restriction kernel for class " + std::to_string(cls));
    }
    HIPCHECK(hipGetLastError());
    return;
  }
  if (cls >= kClsLdsProlong) {
    const dim3 pblk(kLdsProlongThreads);
    switch (cls - kClsLdsProlong) {
      case 0: hipLaunchKernelGGL((interp_lds_prolong_kernel<2>), grid, pblk, sh, s, jobs, tiles); break;
      case 1: hipLaunchKernelGGL((interp_lds_prolong_kernel<3>), grid, pblk, sh, s, jobs, tiles); break;
      case 2: hipLaunchKernelGGL((interp_lds_prolong_kernel<4>), grid, pblk, sh, s, jobs, tiles); break;
      case 3: hipLaunchKernelGGL((interp_lds_prolong_kernel<5>), grid, pblk, sh, s, jobs, tiles); break;
      case 4: hipLaunchKernelGGL((interp_lds_prolong_kernel<6>), grid, pblk, sh, s, jobs, tiles); break;
      default: fail("no LDS-staged prolongation kernel for class " + std::to_string(cls));
    }
    HIPCHECK(hipGetLastError());
    return;
  }
  switch (cls - kClsLds) {
    case 5: hipLaunchKernelGGL((interp_lds_kernel<2>), grid, blk, sh, s, jobs, tiles); break;
    case 6: hipLaunchKernelGGL((interp_lds_kernel<4>), grid, blk, sh, s, jobs, tiles); break;
    case 7: hipLaunchKernelGGL((interp_lds_kernel<6>), grid, blk, sh, s, jobs, tiles); break;
    default: fail("no LDS-staged kernel for class " + std::to_string(cls));
  }
  HIPCHECK(hipGetLastError());
}

int queue_kind(int cls) {
  if (cls == copy_launch(8)) return kQueueCopy;
  if (cls == copy_launch(4)) return kQueueCopy4;
  if (cls == copy_launch(16)) return kQueueCopy16;
  if (cls == kClsGeneric) return kQueueGeneric;
  if (cls >= kClsMerged && cls < kClsMerged + int(kInterpTable.size())) {
    const int i = cls - kClsMerged;
    return (i == 0 || i == 2 || i == 4 || i == 5 || i == 6 || i == 7) ? kQueueMerged + i : -1;
  }
  if (cls >= kClsMergedPair && cls < kClsMergedPair + int(kInterpTable.size())) {
    const int i = cls - kClsMergedPair;
    return (i == 0 || i == 2 || i == 4) ? kQueuePair + i : -1;
  }
  return -1;
}

void launch_queue(uint32_t grid, hipStream_t s, const Job* jobs, const Tile* tiles, const uint32_t* share,
                  uint32_t* counters) {
  if (grid == 0 || grid % 8 != 0) fail("the queue's grid must be a positive multiple of 8");
  hipLaunchKernelGGL(queue_kernel, dim3(grid), dim3(kThreads), 0, s, jobs, tiles, share, counters);
  HIPCHECK(hipGetLastError());
}

void launch_queue_reset(hipStream_t s, uint32_t* counters, uint32_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(queue_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, s, counters, n);
  HIPCHECK(hipGetLastError());
}

int queue_waves_per_cu() { return 4 * kQueueWaves; }

int class_tile_elems(int cls) {
  if (cls > kClsCopyLaunch && cls <= kClsCopyLaunch + 16) return kThreads * (cls - kClsCopyLaunch);
  return cls == kClsGeneric ? kGenericTile : kThreads;
}

}  // namespace tah

namespace {

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

void validate_layout(const tah_schedule_desc* d, const tah_layout* l) {
  if (!d || !l) fail("null argument");
  if (d->dim < 1 || d->dim > 3) fail("schedule dimension must be 1, 2 or 3");
  if (l->nvars < 1) fail("nvars must be positive", kLayout);
  if (l->nblocks != d->nblocks)
    fail("the layout holds " + std::to_string(l->nblocks) + " blocks, the schedule " +
             std::to_string(d->nblocks),
         kLayout);
  for (int k = 0; k < 3; ++k)
    if (l->stored[k] != d->stored[k])
      fail("the layout's stored size differs from the schedule's in dimension " + std::to_string(k + 1),
           kLayout);
  validate_strides(l);
}

}  // namespace

namespace tah {

void validate_strides(const tah_layout* l) {
  if (l->nvars < 1) fail("nvars must be positive", kLayout);
  if (l->nblocks < 0) fail("the block count must not be negative", kLayout);
  for (int k = 0; k < 3; ++k)
    if (l->stored[k] < 1) fail("the layout's stored sizes must be positive", kLayout);
  // Strides positive and the five axes nested without overlap: sorted by
  // stride, each at least the span of the ones inside it. An axis of extent
  // one adds no offset and takes no part.
  int64_t ext[5] = {l->stored[0], l->stored[1], l->stored[2], l->nvars, l->nblocks};
  int order[5] = {0, 1, 2, 3, 4};
  for (int k = 0; k < 5; ++k)
    if (l->stride[k] < 1) fail("layout strides must be positive", kLayout);
  std::sort(order, order + 5, [&](int a, int b) { return l->stride[a] < l->stride[b]; });
  int64_t span = 1;
  for (int k = 0; k < 5; ++k) {
    if (ext[order[k]] == 1) continue;
    if (l->stride[order[k]] < span) fail("the layout's strides overlap", kLayout);
    span = l->stride[order[k]] * ext[order[k]];
  }
  if (l->stride[4] < l->stride[0] * l->stored[0] || l->stride[4] < l->stride[3] * l->nvars)
    fail("blocks must be the outermost axis of the layout", kLayout);
  if (int64_t(l->stored[0]) * l->stored[1] * l->stored[2] * l->nvars > (int64_t(1) << 31))
    fail("a block larger than 2^31 values is not supported", kUnsupported);
}

}  // namespace tah

extern "C" {

tah_status tah_layout_dense(const tah_schedule_desc* d, int32_t nvars, tah_layout* out) {
  return guard([&] {
    if (!d || !out) fail("null argument");
    if (nvars < 1) fail("nvars must be positive");
    tah_layout l{};
    for (int k = 0; k < 3; ++k) l.stored[k] = d->stored[k];
    l.nvars = nvars;
    l.nblocks = d->nblocks;
    l.stride[0] = 1;
    l.stride[1] = l.stored[0];
    l.stride[2] = l.stride[1] * l.stored[1];
    l.stride[3] = l.stride[2] * l.stored[2];
    l.stride[4] = l.stride[3] * nvars;
    *out = l;
  });
}

tah_status tah_layout_arena(const tah_schedule_desc* d, int32_t nvars, tah_layout* out) {
  return guard([&] {
    if (!d || !out) fail("null argument");
    if (nvars < 1) fail("nvars must be positive");
    tah_layout l{};
    for (int k = 0; k < 3; ++k) l.stored[k] = d->stored[k];
    l.nvars = nvars;
    l.nblocks = d->nblocks;
    // Variables innermost: a point's nvars values are contiguous, so every
    // ghost slab row, however short along x, is one run of at least nvars
    // doubles; blocks start on 256-byte boundaries, in Morton order.
    l.stride[3] = 1;
    l.stride[0] = nvars;
    l.stride[1] = l.stride[0] * l.stored[0];
    l.stride[2] = l.stride[1] * l.stored[1];
    l.stride[4] = align_up(l.stride[2] * l.stored[2], 32);
    *out = l;
  });
}

tah_status tah_layout_size(const tah_layout* l, int64_t* elements) {
  return guard([&] {
    if (!l || !elements) fail("null argument");
    validate_strides(l);
    *elements = l->stride[4] * l->nblocks;
  });
}

tah_status tah_ctx_create(int32_t device, tah_ctx** out) {
  return guard([&] {
    if (!out) fail("null argument");
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) fail("no HIP device " + std::to_string(device));
    auto c = std::make_unique<tah_ctx>();
    c->device = device;
    HIPCHECK(hipGetDeviceProperties(&c->prop, device));
    HIPCHECK(hipSetDevice(device));
    noalloc_scratch_init(device);
    *out = c.release();
  });
}

tah_status tah_ctx_destroy(tah_ctx* ctx) {
  if (ctx) batch_forget_ctx(ctx);
  delete ctx;
  return TAH_OK;
}

tah_status tah_plan_create(tah_ctx* ctx, const tah_schedule_desc* d, const tah_layout* l,
                           tah_plan** out) {
  return guard([&] {
    if (!ctx || !out) fail("null argument");
    validate_layout(d, l);
    HIPCHECK(hipSetDevice(ctx->device));
    auto p = std::make_unique<tah_plan>();
    p->ctx = ctx;
    p->layout = *l;
    p->dim = d->dim;
    p->nstages = d->nstages;
    p->tags.assign(d->stage_tags, d->stage_tags + d->nstages);
    for (int32_t k = 0; k < d->nstagegroups; ++k) {
      int32_t a = d->stagegroup_first[k], b = d->stagegroup_last[k];
      if (a < 0 || b < a || b >= d->nstages) fail("stage groups out of range");
      if (k > 0 && a != p->stagegroups.back()[1] + 1) fail("stage groups must tile the stages in order");
      p->stagegroups.push_back({a, b});
    }
    if (d->nstages > 0 && (p->stagegroups.empty() || p->stagegroups.back()[1] != d->nstages - 1))
      fail("stage groups must cover every stage");
    // A regrid transfer: stage REGRID (sources: the old leaves), then at
    // most stage REGRID_KEPT, each a stage group of its own.
    p->regrid = d->nstages > 0 && d->stage_tags[0] == TAH_STAGE_REGRID;
    if (p->regrid) {
      if (d->nsources < 1) fail("a regrid schedule needs the old leaf count (nsources)");
      if (d->nstages > 2 || (d->nstages == 2 && d->stage_tags[1] != TAH_STAGE_REGRID_KEPT) ||
          int(p->stagegroups.size()) != d->nstages)
        fail("a regrid schedule has stage REGRID, then at most REGRID_KEPT, each in a stage group of its own");
      p->nsources = d->nsources;
    } else {
      for (int32_t k = 0; k < d->nstages; ++k)
        if (d->stage_tags[k] == TAH_STAGE_REGRID || d->stage_tags[k] == TAH_STAGE_REGRID_KEPT)
          fail("regrid stages out of place");
      if (d->nsources != 0 && d->nsources != d->nblocks)
        fail("only a regrid schedule's sources index another block count");
    }
    int ncols = 1;
    for (int k = 0; k < d->dim; ++k) ncols *= 3;
    int prev_stage = -1;
    // --- validate and keep host copies ---------------------------------
    for (int32_t gi = 0; gi < d->ngroups; ++gi) {
      const tah_group_desc& g = d->groups[gi];
      PlanGroup h;
      std::string where = "group " + std::to_string(gi) + ": ";
      if (g.kind < 1 || g.kind > 3) fail(where + "unknown kind");
      if (g.stage < 0 || g.stage >= d->nstages) fail(where + "stage out of range");
      if (g.stage < prev_stage) fail("groups must be sorted by stage");
      prev_stage = g.stage;
      if (g.factorcol < 0 || g.factorcol > ncols) fail(where + "factor column out of range");
      if (g.ntransfers < 0) fail(where + "negative transfer count");
      h.kind = g.kind;
      h.factorcol = g.factorcol;
      h.stage = g.stage;
      p->needs_factors = p->needs_factors || g.factorcol != 0;
      h.tblocks.assign(g.targetblocks, g.targetblocks + g.ntransfers);
      h.sblocks.assign(g.sourceblocks, g.sourceblocks + g.ntransfers);
      const int32_t nsrc = p->regrid && g.stage == 0 ? d->nsources : d->nblocks;
      for (int32_t t = 0; t < g.ntransfers; ++t) {
        if (h.tblocks[t] < 0 || h.tblocks[t] >= d->nblocks || h.sblocks[t] < 0 || h.sblocks[t] >= nsrc)
          fail(where + "block index out of range");
        if (t > 0 && h.tblocks[t] < h.tblocks[t - 1]) fail(where + "targetblocks must be non-decreasing");
      }
      bool unit = true;
      int64_t pts = 1;
      for (int k = 0; k < 3; ++k) {
        const tah_stencil1d& s = g.stencils[k];
        if (s.ntarget < 0 || s.order < 1) fail(where + "bad stencil");
        if (s.targetfirst < 0 || s.targetfirst + s.ntarget > l->stored[k])
          fail(where + "target rows outside the stored extent");
        h.srcstart[k].assign(s.srcstart, s.srcstart + s.ntarget);
        h.weights[k].assign(s.weights, s.weights + int64_t(s.order) * s.ntarget);
        for (int32_t r = 0; r < s.ntarget; ++r)
          if (h.srcstart[k][r] < 0 || h.srcstart[k][r] + s.order > l->stored[k])
            fail(where + "source window outside the stored extent");
        bool u = s.order == 1;
        for (double w : h.weights[k]) u = u && w == 1.0;
        unit = unit && u;
        h.n[k] = s.ntarget;
        h.order[k] = s.order;
        h.tfirst[k] = s.targetfirst;
        pts *= s.ntarget;
      }
      if (pts * l->nvars >= (int64_t(1) << 31)) fail(where + "a transfer box too large", kUnsupported);
      h.unit = unit;
      h.points = int32_t(pts);
      p->values += pts * g.ntransfers * l->nvars;
      p->groups.push_back(std::move(h));
    }
    // --- the outer-face hook regions -------------------------------------
    if (d->nboundaries < 0 || (d->nboundaries > 0 && !d->boundaries)) fail("bad hook region list");
    if (p->regrid && d->nboundaries > 0) fail("a regrid schedule has no hook regions");
    int64_t first = 0;
    for (int32_t r = 0; r < d->nboundaries; ++r) {
      const tah_boundary_region& B = d->boundaries[r];
      const std::string where = "hook region " + std::to_string(r) + ": ";
      if (B.block < 0 || B.block >= d->nblocks) fail(where + "block index out of range");
      HookRegion h{};
      h.block = B.block;
      int64_t n = 1;
      for (int k = 0; k < 3; ++k) {
        if (B.lo[k] < 0 || B.hi[k] < B.lo[k] || B.hi[k] > l->stored[k]) fail(where + "box outside the stored extent");
        if (B.clamps[k] < -1 || B.clamps[k] >= l->stored[k])
          fail(where + "clamped index outside the stored extent");
        h.lo[k] = B.lo[k];
        h.n[k] = B.hi[k] - B.lo[k];
        h.clamp[k] = B.clamps[k];
        n *= h.n[k];
      }
      if (B.first != first) fail(where + "its first point does not follow the previous region's");
      if (n * l->nvars >= (int64_t(1) << 31)) fail(where + "too large", kUnsupported);
      h.first = first;
      first += n;
      p->hooks.push_back(h);
      p->hook_slots.push_back(B.slot);
    }
    if (first != d->boundary_points) fail("the hook regions' points do not add up to boundary_points");
    p->hook_points = first;
    build_plan_device(p.get());
    *out = p.release();
  });
}

tah_status tah_plan_destroy(tah_plan* p) {
  delete p;
  return TAH_OK;
}

tah_status tah_plan_values(const tah_plan* p, int64_t* values) {
  return guard([&] {
    if (!p || !values) fail("null argument");
    *values = p->values;
  });
}

tah_status tah_fill_plain(tah_plan* p, double* work, const double* factors, void* stream) {
  return guard([&] {
    if (!p || !work) fail("null argument");
    if (p->regrid) fail("this plan is a regrid transfer's: run it with tah_regrid_transfer");
    if (p->needs_factors && !factors)
      fail("this schedule mirrors ghosts across a reflecting face, but no factor table was given");
    HIPCHECK(hipSetDevice(p->ctx->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const tah_layout& L = p->layout;
    FastDiv dv(uint32_t(L.nvars));
    for (int32_t stage = 0; stage < p->nstages; ++stage) {
      for (size_t gi = 0; gi < p->groups.size(); ++gi) {
        const PlanGroup& h = p->groups[gi];
        if (h.stage != stage) continue;
        const uint64_t per = uint64_t(h.points) * L.nvars;  // values per transfer (< 2^31)
        const int64_t nt = int64_t(h.tblocks.size());
        if (per == 0 || nt == 0) continue;
        // The flat index is 32-bit: a group of 2^31 values or more runs as
        // several launches over runs of its transfers, where the reference
        // switches its inverses to Int64 (src/threading.jl:240-267).
        const int64_t chunk = std::max<int64_t>(1, int64_t((uint64_t(1) << 31) / per));
        for (int64_t t0 = 0; t0 < nt; t0 += chunk) {
          const uint64_t total = per * uint64_t(std::min(chunk, nt - t0));
          const uint32_t nb = uint32_t((total + 255) / 256);
          hipLaunchKernelGGL(plain_kernel, dim3(nb), dim3(256), 0, s, p->devgroups[gi], work, factors,
                             L.stride[0], L.stride[1], L.stride[2], L.stride[3], L.stride[4], L.nvars,
                             dv, uint32_t(total), int32_t(t0));
          HIPCHECK(hipGetLastError());
        }
      }
    }
  });
}

tah_status tah_regrid_transfer(tah_plan* p, const tah_layout* old_layout, const double* old_work,
                               double* new_work, void* stream) {
  return guard([&] {
    if (!p || !old_layout || !old_work || !new_work) fail("null argument");
    if (!p->regrid) fail("this plan is a fill's: run it with tah_fill");
    const tah_layout& L = p->layout;
    if (old_layout->nblocks != p->nsources)
      fail("the old array holds " + std::to_string(old_layout->nblocks) + " blocks, the regrid's sources " +
               std::to_string(p->nsources),
           kLayout);
    bool same = old_layout->nvars == L.nvars;
    for (int k = 0; k < 3; ++k) same = same && old_layout->stored[k] == L.stored[k];
    for (int k = 0; k < 5; ++k) same = same && old_layout->stride[k] == L.stride[k];
    if (!same) fail("the old and the new array must share one layout per block", kLayout);
    HIPCHECK(hipSetDevice(p->ctx->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // Stage REGRID reads the old array; stage REGRID_KEPT the new one.
    const Job j0 = plan_job(p, new_work, nullptr, old_work), j1 = plan_job(p, new_work, nullptr);
    plan_begin(p, s);
    plan_upload_job(p, j0, s);
    if (std::memcmp(&j1, &p->last_job2, sizeof(Job)) != 0) {
      p->last_job2 = j1;
      HIPCHECK(hipMemcpyAsync(p->djob2, &p->last_job2, sizeof(Job), hipMemcpyHostToDevice, s));
    }
    if (p->events.empty())
      for (int i = 0; i < 8; ++i) {
        hipEvent_t e;
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        p->events.push_back(e);
      }
    for (size_t k = 0; k < p->launches.size(); ++k)
      run_launches(p->ctx, p->launches[k], s, k == 0 ? p->djob : p->djob2, p->dtiles, p->dlds_tiles, p->events.data(),
                   p->fork != 0);
    plan_end(p, s);
  });
}

tah_status tah_fill(tah_plan* p, double* work, const double* factors, void* stream) {
  return guard([&] {
    if (!p || !work) fail("null argument");
    if (p->regrid) fail("this plan is a regrid transfer's: run it with tah_regrid_transfer");
    if (p->needs_factors && !factors)
      fail("this schedule mirrors ghosts across a reflecting face, but no factor table was given");
    HIPCHECK(hipSetDevice(p->ctx->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    plan_begin(p, s);
    plan_upload_job(p, plan_job(p, work, factors), s);
    if (p->events.empty())
      for (int i = 0; i < 8; ++i) {
        hipEvent_t e;
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        p->events.push_back(e);
      }
    for (const auto& sg : p->launches)
      run_launches(p->ctx, sg, s, p->djob, p->dtiles, p->dlds_tiles, p->events.data(), p->fork != 0);
    plan_end(p, s);
  });
}

tah_status tah_fill_boundary(tah_plan* p, double* work, const double* factors, int32_t kind, const double* values,
                             const double* fr, void* stream) {
  return guard([&] {
    if (!p || !work) fail("null argument");
    if (p->regrid) fail("this plan is a regrid transfer's: run it with tah_regrid_transfer");
    if (kind != TAH_HOOK_VALUES && kind != TAH_HOOK_RELATIVE)
      fail("the hook kind is TAH_HOOK_VALUES or TAH_HOOK_RELATIVE");
    if (p->needs_factors && !factors)
      fail("this schedule mirrors ghosts across a reflecting face, but no factor table was given");
    if (p->hook_points > 0 && (!values || (kind == TAH_HOOK_RELATIVE && !fr)))
      fail(kind == TAH_HOOK_RELATIVE ? "the relative hook needs f at the ghosts (values) and at their nearest "
                                       "evolved points (fr)"
                                     : "the hook's values are missing");
    if (p->launches.size() != p->stagegroups.size()) fail("plan launches do not match its stage groups", kHip);
    HIPCHECK(hipSetDevice(p->ctx->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    plan_begin(p, s);
    plan_upload_job(p, plan_job(p, work, factors), s);
    if (p->events.empty())
      for (int i = 0; i < 8; ++i) {
        hipEvent_t e;
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        p->events.push_back(e);
      }
    auto hook = [&](int64_t first, int64_t count, int32_t relative) {
      if (count == 0) return;
      hipLaunchKernelGGL(hook_kernel, dim3(uint32_t(count)), dim3(256), 0, s, p->djob, p->dhooks,
                         p->dhook_tiles + first, values, fr, relative);
      HIPCHECK(hipGetLastError());
    };
    for (size_t k = 0; k < p->launches.size(); ++k) {
      run_launches(p->ctx, p->launches[k], s, p->djob, p->dtiles, p->dlds_tiles, p->events.data(), p->fork != 0);
      if (kind == TAH_HOOK_VALUES && int(k) == p->phase1_sg) hook(p->hook_all[0], p->hook_all[1], 0);
      if (kind == TAH_HOOK_RELATIVE) hook(p->hook_after[k][0], p->hook_after[k][1], 1);
    }
    plan_end(p, s);
  });
}

tah_status tah_fill_timed(tah_plan* p, double* work, const double* factors, void* stream, int32_t* n,
                          int32_t* cls, int64_t* values, float* ms) {
  return guard([&] {
    if (!p || !n) fail("null argument");
    if (p->regrid) fail("this plan is a regrid transfer's: run it with tah_regrid_transfer");
    int32_t total = 0;
    for (const auto& sg : p->launches) total += int32_t(sg.size());
    if (!cls && !values && !ms) {
      *n = total;
      return;
    }
    if (!work || !cls || !values || !ms) fail("null argument");
    if (*n < total) fail("the record arrays hold " + std::to_string(*n) + " launches; the fill makes " +
                         std::to_string(total));
    if (p->needs_factors && !factors)
      fail("this schedule mirrors ghosts across a reflecting face, but no factor table was given");
    HIPCHECK(hipSetDevice(p->ctx->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    plan_begin(p, s);
    plan_upload_job(p, plan_job(p, work, factors), s);
    std::vector<hipEvent_t> ev(size_t(total) + 1);
    for (auto& e : ev) HIPCHECK(hipEventCreate(&e));
    try {
      int32_t k = 0;
      HIPCHECK(hipEventRecord(ev[0], s));
      for (const auto& sg : p->launches)
        for (const Launch& L : sg) {
          Launch serial = L;  // timed one by one: the copies' bounded grid is not taken
          serial.grid = 0;
          launch_one(serial, s, p->djob, p->dtiles, p->dlds_tiles);
          HIPCHECK(hipEventRecord(ev[k + 1], s));
          cls[k] = L.cls;
          values[k] = L.values;
          ++k;
        }
      plan_end(p, s);
      HIPCHECK(hipEventSynchronize(ev[total]));
      for (int32_t i = 0; i < total; ++i) HIPCHECK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
    } catch (...) {
      for (auto& e : ev) (void)hipEventDestroy(e);
      throw;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    *n = total;
  });
}

tah_status tah_kernel_class_name(int32_t cls, char* buf, int32_t len) {
  return guard([&] {
    if (!buf || len < 1) fail("null argument");
    std::string name;
    if (cls == kClsGeneric) {
      name = "generic_kernel";
    } else if (cls > kClsCopyLaunch && cls <= kClsCopyLaunch + 16) {
      name = "copy_kernel<" + std::to_string(cls - kClsCopyLaunch) + ">";
    } else if (cls >= kClsMergedPair && cls - kClsMergedPair < int(kInterpTable.size())) {
      const auto& e = kInterpTable[cls - kClsMergedPair];
      name = "interp_pair_kernel<" + std::to_string(e[0]) + ", " + std::to_string(e[1]) + ">";
    } else if (cls >= kClsStream && cls - kClsStream < int(kInterpTable.size())) {
      const auto& e = kInterpTable[cls - kClsStream];
      name = "interp_stream_kernel<" + std::to_string(e[0]) + ", " + std::to_string(stream_ka(e[0], e[1])) + ">";
    } else if (cls >= kClsLdsProlong && cls - kClsLdsProlong < int(kInterpTable.size())) {
      name = "interp_lds_prolong_kernel<" + std::to_string(kInterpTable[cls - kClsLdsProlong][0]) + ">";
    } else if (cls >= kClsLds && cls - kClsLds < int(kInterpTable.size())) {
      name = "interp_lds_kernel<" + std::to_string(kInterpTable[cls - kClsLds][0]) + ">";
    } else if (cls >= kClsMerged && cls - kClsMerged < int(kInterpTable.size())) {
      const auto& e = kInterpTable[cls - kClsMerged];
      name = "interp_all_kernel<" + std::to_string(e[0]) + ", " + std::to_string(e[1]) + ">";
    } else if (cls >= kClsInterp && (cls - kClsInterp) / 8 < int(kInterpTable.size())) {
      const int id = cls - kClsInterp, rb = id % 8;
      const auto& e = kInterpTable[id / 8];
      name = "interp_kernel<" + std::to_string(e[0]) + ", " + std::to_string(e[1]) + ", " +
             std::to_string(1 + (rb & 1)) + ", " + std::to_string(1 + ((rb >> 1) & 1)) + ", " +
             std::to_string(1 + ((rb >> 2) & 1)) + ">";
    } else {
      fail("no kernel class " + std::to_string(cls));
    }
    std::snprintf(buf, size_t(len), "%s", name.c_str());
  });
}

}  // extern "C"
