// Device side of the ghost fill: plans (the schedule's tables in HBM), the
// plain per-group kernel, the staged tile kernel, and their C ABI.
//
// Every value is computed exactly as the reference's transfer_kernel!
// computes it (src/ghosts.jl:372-402): a unit copy is 0 + u (copy_load,
// :127-128); a stencil is one accumulator from +0 with m1 innermost and the
// tap weight ((w1*w2)*w3) (stencil_sum, :58-118), every product and sum
// rounded on its own (explicit _rn intrinsics; the file is also compiled
// with -ffp-contract=off); a factor is acc*f + 0 (scaled, :369-370).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "api_common.hpp"
#include "device.hpp"
#include "forest.hpp"

using namespace tah;

#define HIPCHECK(expr)                                                                    \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw Error(kHip, std::string(#expr) + " failed: " + hipGetErrorString(e_));        \
  } while (0)

namespace tah {

// ---------------------------------------------------------------------------
// The value of one target point of one transfer.

__device__ __forceinline__ double scaled(double acc, const double* __restrict__ factors,
                                         int32_t factorcol, int32_t nvars, int32_t v) {
  if (factorcol == 0) return acc;
  double f = factors[v + int64_t(nvars) * (factorcol - 1)];
  return __dadd_rn(__dmul_rn(acc, f), 0.0);
}

__device__ __forceinline__ double stencil_value(const DevGroup& g, const double* __restrict__ sb,
                                                int64_t s1, int64_t s2, int32_t i, int32_t j,
                                                int32_t k) {
  const int32_t b0 = g.srcstart[0][i], b1 = g.srcstart[1][j], b2 = g.srcstart[2][k];
  if (g.unit) return __dadd_rn(0.0, sb[b0 + s1 * b1 + s2 * b2]);
  const int32_t o0 = g.order[0], o1 = g.order[1], o2 = g.order[2];
  const double* __restrict__ w0 = g.weights[0] + int64_t(o0) * i;
  const double* __restrict__ w1 = g.weights[1] + int64_t(o1) * j;
  const double* __restrict__ w2 = g.weights[2] + int64_t(o2) * k;
  double acc = 0.0;
  for (int32_t m2 = 0; m2 < o2; ++m2) {
    const double c2 = w2[m2];
    const double* __restrict__ row2 = sb + s2 * (b2 + m2) + b0;
    for (int32_t m1 = 0; m1 < o1; ++m1) {
      const double c1 = w1[m1];
      const double* __restrict__ row = row2 + s1 * (b1 + m1);
      for (int32_t m0 = 0; m0 < o0; ++m0) {
        double w = __dmul_rn(__dmul_rn(w0[m0], c1), c2);
        acc = __dadd_rn(acc, __dmul_rn(w, row[m0]));
      }
    }
  }
  return acc;
}

// ---------------------------------------------------------------------------
// The plain fill: one launch per group, one thread per written value, the
// flat index split into (i, j, k, v, t) with i fastest (launch_transfer!'s
// flat form, src/ghosts.jl:507-521).

__global__ void __launch_bounds__(256) plain_kernel(DevGroup g, double* __restrict__ work,
                                                    const double* __restrict__ factors, int64_t s1,
                                                    int64_t s2, int64_t s3, int64_t s4,
                                                    int32_t nvars, FastDiv dv, uint32_t total) {
  uint32_t e = blockIdx.x * 256u + threadIdx.x;
  if (e >= total) return;
  uint32_t r = g.dn[0].div(e);
  int32_t i = int32_t(e - r * uint32_t(g.n[0]));
  uint32_t r2 = g.dn[1].div(r);
  int32_t j = int32_t(r - r2 * uint32_t(g.n[1]));
  uint32_t r3 = g.dn[2].div(r2);
  int32_t k = int32_t(r2 - r3 * uint32_t(g.n[2]));
  uint32_t t = dv.div(r3);
  int32_t v = int32_t(r3 - t * uint32_t(nvars));
  const double* sb = work + s4 * g.sblocks[t] + s3 * v;
  double acc = stencil_value(g, sb, s1, s2, i, j, k);
  acc = scaled(acc, factors, g.factorcol, nvars, v);
  double* tb = work + s4 * g.tblocks[t] + s3 * v;
  tb[(g.tfirst[0] + i) + s1 * (g.tfirst[1] + j) + s2 * (g.tfirst[2] + k)] = acc;
}

// ---------------------------------------------------------------------------
// The staged fill: one launch per stage group over a list of tiles drawn from
// every group (and every job) of the stage, so that small groups no longer
// cost a launch each. A work-group takes one tile of kTileElems values of one
// transfer; the group descriptor is uniform across it.

constexpr int kThreads = 256;
constexpr int kPerThread = 4;
constexpr int kTileElems = kThreads * kPerThread;

__global__ void __launch_bounds__(kThreads) staged_kernel(const Job* __restrict__ jobs,
                                                          const Tile* __restrict__ tiles) {
  const Tile tile = tiles[blockIdx.x];
  const Job& job = jobs[tile.job];
  const DevGroup& g = job.groups[tile.group];
  const int64_t s1 = job.stride[1], s2 = job.stride[2], s3 = job.stride[3], s4 = job.stride[4];
  const int32_t nvars = job.nvars;
  const uint32_t total = uint32_t(g.points) * uint32_t(nvars);
  const double* sbase = job.work + s4 * g.sblocks[tile.transfer];
  double* tbase = job.work + s4 * g.tblocks[tile.transfer];
#pragma unroll
  for (int r = 0; r < kPerThread; ++r) {
    uint32_t e = uint32_t(tile.e0) + uint32_t(r * kThreads) + threadIdx.x;
    if (e >= total) break;
    uint32_t q = g.dn[0].div(e);
    int32_t i = int32_t(e - q * uint32_t(g.n[0]));
    uint32_t q2 = g.dn[1].div(q);
    int32_t j = int32_t(q - q2 * uint32_t(g.n[1]));
    uint32_t v32 = g.dn[2].div(q2);
    int32_t k = int32_t(q2 - v32 * uint32_t(g.n[2]));
    int32_t v = int32_t(v32);
    double acc = stencil_value(g, sbase + s3 * v, s1, s2, i, j, k);
    acc = scaled(acc, job.factors, g.factorcol, nvars, v);
    tbase[s3 * v + (g.tfirst[0] + i) + s1 * (g.tfirst[1] + j) + s2 * (g.tfirst[2] + k)] = acc;
  }
}

}  // namespace tah

// ---------------------------------------------------------------------------
// Plans.

struct tah_ctx {
  int device = 0;
  hipDeviceProp_t prop{};
};

struct HostGroup {
  DevGroup dev;  // pointers filled once the device buffer exists
  int32_t kind, factorcol, ntransfers, unit;
  int32_t n[3], order[3], tfirst[3];
  int32_t stage;
  int64_t off_tblocks, off_sblocks, off_srcstart[3], off_weights[3];  // byte offsets
};

struct tah_plan {
  tah_ctx* ctx = nullptr;
  tah_layout layout{};
  int32_t dim = 3;
  int32_t nstages = 0;
  std::vector<int32_t> tags;
  std::vector<std::array<int32_t, 2>> stagegroups;
  std::vector<HostGroup> groups;
  // Host copies kept for the adjoint's inverted schedule and for tiles.
  std::vector<std::vector<int32_t>> tblocks, sblocks;
  std::vector<std::array<std::vector<int32_t>, 3>> srcstart;
  std::vector<std::array<std::vector<double>, 3>> weights;
  bool needs_factors = false;
  int64_t values = 0;
  // Device memory: tables + DevGroup array + tiles of each stage group.
  char* dtables = nullptr;
  DevGroup* dgroups = nullptr;
  Tile* dtiles = nullptr;
  Job* djob = nullptr;
  std::vector<int64_t> tile_first;  // per stage group, into dtiles; size nsg+1
  ~tah_plan() {
    if (dtables) (void)hipFree(dtables);
    if (dtiles) (void)hipFree(dtiles);
    if (djob) (void)hipFree(djob);
  }
};

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
  if (l->stride[0] != 1) fail("stride[0] must be 1", kLayout);
  if (l->stride[1] < l->stored[0] || l->stride[2] < l->stride[1] * l->stored[1] ||
      l->stride[3] < l->stride[2] * l->stored[2] || l->stride[4] < l->stride[3] * l->nvars)
    fail("the layout's strides overlap", kLayout);
  if (int64_t(l->stored[0]) * l->stored[1] * l->stored[2] * l->nvars > (int64_t(1) << 31))
    fail("a block larger than 2^31 values is not supported", kUnsupported);
}

}  // namespace

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
    l.stride[0] = 1;
    // Rows padded to a multiple of four doubles (32 bytes), planes and
    // blocks to 256 bytes, so that every row starts on a sector.
    l.stride[1] = d->dim >= 2 ? align_up(l.stored[0], 4) : l.stored[0];
    l.stride[2] = l.stride[1] * l.stored[1];
    l.stride[3] = align_up(l.stride[2] * l.stored[2], 32);
    l.stride[4] = l.stride[3] * nvars;
    *out = l;
  });
}

tah_status tah_layout_size(const tah_layout* l, int64_t* elements) {
  return guard([&] {
    if (!l || !elements) fail("null argument");
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
    *out = c.release();
  });
}

tah_status tah_ctx_destroy(tah_ctx* ctx) {
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
    int ncols = 1;
    for (int k = 0; k < d->dim; ++k) ncols *= 3;
    int64_t bytes = 0;
    int prev_stage = -1;
    for (int32_t gi = 0; gi < d->ngroups; ++gi) {
      const tah_group_desc& g = d->groups[gi];
      HostGroup h{};
      if (g.kind < 1 || g.kind > 3) fail("group " + std::to_string(gi) + ": unknown kind");
      if (g.stage < 0 || g.stage >= d->nstages) fail("group " + std::to_string(gi) + ": stage out of range");
      if (g.stage < prev_stage) fail("groups must be sorted by stage");
      prev_stage = g.stage;
      if (g.factorcol < 0 || g.factorcol > ncols) fail("group " + std::to_string(gi) + ": factor column out of range");
      if (g.ntransfers < 0) fail("negative transfer count");
      h.kind = g.kind;
      h.factorcol = g.factorcol;
      h.ntransfers = g.ntransfers;
      h.stage = g.stage;
      p->needs_factors = p->needs_factors || g.factorcol != 0;
      std::vector<int32_t> tb(g.targetblocks, g.targetblocks + g.ntransfers);
      std::vector<int32_t> sb(g.sourceblocks, g.sourceblocks + g.ntransfers);
      for (int32_t t = 0; t < g.ntransfers; ++t) {
        if (tb[t] < 0 || tb[t] >= d->nblocks || sb[t] < 0 || sb[t] >= d->nblocks)
          fail("group " + std::to_string(gi) + ": block index out of range");
        if (t > 0 && tb[t] < tb[t - 1]) fail("group " + std::to_string(gi) + ": targetblocks must be non-decreasing");
      }
      int unit = 1;
      int64_t pts = 1;
      std::array<std::vector<int32_t>, 3> ss;
      std::array<std::vector<double>, 3> ww;
      for (int k = 0; k < 3; ++k) {
        const tah_stencil1d& s = g.stencils[k];
        if (s.ntarget < 0 || s.order < 1) fail("group " + std::to_string(gi) + ": bad stencil");
        if (s.targetfirst < 0 || s.targetfirst + s.ntarget > l->stored[k])
          fail("group " + std::to_string(gi) + ": target rows outside the stored extent");
        ss[k].assign(s.srcstart, s.srcstart + s.ntarget);
        ww[k].assign(s.weights, s.weights + int64_t(s.order) * s.ntarget);
        for (int32_t r = 0; r < s.ntarget; ++r)
          if (ss[k][r] < 0 || ss[k][r] + s.order > l->stored[k])
            fail("group " + std::to_string(gi) + ": source window outside the stored extent");
        bool u = s.order == 1;
        for (double w : ww[k]) u = u && w == 1.0;
        unit = unit && u;
        h.n[k] = s.ntarget;
        h.order[k] = s.order;
        h.tfirst[k] = s.targetfirst;
        pts *= s.ntarget;
      }
      if (pts * l->nvars >= (int64_t(1) << 31)) fail("a transfer box too large", kUnsupported);
      h.unit = unit;
      p->values += pts * g.ntransfers * l->nvars;
      h.off_tblocks = bytes; bytes = align_up(bytes + 4 * int64_t(g.ntransfers), 256);
      h.off_sblocks = bytes; bytes = align_up(bytes + 4 * int64_t(g.ntransfers), 256);
      for (int k = 0; k < 3; ++k) {
        h.off_srcstart[k] = bytes; bytes = align_up(bytes + 4 * int64_t(ss[k].size()), 256);
        h.off_weights[k] = bytes; bytes = align_up(bytes + 8 * int64_t(ww[k].size()), 256);
      }
      p->groups.push_back(h);
      p->tblocks.push_back(std::move(tb));
      p->sblocks.push_back(std::move(sb));
      p->srcstart.push_back(std::move(ss));
      p->weights.push_back(std::move(ww));
    }
    int64_t off_groups = bytes;
    bytes = align_up(bytes + int64_t(sizeof(DevGroup)) * std::max<size_t>(1, p->groups.size()), 256);
    std::vector<char> host(bytes, 0);
    HIPCHECK(hipMalloc(&p->dtables, bytes));
    std::vector<DevGroup> dg;
    for (size_t gi = 0; gi < p->groups.size(); ++gi) {
      HostGroup& h = p->groups[gi];
      std::memcpy(host.data() + h.off_tblocks, p->tblocks[gi].data(), 4 * p->tblocks[gi].size());
      std::memcpy(host.data() + h.off_sblocks, p->sblocks[gi].data(), 4 * p->sblocks[gi].size());
      DevGroup g{};
      g.kind = h.kind;
      g.factorcol = h.factorcol;
      g.ntransfers = h.ntransfers;
      g.unit = h.unit;
      g.points = h.n[0] * h.n[1] * h.n[2];
      g.tblocks = reinterpret_cast<const int32_t*>(p->dtables + h.off_tblocks);
      g.sblocks = reinterpret_cast<const int32_t*>(p->dtables + h.off_sblocks);
      for (int k = 0; k < 3; ++k) {
        std::memcpy(host.data() + h.off_srcstart[k], p->srcstart[gi][k].data(), 4 * p->srcstart[gi][k].size());
        std::memcpy(host.data() + h.off_weights[k], p->weights[gi][k].data(), 8 * p->weights[gi][k].size());
        g.n[k] = h.n[k];
        g.order[k] = h.order[k];
        g.tfirst[k] = h.tfirst[k];
        g.dn[k] = FastDiv(uint32_t(std::max(1, h.n[k])));
        g.srcstart[k] = reinterpret_cast<const int32_t*>(p->dtables + h.off_srcstart[k]);
        g.weights[k] = reinterpret_cast<const double*>(p->dtables + h.off_weights[k]);
      }
      h.dev = g;
      dg.push_back(g);
    }
    if (!dg.empty()) std::memcpy(host.data() + off_groups, dg.data(), sizeof(DevGroup) * dg.size());
    p->dgroups = reinterpret_cast<DevGroup*>(p->dtables + off_groups);
    HIPCHECK(hipMemcpy(p->dtables, host.data(), bytes, hipMemcpyHostToDevice));

    // Tiles of each stage group, in group order.
    std::vector<Tile> tiles;
    p->tile_first.push_back(0);
    for (auto& sg : p->stagegroups) {
      for (size_t gi = 0; gi < p->groups.size(); ++gi) {
        const HostGroup& h = p->groups[gi];
        if (h.stage < sg[0] || h.stage > sg[1]) continue;
        int64_t elems = int64_t(h.dev.points) * l->nvars;
        if (elems == 0) continue;
        for (int32_t t = 0; t < h.ntransfers; ++t)
          for (int64_t e = 0; e < elems; e += kTileElems)
            tiles.push_back(Tile{0, int32_t(gi), t, int32_t(e)});
      }
      p->tile_first.push_back(int64_t(tiles.size()));
    }
    if (!tiles.empty()) {
      HIPCHECK(hipMalloc(&p->dtiles, sizeof(Tile) * tiles.size()));
      HIPCHECK(hipMemcpy(p->dtiles, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice));
    }
    HIPCHECK(hipMalloc(&p->djob, sizeof(Job)));
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
    if (p->needs_factors && !factors)
      fail("this schedule mirrors ghosts across a reflecting face, but no factor table was given");
    HIPCHECK(hipSetDevice(p->ctx->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const tah_layout& L = p->layout;
    FastDiv dv(uint32_t(L.nvars));
    for (int32_t stage = 0; stage < p->nstages; ++stage) {
      for (const HostGroup& h : p->groups) {
        if (h.stage != stage) continue;
        uint64_t total = uint64_t(h.dev.points) * L.nvars * h.ntransfers;
        if (total == 0) continue;
        if (total >= (uint64_t(1) << 32)) fail("a group larger than 2^32 values", kUnsupported);
        uint32_t nb = uint32_t((total + 255) / 256);
        hipLaunchKernelGGL(plain_kernel, dim3(nb), dim3(256), 0, s, h.dev, work, factors,
                           L.stride[1], L.stride[2], L.stride[3], L.stride[4], L.nvars, dv,
                           uint32_t(total));
        HIPCHECK(hipGetLastError());
      }
    }
  });
}

tah_status tah_fill(tah_plan* p, double* work, const double* factors, void* stream) {
  return guard([&] {
    if (!p || !work) fail("null argument");
    if (p->needs_factors && !factors)
      fail("this schedule mirrors ghosts across a reflecting face, but no factor table was given");
    HIPCHECK(hipSetDevice(p->ctx->device));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    Job job{};
    job.groups = p->dgroups;
    job.work = work;
    job.factors = factors;
    for (int k = 0; k < 5; ++k) job.stride[k] = p->layout.stride[k];
    job.nvars = p->layout.nvars;
    HIPCHECK(hipMemcpyAsync(p->djob, &job, sizeof(Job), hipMemcpyHostToDevice, s));
    for (size_t k = 0; k + 1 < p->tile_first.size(); ++k) {
      int64_t a = p->tile_first[k], b = p->tile_first[k + 1];
      if (b == a) continue;
      hipLaunchKernelGGL(staged_kernel, dim3(uint32_t(b - a)), dim3(kThreads), 0, s, p->djob,
                         p->dtiles + a);
      HIPCHECK(hipGetLastError());
    }
  });
}

tah_status tah_fill_batched(tah_ctx* ctx, int32_t n, tah_plan* const* plans, double* const* works,
                            const double* const* factors, void* stream) {
  (void)ctx; (void)n; (void)plans; (void)works; (void)factors; (void)stream;
  set_last_error("tah_fill_batched is not implemented yet");
  return TAH_EUNSUPPORTED;
}

tah_status tah_graph_create(tah_ctx* ctx, int32_t n, tah_plan* const* plans, double* const* works,
                            const double* const* factors, tah_graph** out) {
  (void)ctx; (void)n; (void)plans; (void)works; (void)factors; (void)out;
  set_last_error("tah_graph_create is not implemented yet");
  return TAH_EUNSUPPORTED;
}

tah_status tah_graph_launch(tah_graph* g, void* stream) {
  (void)g; (void)stream;
  set_last_error("tah_graph_launch is not implemented yet");
  return TAH_EUNSUPPORTED;
}

tah_status tah_graph_destroy(tah_graph* g) {
  (void)g;
  return TAH_OK;
}

tah_status tah_fill_adjoint(tah_plan* p, double* adj, const double* factors, void* stream) {
  (void)p; (void)adj; (void)factors; (void)stream;
  set_last_error("tah_fill_adjoint is not implemented yet");
  return TAH_EUNSUPPORTED;
}

tah_status tah_tune(tah_plan* p, double* work, const double* factors, const char* cache_path) {
  (void)p; (void)work; (void)factors; (void)cache_path;
  set_last_error("tah_tune is not implemented yet");
  return TAH_EUNSUPPORTED;
}

}  // extern "C"
