// The adjoint (transpose) of the fill.
//
// The fill is linear and runs in stages, each overwriting its targets from
// sources it never writes. Its transpose therefore runs the stages in
// reverse: in each, every source point gathers w * f * adj[target] over the
// transfers that read it, and then every target of the stage is zeroed.
// The gather is deterministic and uses no float atomics: one workgroup owns
// a source block and visits its incoming transfers in a fixed order (group,
// then transfer), with a barrier between transfers, so that no two threads
// ever update one point at once; within a transfer a point sums its
// (target, tap) pairs in a fixed order from a host-built inverted stencil.
// The reference has no adjoint: this is new, pinned by the dot-product
// identity and by the CPU transpose of the oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "api_common.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace tah {

// One dimension of a group's transposed stencil: for source row x in the
// read box lo..hi, the (target row, tap) pairs reading it, packed r*16 + m.
struct alignas(16) DevGroupT {
  int32_t lo[3], ext[3];
  int32_t pts;
  int32_t pad;
  FastDiv dx[2];
  const int32_t* start[3];  // ext[d] + 1 offsets into list[d]
  const int32_t* list[3];
};

struct alignas(16) AdjItem {
  int32_t sblock, first, count, pad;
};

struct AdjointTables {
  char* dmem = nullptr;
  DevGroupT* dgt = nullptr;
  AdjItem* ditems = nullptr;
  int32_t* dentries = nullptr;  // (group, transfer) pairs
  Tile* dzero = nullptr;
  // per stage: item range and zero-tile range
  std::vector<std::array<int64_t, 4>> stage_ranges;
};

void free_adjoint(tah_plan* p) {
  if (!p->adj) return;
  if (p->adj->dmem) (void)hipFree(p->adj->dmem);
  delete p->adj;
  p->adj = nullptr;
}

__global__ void __launch_bounds__(kThreads) adjoint_gather_kernel(const Job* __restrict__ jobs,
                                                                  const DevGroupT* __restrict__ gts,
                                                                  const AdjItem* __restrict__ items,
                                                                  const int32_t* __restrict__ entries) {
  const AdjItem item = items[blockIdx.x];
  const Job& job = jobs[0];
  const int32_t s0 = int32_t(job.stride[0]), s1 = int32_t(job.stride[1]),
                s2 = int32_t(job.stride[2]), s3 = int32_t(job.stride[3]);
  const int64_t s4 = job.stride[4];
  const int32_t nvars = job.nvars;
  double* __restrict__ sb = job.work + s4 * item.sblock;
  for (int32_t e = 0; e < item.count; ++e) {
    const int32_t gi = entries[2 * (item.first + e)];
    const int32_t t = entries[2 * (item.first + e) + 1];
    const DevGroup& g = job.groups[gi];
    const DevGroupT& T = gts[gi];
    const double* __restrict__ tb = job.work + s4 * g.tblocks[t];
    const uint32_t n = uint32_t(T.pts) * uint32_t(nvars);
    for (uint32_t idx = threadIdx.x; idx < n; idx += kThreads) {
      uint32_t r;
      int32_t v;
      if (job.vfast) {
        r = job.dnv.div(idx);
        v = int32_t(idx - r * uint32_t(nvars));
      } else {
        v = int32_t(idx / uint32_t(T.pts));
        r = idx - uint32_t(v) * uint32_t(T.pts);
      }
      const uint32_t q = T.dx[0].div(r);
      const int32_t x = int32_t(r - q * uint32_t(T.ext[0]));
      const uint32_t zq = T.dx[1].div(q);
      const int32_t y = int32_t(q - zq * uint32_t(T.ext[1]));
      const int32_t z = int32_t(zq);
      double acc = 0.0;
      if (g.unit == 2) {
        // A copy: along a dimension with step +-1 one target row reads this
        // point, along a clamped one (step 0) every row does; rows in order.
        const int32_t p0 = T.lo[0] + x, p1 = T.lo[1] + y, p2 = T.lo[2] + z;
        int32_t a[3], b[3];
        const int32_t pd[3] = {p0, p1, p2};
        bool any = true;
        for (int d = 0; d < 3; ++d) {
          if (g.cs[d] == 0) {
            a[d] = 0;
            b[d] = g.n[d];
          } else {
            const int32_t r = (pd[d] - g.ca[d]) * g.cs[d];
            a[d] = r;
            b[d] = r + 1;
            any = any && r >= 0 && r < g.n[d];
          }
        }
        if (any)
          for (int32_t r2 = a[2]; r2 < b[2]; ++r2)
            for (int32_t r1 = a[1]; r1 < b[1]; ++r1)
              for (int32_t r0 = a[0]; r0 < b[0]; ++r0)
                acc = __dadd_rn(acc, tb[s3 * v + s0 * (g.tfirst[0] + r0) + s1 * (g.tfirst[1] + r1) +
                                       s2 * (g.tfirst[2] + r2)]);
      } else {
      for (int32_t j2 = T.start[2][z]; j2 < T.start[2][z + 1]; ++j2) {
        const int32_t p2 = T.list[2][j2], r2 = p2 >> 4, m2 = p2 & 15;
        const double w2 = g.weights[2][g.order[2] * r2 + m2];
        for (int32_t j1 = T.start[1][y]; j1 < T.start[1][y + 1]; ++j1) {
          const int32_t p1 = T.list[1][j1], r1 = p1 >> 4, m1 = p1 & 15;
          const double w1 = g.weights[1][g.order[1] * r1 + m1];
          for (int32_t j0 = T.start[0][x]; j0 < T.start[0][x + 1]; ++j0) {
            const int32_t p0 = T.list[0][j0], r0 = p0 >> 4, m0 = p0 & 15;
            const double w = __dmul_rn(__dmul_rn(g.weights[0][g.order[0] * r0 + m0], w1), w2);
            acc = __dadd_rn(acc, __dmul_rn(w, tb[s3 * v + s0 * (g.tfirst[0] + r0) +
                                                  s1 * (g.tfirst[1] + r1) + s2 * (g.tfirst[2] + r2)]));
          }
        }
      }
      }
      if (g.factorcol) acc = __dmul_rn(acc, job.factors[v + int64_t(nvars) * (g.factorcol - 1)]);
      double* dst = sb + (s3 * v + s0 * (T.lo[0] + x) + s1 * (T.lo[1] + y) + s2 * (T.lo[2] + z));
      *dst = __dadd_rn(*dst, acc);
    }
    __syncthreads();
  }
}

// Zero the target boxes of one stage (tiles as the copy kernel's).
__global__ void __launch_bounds__(kThreads) zero_kernel(const Job* __restrict__ jobs,
                                                        const Tile* __restrict__ tiles) {
  const Tile tile = tiles[blockIdx.x];
  const Job& job = jobs[0];
  const DevGroup& g = job.groups[tile.item];
  const int32_t s0 = int32_t(job.stride[0]), s1 = int32_t(job.stride[1]),
                s2 = int32_t(job.stride[2]), s3 = int32_t(job.stride[3]);
  const uint32_t total = uint32_t(g.points) * uint32_t(job.nvars);
  double* __restrict__ tbase = job.work + job.stride[4] * g.tblocks[tile.transfer];
  for (int r = 0; r < kCopyPerThread; ++r) {
    const uint32_t e = uint32_t(tile.e0) + uint32_t(r * kThreads) + threadIdx.x;
    if (e >= total) break;
    int32_t i, j, k, v;
    split_element(g, job, e, i, j, k, v);
    tbase[s3 * v + s0 * (g.tfirst[0] + i) + s1 * (g.tfirst[1] + j) + s2 * (g.tfirst[2] + k)] = 0.0;
  }
}

namespace {

void build_adjoint(tah_plan* p) {
  auto* A = new AdjointTables();
  p->adj = A;
  std::vector<char> host;
  auto put = [&](const void* src, int64_t bytes) {
    int64_t off = (int64_t(host.size()) + 255) / 256 * 256;
    host.resize(off + std::max<int64_t>(bytes, 8));
    if (bytes) std::memcpy(host.data() + off, src, bytes);
    return off;
  };
  struct GT {
    int32_t lo[3], ext[3];
    int64_t start[3], list[3];
  };
  std::vector<GT> gts(p->groups.size());
  for (size_t gi = 0; gi < p->groups.size(); ++gi) {
    const PlanGroup& g = p->groups[gi];
    GT& h = gts[gi];
    for (int d = 0; d < 3; ++d) {
      int lo = 1 << 30, hi = -1;
      for (int32_t s : g.srcstart[d]) {
        lo = std::min(lo, s);
        hi = std::max(hi, s + g.order[d] - 1);
      }
      if (g.srcstart[d].empty()) lo = hi = 0;
      h.lo[d] = lo;
      h.ext[d] = hi - lo + 1;
      if (g.order[d] > 16) fail("the adjoint supports stencils of at most 16 taps", kUnsupported);
      std::vector<std::vector<int32_t>> by(h.ext[d]);
      for (int32_t r = 0; r < int32_t(g.srcstart[d].size()); ++r)
        for (int m = 0; m < g.order[d]; ++m) by[g.srcstart[d][r] + m - lo].push_back(r * 16 + m);
      std::vector<int32_t> start{0}, list;
      for (auto& v : by) {
        list.insert(list.end(), v.begin(), v.end());
        start.push_back(int32_t(list.size()));
      }
      h.start[d] = put(start.data(), 4 * int64_t(start.size()));
      h.list[d] = put(list.data(), 4 * int64_t(list.size()));
    }
  }
  // Items and entries, stage by stage.
  std::vector<AdjItem> items;
  std::vector<int32_t> entries;
  std::vector<Tile> zeros;
  const int32_t nvars = p->layout.nvars;
  for (int32_t st = 0; st < p->nstages; ++st) {
    std::array<int64_t, 4> rng{int64_t(items.size()), 0, int64_t(zeros.size()), 0};
    std::vector<std::array<int32_t, 3>> es;  // sblock, group, transfer
    for (size_t gi = 0; gi < p->groups.size(); ++gi) {
      const PlanGroup& g = p->groups[gi];
      if (g.stage != st || g.points == 0) continue;
      for (int32_t t = 0; t < int32_t(g.tblocks.size()); ++t) {
        es.push_back({g.sblocks[t], int32_t(gi), t});
        int64_t elems = int64_t(g.points) * nvars;
        for (int64_t e = 0; e < elems; e += kCopyTile) zeros.push_back(Tile{0, int32_t(gi), t, int32_t(e), 0, 0, 0, 0});
      }
    }
    std::stable_sort(es.begin(), es.end(), [](const std::array<int32_t, 3>& a, const std::array<int32_t, 3>& b) {
      return a[0] < b[0];
    });
    for (size_t i = 0; i < es.size();) {
      size_t j = i;
      while (j < es.size() && es[j][0] == es[i][0]) ++j;
      items.push_back(AdjItem{es[i][0], int32_t(entries.size() / 2), int32_t(j - i), 0});
      for (size_t k = i; k < j; ++k) {
        entries.push_back(es[k][1]);
        entries.push_back(es[k][2]);
      }
      i = j;
    }
    rng[1] = int64_t(items.size()) - rng[0];
    rng[3] = int64_t(zeros.size()) - rng[2];
    A->stage_ranges.push_back(rng);
  }
  int64_t off_items = put(items.data(), int64_t(sizeof(AdjItem) * items.size()));
  int64_t off_entries = put(entries.data(), 4 * int64_t(entries.size()));
  int64_t off_zero = put(zeros.data(), int64_t(sizeof(Tile) * zeros.size()));
  int64_t off_gt = put(nullptr, int64_t(sizeof(DevGroupT) * std::max<size_t>(1, gts.size())));
  HIPCHECK(hipMalloc(&A->dmem, host.size()));
  char* base = A->dmem;
  std::vector<DevGroupT> dg(gts.size());
  for (size_t gi = 0; gi < gts.size(); ++gi) {
    DevGroupT d{};
    for (int k = 0; k < 3; ++k) {
      d.lo[k] = gts[gi].lo[k];
      d.ext[k] = gts[gi].ext[k];
      d.start[k] = reinterpret_cast<const int32_t*>(base + gts[gi].start[k]);
      d.list[k] = reinterpret_cast<const int32_t*>(base + gts[gi].list[k]);
    }
    d.pts = d.ext[0] * d.ext[1] * d.ext[2];
    d.dx[0] = FastDiv(uint32_t(std::max(1, d.ext[0])));
    d.dx[1] = FastDiv(uint32_t(std::max(1, d.ext[1])));
    dg[gi] = d;
  }
  if (!dg.empty()) std::memcpy(host.data() + off_gt, dg.data(), sizeof(DevGroupT) * dg.size());
  HIPCHECK(hipMemcpy(A->dmem, host.data(), host.size(), hipMemcpyHostToDevice));
  A->dgt = reinterpret_cast<DevGroupT*>(base + off_gt);
  A->ditems = reinterpret_cast<AdjItem*>(base + off_items);
  A->dentries = reinterpret_cast<int32_t*>(base + off_entries);
  A->dzero = reinterpret_cast<Tile*>(base + off_zero);
}

}  // namespace
}  // namespace tah

using namespace tah;

extern "C" tah_status tah_fill_adjoint(tah_plan* p, double* adj, const double* factors, void* stream) {
  return guard([&] {
    if (!p || !adj) fail("null argument");
    if (p->needs_factors && !factors)
      fail("this schedule mirrors ghosts across a reflecting face, but no factor table was given");
    HIPCHECK(hipSetDevice(p->ctx->device));
    if (!p->adj) build_adjoint(p);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    Job job = plan_job(p, adj, factors);
    if (std::memcmp(&job, &p->last_job, sizeof(Job)) != 0) {
      p->last_job = job;
      HIPCHECK(hipMemcpyAsync(p->djob, &p->last_job, sizeof(Job), hipMemcpyHostToDevice, s));
    }
    for (int32_t st = p->nstages - 1; st >= 0; --st) {
      const auto& r = p->adj->stage_ranges[st];
      if (r[1] > 0) {
        hipLaunchKernelGGL(adjoint_gather_kernel, dim3(uint32_t(r[1])), dim3(kThreads), 0, s, p->djob,
                           p->adj->dgt, p->adj->ditems + r[0], p->adj->dentries);
        HIPCHECK(hipGetLastError());
      }
      if (r[3] > 0) {
        hipLaunchKernelGGL(zero_kernel, dim3(uint32_t(r[3])), dim3(kThreads), 0, s, p->djob,
                           p->adj->dzero + r[2]);
        HIPCHECK(hipGetLastError());
      }
    }
  });
}
