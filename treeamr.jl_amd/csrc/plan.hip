// Plan construction: kernel classes, device tables, tiles.
#include "plan.hpp"

#include <algorithm>
#include <cstring>

namespace tah {

namespace {

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

struct Blob {
  std::vector<char> host;
  int64_t put(const void* p, int64_t bytes) {
    int64_t off = align_up(int64_t(host.size()), 256);
    host.resize(off + std::max<int64_t>(bytes, 8));
    if (bytes) std::memcpy(host.data() + off, p, bytes);
    return off;
  }
};

// Pair the target rows of one dimension whose source windows start `shift`
// apart (in either row order); the rest stay single.
void pair_rows(const std::vector<int32_t>& ss, int shift, std::vector<int32_t>& pairs,
               std::vector<int32_t>& singles) {
  int n = int(ss.size());
  for (int i = 0; i < n;) {
    if (i + 1 < n && ss[i + 1] - ss[i] == shift) {
      pairs.push_back(i);
      pairs.push_back(i + 1);
      i += 2;
    } else if (i + 1 < n && shift > 0 && ss[i] - ss[i + 1] == shift) {
      pairs.push_back(i + 1);
      pairs.push_back(i);
      i += 2;
    } else {
      singles.push_back(i);
      i += 1;
    }
  }
}

// A unit copy whose source row is affine in the target row along every
// dimension: ca + cs*r with cs in {-1, 0, 1}.
bool affine_copy(const PlanGroup& g, int32_t ca[3], int32_t cs[3]) {
  if (!g.unit) return false;
  for (int d = 0; d < 3; ++d) {
    const auto& s = g.srcstart[d];
    ca[d] = s.empty() ? 0 : s[0];
    cs[d] = s.size() > 1 ? s[1] - s[0] : 0;
    if (cs[d] < -1 || cs[d] > 1) return false;
    for (size_t r = 0; r < s.size(); ++r)
      if (s[r] != ca[d] + cs[d] * int32_t(r)) return false;
  }
  return true;
}

}  // namespace

void classify(tah_plan* p) {
  for (size_t gi = 0; gi < p->groups.size(); ++gi) {
    PlanGroup& g = p->groups[gi];
    g.cls = kClsGeneric;
    int32_t ca[3], cs[3];
    if (g.points == 0) continue;
    if (affine_copy(g, ca, cs)) {
      g.cls = kClsCopy;
      continue;
    }
    if (g.kind == 1) continue;
    int P = g.order[0];
    if (g.order[1] != P || g.order[2] != P) continue;
    int shift = g.kind == 2 ? 2 : 0;
    int idx = interp_index(P, shift);
    if (idx < 0) continue;
    std::array<std::vector<int32_t>, 3> pairs, singles;
    for (int d = 0; d < 3; ++d) pair_rows(g.srcstart[d], shift, pairs[d], singles[d]);
    g.cls = kClsInterp;  // marks "split into subs"
    for (int rbits = 0; rbits < 8; ++rbits) {
      PlanSub s;
      s.group = int32_t(gi);
      s.cls = kClsInterp + 8 * idx + rbits;
      bool empty = false;
      for (int d = 0; d < 3; ++d) {
        bool pr = (rbits >> d) & 1;
        s.rows[d] = pr ? pairs[d] : singles[d];
        s.cnt[d] = int32_t(s.rows[d].size()) / (pr ? 2 : 1);
        if (s.cnt[d] == 0) empty = true;
      }
      if (empty) continue;
      g.subs.push_back(int32_t(p->subs.size()));
      p->subs.push_back(std::move(s));
    }
  }
}

// The tiles of one interpolation sub-box (the same for every transfer of its
// group): each a run of units with the source box it stages in LDS. The
// tile's first flat element rides in pad[0].
std::vector<TileBox> interp_tile_shapes(const tah_plan* p, const PlanSub& s, int32_t nvars) {
  const PlanGroup& g = p->groups[s.group];
  const int id = s.cls - kClsInterp, rbits = id % 8;
  const int P = kInterpTable[id / 8][0], shift = kInterpTable[id / 8][1];
  int R[3], W[3];
  for (int d = 0; d < 3; ++d) {
    R[d] = ((rbits >> d) & 1) ? 2 : 1;
    W[d] = P + (R[d] == 2 ? shift : 0);
  }
  const int64_t U = int64_t(s.cnt[0]) * s.cnt[1] * s.cnt[2];
  auto start = [&](int d, int64_t ud) { return g.srcstart[d][s.rows[d][ud * R[d]]]; };
  auto box_of = [&](int64_t u0, int64_t u1) {
    TileBox b{};
    int32_t lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};
    for (int64_t u = u0; u < u1; ++u) {
      int64_t ud[3] = {u % s.cnt[0], (u / s.cnt[0]) % s.cnt[1], u / (int64_t(s.cnt[0]) * s.cnt[1])};
      for (int d = 0; d < 3; ++d) {
        int32_t a = start(d, ud[d]);
        lo[d] = std::min(lo[d], a);
        hi[d] = std::max(hi[d], a + W[d]);
      }
    }
    for (int d = 0; d < 3; ++d) {
      b.lo[d] = lo[d];
      b.ext[d] = hi[d] - lo[d];
    }
    return b;
  };
  auto vol = [](const TileBox& b) { return int64_t(b.ext[0]) * b.ext[1] * b.ext[2]; };
  std::vector<TileBox> out;
  TileBox full = box_of(0, U);
  if (U <= kThreads && vol(full) <= kMaxBox) {
    int64_t nvt = std::max<int64_t>(1, std::min<int64_t>({int64_t(nvars), kThreads / U, kMaxBox / vol(full)}));
    for (int64_t v0 = 0; v0 < nvars; v0 += nvt) {
      TileBox b = full;
      b.v0 = int32_t(v0);
      b.nv = int32_t(std::min<int64_t>(nvt, nvars - v0));
      b.count = int32_t(b.nv * U);
      b.pad[0] = int32_t(v0 * U);
      out.push_back(b);
    }
    return out;
  }
  int64_t cs = kThreads;
  std::vector<TileBox> chunks;
  while (true) {
    chunks.clear();
    bool fits = true;
    for (int64_t u0 = 0; u0 < U; u0 += cs) {
      int64_t u1 = std::min(U, u0 + cs);
      TileBox b = box_of(u0, u1);
      b.count = int32_t(u1 - u0);
      b.pad[0] = int32_t(u0);
      if (vol(b) > kMaxBox) fits = false;
      chunks.push_back(b);
    }
    if (fits || cs == 1) break;
    cs /= 2;
  }
  for (int32_t v = 0; v < nvars; ++v)
    for (TileBox b : chunks) {
      b.v0 = v;
      b.nv = 1;
      b.pad[0] = int32_t(v * U) + b.pad[0];
      out.push_back(b);
    }
  return out;
}

void build_plan_device(tah_plan* p) {
  classify(p);
  Blob blob;
  struct Offs {
    int64_t tb, sb, ss[3], w[3];
  };
  std::vector<Offs> offs(p->groups.size());
  for (size_t gi = 0; gi < p->groups.size(); ++gi) {
    const PlanGroup& g = p->groups[gi];
    Offs& o = offs[gi];
    o.tb = blob.put(g.tblocks.data(), 4 * int64_t(g.tblocks.size()));
    o.sb = blob.put(g.sblocks.data(), 4 * int64_t(g.sblocks.size()));
    for (int d = 0; d < 3; ++d) {
      o.ss[d] = blob.put(g.srcstart[d].data(), 4 * int64_t(g.srcstart[d].size()));
      o.w[d] = blob.put(g.weights[d].data(), 8 * int64_t(g.weights[d].size()));
    }
  }
  std::vector<std::array<int64_t, 3>> subrows(p->subs.size());
  for (size_t si = 0; si < p->subs.size(); ++si)
    for (int d = 0; d < 3; ++d)
      subrows[si][d] = blob.put(p->subs[si].rows[d].data(), 4 * int64_t(p->subs[si].rows[d].size()));
  int64_t off_groups = blob.put(nullptr, int64_t(sizeof(DevGroup)) * std::max<size_t>(1, p->groups.size()));
  int64_t off_subs = blob.put(nullptr, int64_t(sizeof(DevSub)) * std::max<size_t>(1, p->subs.size()));
  HIPCHECK(hipMalloc(&p->dtables, blob.host.size()));
  char* base = p->dtables;
  p->devgroups.clear();
  for (size_t gi = 0; gi < p->groups.size(); ++gi) {
    const PlanGroup& h = p->groups[gi];
    const Offs& o = offs[gi];
    DevGroup g{};
    g.kind = h.kind;
    g.factorcol = h.factorcol;
    g.ntransfers = int32_t(h.tblocks.size());
    g.unit = h.unit ? 1 : 0;
    g.points = h.points;
    g.tblocks = reinterpret_cast<const int32_t*>(base + o.tb);
    g.sblocks = reinterpret_cast<const int32_t*>(base + o.sb);
    int32_t ca[3] = {0, 0, 0}, cs[3] = {0, 0, 0};
    if (h.cls == kClsCopy) affine_copy(h, ca, cs);
    for (int d = 0; d < 3; ++d) {
      g.n[d] = h.n[d];
      g.order[d] = h.order[d];
      g.tfirst[d] = h.tfirst[d];
      g.ca[d] = ca[d];
      g.cs[d] = cs[d];
      g.dn[d] = FastDiv(uint32_t(std::max(1, h.n[d])));
      g.srcstart[d] = reinterpret_cast<const int32_t*>(base + o.ss[d]);
      g.weights[d] = reinterpret_cast<const double*>(base + o.w[d]);
    }
    p->devgroups.push_back(g);
  }
  if (!p->devgroups.empty())
    std::memcpy(blob.host.data() + off_groups, p->devgroups.data(), sizeof(DevGroup) * p->devgroups.size());
  std::vector<DevSub> ds;
  for (size_t si = 0; si < p->subs.size(); ++si) {
    const PlanSub& s = p->subs[si];
    DevSub d{};
    d.group = s.group;
    d.units = s.cnt[0] * s.cnt[1] * s.cnt[2];
    for (int k = 0; k < 3; ++k) {
      d.cnt[k] = s.cnt[k];
      d.rows[k] = reinterpret_cast<const int32_t*>(base + subrows[si][k]);
    }
    d.dc[0] = FastDiv(uint32_t(s.cnt[0]));
    d.dc[1] = FastDiv(uint32_t(s.cnt[1]));
    d.du = FastDiv(uint32_t(d.units));
    ds.push_back(d);
  }
  if (!ds.empty()) std::memcpy(blob.host.data() + off_subs, ds.data(), sizeof(DevSub) * ds.size());
  p->dgroups = reinterpret_cast<DevGroup*>(base + off_groups);
  p->dsubs = reinterpret_cast<DevSub*>(base + off_subs);
  HIPCHECK(hipMemcpy(p->dtables, blob.host.data(), blob.host.size(), hipMemcpyHostToDevice));

  // Tiles: per stage group, per class; within a launch ordered by target
  // block, so that the tiles of one block run close together in time.
  const int32_t nvars = p->layout.nvars;
  p->tiles.clear();
  p->boxes.clear();
  p->launches.clear();
  for (const auto& sg : p->stagegroups) {
    std::vector<Launch> ls;
    std::vector<int> classes;
    for (const PlanGroup& g : p->groups) {
      if (g.stage < sg[0] || g.stage > sg[1] || g.points == 0 || g.tblocks.empty()) continue;
      if (g.cls == kClsInterp) {
        for (int32_t si : g.subs) classes.push_back(p->subs[si].cls);
      } else {
        classes.push_back(g.cls);
      }
    }
    std::sort(classes.begin(), classes.end());
    classes.erase(std::unique(classes.begin(), classes.end()), classes.end());
    for (int cls : classes) {
      struct Entry {
        int32_t block;
        Tile t;
        TileBox b;
      };
      std::vector<Entry> tl;
      uint32_t lds = 0;
      int64_t per = class_tile_elems(cls);
      for (size_t gi = 0; gi < p->groups.size(); ++gi) {
        const PlanGroup& g = p->groups[gi];
        if (g.stage < sg[0] || g.stage > sg[1] || g.points == 0) continue;
        if (g.cls == kClsInterp) {
          for (int32_t si : g.subs) {
            const PlanSub& s = p->subs[si];
            if (s.cls != cls) continue;
            std::vector<TileBox> shapes = interp_tile_shapes(p, s, nvars);
            for (const TileBox& b : shapes)
              lds = std::max<uint32_t>(lds, uint32_t(8 * int64_t(b.ext[0]) * b.ext[1] * b.ext[2] * b.nv));
            for (int32_t t = 0; t < int32_t(g.tblocks.size()); ++t)
              for (const TileBox& b : shapes)
                tl.push_back(Entry{g.tblocks[t], Tile{0, si, t, b.pad[0]}, b});
          }
        } else if (g.cls == cls) {
          int64_t elems = int64_t(g.points) * nvars;
          for (int32_t t = 0; t < int32_t(g.tblocks.size()); ++t)
            for (int64_t e = 0; e < elems; e += per)
              tl.push_back(Entry{g.tblocks[t], Tile{0, int32_t(gi), t, int32_t(e)}, TileBox{}});
        }
      }
      std::stable_sort(tl.begin(), tl.end(),
                       [](const Entry& x, const Entry& y) { return x.block < y.block; });
      Launch L{cls, int64_t(p->tiles.size()), int64_t(tl.size()), lds};
      for (const Entry& e : tl) {
        p->tiles.push_back(e.t);
        TileBox b = e.b;
        b.pad[0] = 0;
        p->boxes.push_back(b);
      }
      if (L.count > 0) ls.push_back(L);
    }
    p->launches.push_back(std::move(ls));
  }
  if (!p->tiles.empty()) {
    HIPCHECK(hipMalloc(&p->dtiles, sizeof(Tile) * p->tiles.size()));
    HIPCHECK(hipMemcpy(p->dtiles, p->tiles.data(), sizeof(Tile) * p->tiles.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMalloc(&p->dboxes, sizeof(TileBox) * p->boxes.size()));
    HIPCHECK(hipMemcpy(p->dboxes, p->boxes.data(), sizeof(TileBox) * p->boxes.size(), hipMemcpyHostToDevice));
  }
  HIPCHECK(hipMalloc(&p->djob, sizeof(Job)));
  std::memset(&p->last_job, 0, sizeof(Job));
}

Job plan_job(const tah_plan* p, double* work, const double* factors) {
  Job job;
  std::memset(&job, 0, sizeof(Job));
  job.groups = p->dgroups;
  job.subs = p->dsubs;
  job.work = work;
  job.factors = factors;
  for (int k = 0; k < 5; ++k) job.stride[k] = p->layout.stride[k];
  job.nvars = p->layout.nvars;
  return job;
}

}  // namespace tah

tah_plan::~tah_plan() {
  if (dtables) (void)hipFree(dtables);
  if (dtiles) (void)hipFree(dtiles);
  if (dboxes) (void)hipFree(dboxes);
  if (djob) (void)hipFree(djob);
}
