// Plan construction: kernel classes, weight-product tables, device tables,
// tiles.
#include "plan.hpp"

#include <algorithm>
#include <cstring>
#include <map>

namespace tah {

// A unit copy whose source row is affine in the target row along every
// dimension: ca + cs*r with |cs| <= max_step (2 takes in vertex injection).
bool affine_copy(const PlanGroup& g, int32_t ca[3], int32_t cs[3], int max_step) {
  if (!g.unit) return false;
  for (int d = 0; d < 3; ++d) {
    const auto& s = g.srcstart[d];
    ca[d] = s.empty() ? 0 : s[0];
    cs[d] = s.size() > 1 ? s[1] - s[0] : 0;
    if (cs[d] < -max_step || cs[d] > max_step) return false;
    for (size_t r = 0; r < s.size(); ++r)
      if (s[r] != ca[d] + cs[d] * int32_t(r)) return false;
  }
  return true;
}

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
// apart (the lower window first); the rest stay single.
void pair_rows(const std::vector<int32_t>& ss, int shift, std::vector<std::array<int32_t, 2>>& pairs,
               std::vector<std::array<int32_t, 2>>& singles) {
  int n = int(ss.size());
  for (int i = 0; i < n;) {
    if (i + 1 < n && ss[i + 1] - ss[i] == shift) {
      pairs.push_back({i, i + 1});
      i += 2;
    } else if (i + 1 < n && shift > 0 && ss[i] - ss[i + 1] == shift) {
      pairs.push_back({i + 1, i});
      i += 2;
    } else {
      singles.push_back({i, -1});
      i += 1;
    }
  }
}

// The chunks of group g (index gi) for LDS-staged launches: each dimension
// of more than 4 target rows cut into runs of at most c rows (8 where two
// dimensions are that long, so that a face's chunk holds 3*8*8 targets; 16
// where one is), the others whole. A chunk's source box is the union of its
// rows' windows. None where a chunk would hold fewer than 64 targets (a
// corner: the unit kernels fill it better) or its box would not fit in LDS
// as pairs of variables.
std::vector<LdsChunk> lds_chunks_of(const PlanGroup& g, int32_t gi) {
  std::vector<LdsChunk> out;
  int nlong = 0;
  for (int d = 0; d < 3; ++d) nlong += g.n[d] > 4 ? 1 : 0;
  const int c = nlong >= 2 ? 8 : 16;
  int C[3];
  int64_t per = 1;
  for (int d = 0; d < 3; ++d) {
    C[d] = g.n[d] > 4 ? c : std::max(1, g.n[d]);
    per *= std::min(C[d], std::max(1, g.n[d]));
  }
  if (per < 64) return out;
  for (int z0 = 0; z0 < g.n[2]; z0 += C[2])
    for (int y0 = 0; y0 < g.n[1]; y0 += C[1])
      for (int x0 = 0; x0 < g.n[0]; x0 += C[0]) {
        LdsChunk ch{};
        ch.group = gi;
        const int lo[3] = {x0, y0, z0};
        int64_t staged = 1;
        for (int d = 0; d < 3; ++d) {
          ch.tlo[d] = lo[d];
          ch.tn[d] = std::min(C[d], g.n[d] - lo[d]);
          int32_t smin = INT32_MAX, smax = INT32_MIN;
          for (int r = lo[d]; r < lo[d] + ch.tn[d]; ++r) {
            smin = std::min(smin, g.srcstart[d][size_t(r)]);
            smax = std::max(smax, g.srcstart[d][size_t(r)]);
          }
          ch.slo[d] = smin;
          ch.sext[d] = smax + g.order[d] - smin;
          staged *= ch.sext[d];
        }
        // every window an even number of rows into the box, for the LDS
        // layout of even rows, then odd ones
        for (int d = 0; d < 3; ++d)
          for (int r = ch.tlo[d]; r < ch.tlo[d] + ch.tn[d]; ++r)
            if ((g.srcstart[d][size_t(r)] - ch.slo[d]) % 2 != 0) return {};
        for (int d = 0; d < 3; ++d) ch.half[d] = (ch.sext[d] + 1) / 2;
        // odd pitches, so that rows and planes apart start in other banks
        ch.pitch[0] = (2 * ch.half[0]) | 1;
        ch.pitch[1] = (2 * ch.half[1] * ch.pitch[0]) | 1;
        ch.slots = 2 * ch.half[2] * ch.pitch[1];
        if (16 * int64_t(ch.slots) > kLdsMaxBytes) return {};
        ch.points = ch.tn[0] * ch.tn[1] * ch.tn[2];
        ch.staged = int32_t(staged);
        // threads' targets fastest along the first dimension of 8 rows or more
        int fast = 0;
        while (fast < 2 && ch.tn[fast] < 8) ++fast;
        if (ch.tn[fast] < 8) fast = 0;
        ch.ord[0] = fast;
        ch.ord[1] = fast == 0 ? 1 : 0;
        ch.ord[2] = 3 - ch.ord[0] - ch.ord[1];
        ch.dt[0] = FastDiv(uint32_t(ch.tn[ch.ord[0]]));
        ch.dt[1] = FastDiv(uint32_t(ch.tn[ch.ord[1]]));
        ch.dsx[0] = FastDiv(uint32_t(ch.sext[0]));
        ch.dsx[1] = FastDiv(uint32_t(ch.sext[1]));
        out.push_back(ch);
      }
  return out;
}

// The prolongation chunks of group g (index gi) over its unit classes cls:
// their units binned by 16 target rows per dimension (by each unit's first
// rows); a bin is a chunk whose source box is the union of its units'
// windows, its units cut per class into waves of at most 64. Appends to
// chunks, waves and units; false, with nothing appended, where a box would
// not fit in LDS as pairs of variables.
bool lds_unit_chunks_of(const tah_plan* p, const PlanGroup& g, int32_t gi, const std::vector<int32_t>& cls,
                        std::vector<LdsUnitChunk>& chunks, std::vector<LdsWave>& waves,
                        std::vector<uint32_t>& units) {
  std::map<std::array<int32_t, 3>, std::vector<std::pair<int32_t, int32_t>>> bins;  // (class, unit)
  for (int32_t ci : cls) {
    const PlanWClass& c = p->wclasses[size_t(ci)];
    for (int32_t k = 0; k < c.nunits(); ++k) {
      const int32_t* u = &c.units[8 * size_t(k)];
      bins[{u[0] / 16, u[2] / 16, u[4] / 16}].push_back({ci, k});
    }
  }
  const size_t c0 = chunks.size(), w0 = waves.size(), u0 = units.size();
  for (const auto& bin : bins) {
    LdsUnitChunk ch{};
    ch.group = gi;
    int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (const auto& cu : bin.second) {
      const int32_t* u = &p->wclasses[size_t(cu.first)].units[8 * size_t(cu.second)];
      for (int d = 0; d < 3; ++d) {
        const int32_t s = g.srcstart[d][size_t(u[2 * d])];
        lo[d] = std::min(lo[d], s);
        hi[d] = std::max(hi[d], s);
      }
    }
    for (int d = 0; d < 3; ++d) {
      ch.slo[d] = lo[d];
      ch.sext[d] = hi[d] + g.order[d] - lo[d];
    }
    ch.pitch[0] = ch.sext[0] | 1;
    ch.pitch[1] = (ch.sext[1] * ch.pitch[0]) | 1;
    ch.slots = ch.sext[2] * ch.pitch[1];
    if (16 * int64_t(ch.slots) > kLdsMaxBytes) {
      chunks.resize(c0);
      waves.resize(w0);
      units.resize(u0);
      return false;
    }
    ch.staged = ch.sext[0] * ch.sext[1] * ch.sext[2];
    ch.dsx[0] = FastDiv(uint32_t(ch.sext[0]));
    ch.dsx[1] = FastDiv(uint32_t(ch.sext[1]));
    ch.wave0 = int32_t(waves.size());
    std::map<int32_t, std::vector<int32_t>> per;  // class -> its units in the bin
    for (const auto& cu : bin.second) per[cu.first].push_back(cu.second);
    for (const auto& cl : per)
      for (size_t i = 0; i < cl.second.size(); i += 64) {
        const int32_t nu = int32_t(std::min<size_t>(64, cl.second.size() - i));
        waves.push_back(LdsWave{cl.first, int32_t(units.size() / 2), nu, 0});
        for (int32_t j = 0; j < nu; ++j) {
          const int32_t* u = &p->wclasses[size_t(cl.first)].units[8 * size_t(cl.second[i + size_t(j)])];
          units.push_back(uint32_t(u[0] & 255) | uint32_t(u[1] & 255) << 8 | uint32_t(u[2] & 255) << 16 |
                          uint32_t(u[3] & 255) << 24);
          units.push_back(uint32_t(u[4] & 255) | uint32_t(u[5] & 255) << 8);
        }
      }
    ch.nwave = int32_t(waves.size()) - ch.wave0;
    chunks.push_back(ch);
  }
  return true;
}

// The plane records of interpolating group g streamed along z (StreamPlane):
// its z rows sorted by window start (stable) enter as their windows begin
// and leave, stored, as they end. Appends to planes; *zlo is the first
// plane's source row. False, with nothing appended, where more than
// stream_ka(P, shift) windows hold one plane.
bool stream_planes_of(const PlanGroup& g, std::vector<StreamPlane>& planes, int32_t* zlo) {
  const int32_t n2 = g.n[2], P = g.order[2];
  std::vector<int32_t> order(static_cast<size_t>(n2));
  for (int32_t k = 0; k < n2; ++k) order[size_t(k)] = k;
  const auto& s = g.srcstart[2];
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return s[size_t(x)] < s[size_t(y)]; });
  auto start = [&](int32_t o) { return s[size_t(order[size_t(o)])]; };
  const size_t first = planes.size();
  *zlo = start(0);
  int32_t ol = 0, oh = -1;
  for (int32_t z = start(0); z <= start(n2 - 1) + P - 1; ++z) {
    StreamPlane pl{};
    pl.had = oh - ol + 1;
    while (oh + 1 < n2 && start(oh + 1) <= z) ++oh;
    pl.nact = oh - ol + 1;
    if (pl.nact > stream_ka(P, g.kind == 2 ? 2 : 0) || pl.nact > kStreamKA) {
      planes.resize(first);
      return false;
    }
    for (int32_t q = 0; q < pl.nact; ++q) {
      const int32_t k = order[size_t(ol + q)];
      pl.c2[q] = g.weights[2][size_t(P) * k + size_t(z - s[size_t(k)])];
    }
    while (ol + pl.nret <= oh && start(ol + pl.nret) + P - 1 == z) {
      pl.k[pl.nret] = order[size_t(ol + pl.nret)];
      ++pl.nret;
    }
    ol += pl.nret;
    planes.push_back(pl);
  }
  return true;
}

// The chunks of interpolating group g (index gi) streamed along z, for tiles
// of nvt variables: the target rows along x and along y cut into balanced
// runs, so that a chunk's targets, each taking one thread per pair of
// variables, fill at most a workgroup, and its plane (the union of its
// windows along x and y, every variable of a tile) takes at most
// kStreamSlots pairs per thread. Of the cuts that fit, the one with the
// fewest chunks, then the fewest staged points. None where none fits.
std::vector<StreamChunk> stream_chunks_of(const PlanGroup& g, int32_t gi, int32_t nvt, int32_t plane0,
                                          int32_t nplanes, int32_t zlo) {
  const int32_t L = (nvt + 1) / 2;
  const int32_t items = kThreads / L;
  if (items < 1) return {};
  const int32_t n0 = g.n[0], n1 = g.n[1];
  auto part = [](int32_t n, int32_t m, int32_t p) { return int32_t(int64_t(n) * p / m); };
  auto extent = [&](int d, int32_t r0, int32_t r1, int32_t* lo) {
    int32_t smin = INT32_MAX, smax = INT32_MIN;
    for (int32_t r = r0; r < r1; ++r) {
      smin = std::min(smin, g.srcstart[d][size_t(r)]);
      smax = std::max(smax, g.srcstart[d][size_t(r)]);
    }
    *lo = smin;
    return smax + g.order[d] - smin;
  };
  int64_t best_chunks = INT64_MAX, best_staged = INT64_MAX;
  int32_t best_m0 = 0, best_m1 = 0;
  for (int32_t c0 = 1; c0 <= std::min(n0, items); ++c0) {
    const int32_t c1 = std::min(n1, items / c0);
    if (c1 < 1) continue;
    const int32_t m0 = (n0 + c0 - 1) / c0, m1 = (n1 + c1 - 1) / c1;
    int64_t sum0 = 0, sum1 = 0;
    int32_t max0 = 0, max1 = 0, lo;
    for (int32_t q = 0; q < m0; ++q) {
      const int32_t e = extent(0, part(n0, m0, q), part(n0, m0, q + 1), &lo);
      sum0 += e;
      max0 = std::max(max0, e);
    }
    for (int32_t q = 0; q < m1; ++q) {
      const int32_t e = extent(1, part(n1, m1, q), part(n1, m1, q + 1), &lo);
      sum1 += e;
      max1 = std::max(max1, e);
    }
    if (int64_t(max0) * max1 * L > int64_t(kStreamSlots) * kThreads) continue;
    if (2 * int64_t(max0) * max1 * 2 * L * 8 > kLdsMaxBytes) continue;
    const int64_t chunks = int64_t(m0) * m1, staged = sum0 * sum1;
    if (chunks < best_chunks || (chunks == best_chunks && staged < best_staged)) {
      best_chunks = chunks;
      best_staged = staged;
      best_m0 = m0;
      best_m1 = m1;
    }
  }
  std::vector<StreamChunk> out;
  if (best_m0 == 0) return out;
  for (int32_t qy = 0; qy < best_m1; ++qy)
    for (int32_t qx = 0; qx < best_m0; ++qx) {
      StreamChunk ch{};
      ch.group = gi;
      ch.tlo[0] = part(n0, best_m0, qx);
      ch.tn[0] = part(n0, best_m0, qx + 1) - ch.tlo[0];
      ch.tlo[1] = part(n1, best_m1, qy);
      ch.tn[1] = part(n1, best_m1, qy + 1) - ch.tlo[1];
      ch.sext[0] = extent(0, ch.tlo[0], ch.tlo[0] + ch.tn[0], &ch.slo[0]);
      ch.sext[1] = extent(1, ch.tlo[1], ch.tlo[1] + ch.tn[1], &ch.slo[1]);
      ch.plane0 = plane0;
      ch.nplanes = nplanes;
      ch.zlo = zlo;
      ch.nvt = nvt;
      // What interp_stream_kernel assumes of its chunk, checked before any
      // launch: its targets' threads fit the workgroup, its plane fits the
      // staging slots and LDS, and every target's window lies in the plane.
      const int32_t P = g.order[0];
      bool ok = ch.tn[0] >= 1 && ch.tn[1] >= 1 && int64_t(ch.tn[0]) * ch.tn[1] * L <= kThreads &&
                int64_t(ch.sext[0]) * ch.sext[1] * L <= int64_t(kStreamSlots) * kThreads &&
                2 * int64_t(ch.sext[0]) * ch.sext[1] * 2 * L * 8 <= kLdsMaxBytes && ch.slo[0] >= 0 && ch.slo[1] >= 0;
      for (int d = 0; d < 2 && ok; ++d)
        for (int32_t r = ch.tlo[d]; r < ch.tlo[d] + ch.tn[d]; ++r) {
          const int32_t o = g.srcstart[d][size_t(r)] - ch.slo[d];
          ok = ok && o >= 0 && o + P <= ch.sext[d];
        }
      if (!ok) fail("internal: a streamed interpolation chunk does not fit its kernel");
      // the smallest workgroup its threads and plane fit (chunks of few
      // targets, such as edges along z and corners, keep their threads busy)
      ch.threads = kThreads;
      for (int32_t T : {64, 128})
        if (ch.threads == kThreads && int64_t(ch.tn[0]) * ch.tn[1] * L <= T &&
            int64_t(ch.sext[0]) * ch.sext[1] * L <= int64_t(kStreamSlots) * T)
          ch.threads = T;
      out.push_back(ch);
    }
  return out;
}

// Bitwise key of a vector of doubles.
std::vector<uint64_t> bits_of(const double* w, size_t n) {
  std::vector<uint64_t> k(n);
  std::memcpy(k.data(), w, n * sizeof(double));
  return k;
}

}  // namespace

void classify(tah_plan* p) {
  std::map<std::vector<uint64_t>, int64_t> tables;  // W content -> offset in wblob
  p->wblob.clear();
  p->wclasses.clear();
  for (size_t gi = 0; gi < p->groups.size(); ++gi) {
    PlanGroup& g = p->groups[gi];
    g.cls = kClsGeneric;
    g.wclasses.clear();
    int32_t ca[3], cs[3];
    if (g.points == 0) continue;
    if (affine_copy(g, ca, cs, 2)) {
      g.cls = kClsCopy;
      continue;
    }
    if (g.kind == 1) continue;
    const int P = g.order[0];
    if (g.order[1] != P || g.order[2] != P) continue;
    const int shift = g.kind == 2 ? 2 : 0;
    const int idx = interp_index(P, shift);
    if (idx < 0) continue;
    g.cls = kClsInterp;
    std::array<std::vector<std::array<int32_t, 2>>, 3> pairs, singles;
    for (int d = 0; d < 3; ++d) pair_rows(g.srcstart[d], shift, pairs[d], singles[d]);
    for (int rbits = 0; rbits < 8; ++rbits) {
      int R[3];
      std::array<const std::vector<std::array<int32_t, 2>>*, 3> E;
      bool empty = false;
      for (int d = 0; d < 3; ++d) {
        R[d] = ((rbits >> d) & 1) ? 2 : 1;
        E[d] = R[d] == 2 ? &pairs[d] : &singles[d];
        empty = empty || E[d]->empty();
      }
      if (empty) continue;
      // Signature of every entry: the weights of its rows.
      std::array<std::vector<int>, 3> sig;
      for (int d = 0; d < 3; ++d) {
        std::map<std::vector<uint64_t>, int> ids;
        for (const auto& e : *E[d]) {
          std::vector<uint64_t> key;
          for (int r = 0; r < R[d]; ++r) {
            auto k = bits_of(&g.weights[d][size_t(P) * e[r]], P);
            key.insert(key.end(), k.begin(), k.end());
          }
          auto it = ids.emplace(key, int(ids.size())).first;
          sig[d].push_back(it->second);
        }
      }
      std::map<std::array<int, 3>, int32_t> byclass;
      for (size_t ez = 0; ez < E[2]->size(); ++ez)
        for (size_t ey = 0; ey < E[1]->size(); ++ey)
          for (size_t ex = 0; ex < E[0]->size(); ++ex) {
            std::array<int, 3> key{sig[0][ex], sig[1][ey], sig[2][ez]};
            auto it = byclass.find(key);
            if (it == byclass.end()) {
              PlanWClass wc;
              wc.group = int32_t(gi);
              wc.cls = kClsInterp + 8 * idx + rbits;
              // The weight products of this class, in the kernel's visit order.
              const auto& ex0 = (*E[0])[ex];
              const auto& ey0 = (*E[1])[ey];
              const auto& ez0 = (*E[2])[ez];
              const int U0 = P + (R[0] == 2 ? shift : 0), U1 = P + (R[1] == 2 ? shift : 0),
                        U2 = P + (R[2] == 2 ? shift : 0);
              auto w = [&](int d, int row, int k) { return g.weights[d][size_t(P) * row + k]; };
              std::vector<double> W;
              for (int m2 = 0; m2 < U2; ++m2)
                for (int m1 = 0; m1 < U1; ++m1)
                  for (int c = 0; c < R[2]; ++c) {
                    int k2 = m2 - c * shift;
                    if (k2 < 0 || k2 >= P) continue;
                    for (int b = 0; b < R[1]; ++b) {
                      int k1 = m1 - b * shift;
                      if (k1 < 0 || k1 >= P) continue;
                      for (int a = 0; a < R[0]; ++a)
                        for (int m0 = 0; m0 < U0; ++m0) {
                          int k0 = m0 - a * shift;
                          if (k0 < 0 || k0 >= P) continue;
                          double p01 = w(0, ex0[a], k0) * w(1, ey0[b], k1);
                          W.push_back(p01 * w(2, ez0[c], k2));
                        }
                    }
                  }
              auto key2 = bits_of(W.data(), W.size());
              auto t = tables.find(key2);
              if (t == tables.end()) {
                int64_t off = int64_t(p->wblob.size());
                p->wblob.insert(p->wblob.end(), W.begin(), W.end());
                // keep tables 256-byte aligned
                while (p->wblob.size() % 32) p->wblob.push_back(0.0);
                t = tables.emplace(key2, off).first;
              }
              wc.wtable = t->second;
              it = byclass.emplace(key, int32_t(p->wclasses.size())).first;
              g.wclasses.push_back(it->second);
              p->wclasses.push_back(std::move(wc));
            }
            PlanWClass& wc = p->wclasses[it->second];
            const auto& a = (*E[0])[ex];
            const auto& b = (*E[1])[ey];
            const auto& c = (*E[2])[ez];
            int32_t u[8] = {a[0], a[1], b[0], b[1], c[0], c[1], 0, 0};
            wc.units.insert(wc.units.end(), u, u + 8);
          }
    }
  }
}

// The hook's tiles and the boxes the relative form reads (tah_plan). A
// region whose slot names no stage of the plan is never run by the
// relative form, as in the reference, whose relative batches run only
// after a stage of their slot (src/ghosts.jl:1527-1531, 1546-1547).
void build_hooks(tah_plan* p) {
  const int nsg = int(p->stagegroups.size());
  p->phase1_sg = -1;
  for (int k = 0; k < nsg; ++k)
    for (int st = p->stagegroups[k][0]; st <= p->stagegroups[k][1]; ++st)
      if (p->tags[st] == TAH_STAGE_PHASE1) p->phase1_sg = k;
  auto sg_of_slot = [&](int32_t slot) {
    for (int k = 0; k < nsg; ++k)
      for (int st = p->stagegroups[k][0]; st <= p->stagegroups[k][1]; ++st)
        if (p->tags[st] == slot) return k;
    return -1;
  };
  const int32_t nvars = p->layout.nvars;
  std::vector<HookTile> tiles;
  auto push = [&](int32_t r) {
    const HookRegion& R = p->hooks[size_t(r)];
    const int64_t total = int64_t(R.n[0]) * R.n[1] * R.n[2] * nvars;
    for (int64_t e = 0; e < total; e += kHookTile) tiles.push_back(HookTile{r, int32_t(e)});
  };
  for (size_t r = 0; r < p->hooks.size(); ++r) push(int32_t(r));
  p->hook_all[0] = 0;
  p->hook_all[1] = int64_t(tiles.size());
  p->hook_after.assign(size_t(nsg), {0, 0});
  p->hook_reads.assign(size_t(nsg), {});
  for (int k = 0; k < nsg; ++k) {
    p->hook_after[size_t(k)][0] = int64_t(tiles.size());
    for (size_t r = 0; r < p->hooks.size(); ++r) {
      if (sg_of_slot(p->hook_slots[r]) != k) continue;
      push(int32_t(r));
      const HookRegion& R = p->hooks[r];
      std::array<int32_t, 7> box{R.block, 0, 0, 0, 0, 0, 0};
      for (int d = 0; d < 3; ++d) {
        box[1 + d] = R.clamp[d] >= 0 ? R.clamp[d] : R.lo[d];
        box[4 + d] = R.clamp[d] >= 0 ? R.clamp[d] : R.lo[d] + R.n[d] - 1;
      }
      p->hook_reads[size_t(k)].push_back(box);
    }
    p->hook_after[size_t(k)][1] = int64_t(tiles.size()) - p->hook_after[size_t(k)][0];
  }
  if (tiles.size() > size_t(INT32_MAX / 2)) fail("too many hook tiles", kUnsupported);
  if (!p->hooks.empty()) {
    HIPCHECK(hipMalloc(&p->dhooks, sizeof(HookRegion) * p->hooks.size()));
    HIPCHECK(hipMemcpy(p->dhooks, p->hooks.data(), sizeof(HookRegion) * p->hooks.size(), hipMemcpyHostToDevice));
  }
  if (!tiles.empty()) {
    HIPCHECK(hipMalloc(&p->dhook_tiles, sizeof(HookTile) * tiles.size()));
    HIPCHECK(hipMemcpy(p->dhook_tiles, tiles.data(), sizeof(HookTile) * tiles.size(), hipMemcpyHostToDevice));
  }
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
  std::vector<int64_t> unit_off(p->wclasses.size());
  for (size_t ci = 0; ci < p->wclasses.size(); ++ci)
    unit_off[ci] = blob.put(p->wclasses[ci].units.data(), 4 * int64_t(p->wclasses[ci].units.size()));
  int64_t off_w = blob.put(p->wblob.data(), 8 * int64_t(p->wblob.size()));
  int64_t off_groups = blob.put(nullptr, int64_t(sizeof(DevGroup)) * std::max<size_t>(1, p->groups.size()));
  int64_t off_wc = blob.put(nullptr, int64_t(sizeof(DevWClass)) * std::max<size_t>(1, p->wclasses.size()));
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
    // 2: a unit copy whose rows are affine (ca, cs below hold its map).
    g.unit = h.cls == kClsCopy ? 2 : h.unit ? 1 : 0;
    g.points = h.points;
    g.tblocks = reinterpret_cast<const int32_t*>(base + o.tb);
    g.sblocks = reinterpret_cast<const int32_t*>(base + o.sb);
    int32_t ca[3] = {0, 0, 0}, cs[3] = {0, 0, 0};
    if (h.cls == kClsCopy) affine_copy(h, ca, cs, 2);
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
    g.dper = FastDiv(uint32_t(std::max<int64_t>(1, int64_t(h.points) * p->layout.nvars)));
    p->devgroups.push_back(g);
  }
  if (!p->devgroups.empty())
    std::memcpy(blob.host.data() + off_groups, p->devgroups.data(), sizeof(DevGroup) * p->devgroups.size());
  std::vector<DevWClass> dw;
  for (size_t ci = 0; ci < p->wclasses.size(); ++ci) {
    const PlanWClass& c = p->wclasses[ci];
    DevWClass d{};
    d.group = c.group;
    d.nunits = c.nunits();
    d.units = reinterpret_cast<const int32_t*>(base + unit_off[ci]);
    d.W = reinterpret_cast<const double*>(base + off_w) + c.wtable;
    d.rbits = (c.cls - kClsInterp) % 8;
    dw.push_back(d);
  }
  if (!dw.empty()) std::memcpy(blob.host.data() + off_wc, dw.data(), sizeof(DevWClass) * dw.size());
  p->dgroups = reinterpret_cast<DevGroup*>(base + off_groups);
  p->dwclasses = reinterpret_cast<DevWClass*>(base + off_wc);
  HIPCHECK(hipMemcpy(p->dtables, blob.host.data(), blob.host.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMalloc(&p->djob, sizeof(Job)));
  std::memset(&p->last_job, 0, sizeof(Job));
  HIPCHECK(hipEventCreateWithFlags(&p->done, hipEventDisableTiming));
  build_hooks(p);
  // Where variables are not innermost (the reference's dense layout) no
  // variable-pair kernel applies, and prolongation streamed along z through
  // LDS is the faster form (the large forest: 8.5 ms against 17.5 ms for the
  // unit kernel); with variables innermost the pair kernel is (4.5 ms
  // against 7.4 ms).
  if (!(p->layout.stride[3] < p->layout.stride[0])) p->lds |= 8;
  families_from_options(p);
  if (p->regrid) {
    HIPCHECK(hipMalloc(&p->djob2, sizeof(Job)));
    std::memset(&p->last_job2, 0, sizeof(Job));
  }
  build_launches(p);
}

void build_launches(tah_plan* p) {
  // Which stage group each transfer runs in: its own, or the last one when
  // no later stage reads what it writes nor writes what it reads, so that it
  // can run beside the last group's work (most of phase 1 beside the
  // prolongations).
  const int nsg = int(p->stagegroups.size());
  std::vector<int> sg_of_stage(std::max(1, p->nstages), 0);
  for (int k = 0; k < nsg; ++k)
    for (int st = p->stagegroups[k][0]; st <= p->stagegroups[k][1]; ++st) sg_of_stage[st] = k;
  struct Box3 {
    int lo[3], hi[3];
  };
  auto overlaps = [](const Box3& x, const Box3& y) {
    for (int d = 0; d < 3; ++d)
      if (std::max(x.lo[d], y.lo[d]) > std::min(x.hi[d], y.hi[d])) return false;
    return true;
  };
  auto wbox = [](const PlanGroup& g) {
    Box3 b;
    for (int d = 0; d < 3; ++d) {
      b.lo[d] = g.tfirst[d];
      b.hi[d] = g.tfirst[d] + g.n[d] - 1;
    }
    return b;
  };
  auto rbox = [](const PlanGroup& g) {
    Box3 b;
    for (int d = 0; d < 3; ++d) {
      b.lo[d] = *std::min_element(g.srcstart[d].begin(), g.srcstart[d].end());
      b.hi[d] = *std::max_element(g.srcstart[d].begin(), g.srcstart[d].end()) + g.order[d] - 1;
    }
    return b;
  };
  std::vector<std::vector<int>> run_in(p->groups.size());
  for (size_t gi = 0; gi < p->groups.size(); ++gi)
    run_in[gi].assign(p->groups[gi].tblocks.size(), sg_of_stage[p->groups[gi].stage]);
  p->deferred = 0;
  {
    std::map<int32_t, std::vector<Box3>> later_reads, later_writes;
    for (int k = nsg - 1; k >= 0; --k) {
      // the relative hook after this group reads what this group and the
      // ones before it wrote
      if (size_t(k) < p->hook_reads.size())
        for (const auto& h : p->hook_reads[size_t(k)])
          later_reads[h[0]].push_back(Box3{{h[1], h[2], h[3]}, {h[4], h[5], h[6]}});
      if (k < nsg - 1) {
        for (size_t gi = 0; gi < p->groups.size(); ++gi) {
          const PlanGroup& g = p->groups[gi];
          if (g.points == 0 || sg_of_stage[g.stage] != k) continue;
          const Box3 wb = wbox(g), rb = rbox(g);
          // (a regrid's stages read different arrays: nothing moves)
          for (size_t t = 0; t < g.tblocks.size() && p->defer && !p->regrid; ++t) {
            bool ok = true;
            auto ir = later_reads.find(g.tblocks[t]);
            if (ir != later_reads.end())
              for (const Box3& x : ir->second) ok = ok && !overlaps(x, wb);
            auto iw = later_writes.find(g.sblocks[t]);
            if (ok && iw != later_writes.end())
              for (const Box3& x : iw->second) ok = ok && !overlaps(x, rb);
            if (ok) {
              run_in[gi][t] = nsg - 1;
              ++p->deferred;
            }
          }
        }
      }
      // Then this group's own accesses join the later ones.
      for (size_t gi = 0; gi < p->groups.size(); ++gi) {
        const PlanGroup& g = p->groups[gi];
        if (g.points == 0 || sg_of_stage[g.stage] != k) continue;
        const Box3 wb = wbox(g), rb = rbox(g);
        for (size_t t = 0; t < g.tblocks.size(); ++t) {
          later_writes[g.tblocks[t]].push_back(wb);
          later_reads[g.sblocks[t]].push_back(rb);
        }
      }
    }
  }

  // Tiles: per stage group, per class; within a launch ordered by target
  // block (restriction: by source block; copies: by the plan option
  // copy_order), so that the tiles of one block run close together in time.
  // (Copies by source block: 6.60 ms per fill on the large forest against
  // 6.70 ms by target block and 6.64 ms by the lower of the two; the
  // 64-forest batch's forked fill 20.29-20.30 ms against 19.92-19.96 ms by
  // target block; interleaved runs on one box.)
  const bool copy_by_source = p->copy_order == 0;
  const int32_t nvars = p->layout.nvars;
  const bool vfast = p->layout.stride[3] < p->layout.stride[0];
  p->tiles.clear();
  p->lds_tiles.clear();
  p->launches.clear();
  // The launch an interpolation class runs in: its own, or with every
  // class of its order and shift (tah_plan::merge).
  // Merged prolongations take tiles of variable pairs where the layout
  // allows: variables innermost, an even count, every other stride even.
  // (Restriction over pairs measured slower than over single variables.)
  const int64_t* st = p->layout.stride;
  // (an odd count, or odd strides, take pairs of single loads, the last
  // variable of an odd count alone: the per-call flag pairs is then off)
  const bool pair_layout = st[3] == 1 && nvars > 1;
  auto launch_of = [&](int cls) {
    if (cls < kClsInterp || cls >= kClsMerged) return cls;
    const int index = (cls - kClsInterp) / 8;
    if (!p->fam_merge[size_t(index)]) return cls;
    const bool pair = pair_layout && p->fam_pairs[size_t(index)] && kInterpTable[size_t(index)][1] == 0;
    return (pair ? kClsMergedPair : kClsMerged) + index;
  };
  // The family of a group (-1: an order no specialised kernel takes).
  auto family_of = [&](const PlanGroup& g) { return interp_index(g.order[0], g.kind == 2 ? 2 : 0); };
  // a group's family's LDS staging: 0 none, 1 LDS-staged chunks, 2
  // (restriction) streamed along z
  auto fam_lds = [&](const PlanGroup& g) {
    const int f = family_of(g);
    return f >= 0 ? p->fam_lds[size_t(f)] : 0;
  };
  // The launch a copy or generic group runs in.
  auto group_launch = [&](int cls) { return cls == kClsCopy ? copy_launch(p->copy_per_thread) : cls; };
  // LDS-staged restriction (tah_plan::lds): the groups whose boxes cut into
  // chunks of workgroup size, in a layout that holds variables innermost in
  // even number; their chunks, and the launch class of each such group.
  const bool pair_capable = st[3] == 1 && nvars % 2 == 0 && st[0] % 2 == 0 && st[1] % 2 == 0 &&
                            st[2] % 2 == 0 && st[4] % 2 == 0;
  p->lds_chunks.clear();
  std::vector<std::array<int32_t, 2>> lds_of(p->groups.size(), {0, 0});  // first chunk, count
  if (pair_capable)
    for (size_t gi = 0; gi < p->groups.size(); ++gi) {
      const PlanGroup& g = p->groups[gi];
      if (g.cls == kClsInterp && g.kind == 2 && fam_lds(g) == 1) {
        const auto cs = lds_chunks_of(g, int32_t(gi));
        lds_of[gi] = {int32_t(p->lds_chunks.size()), int32_t(cs.size())};
        p->lds_chunks.insert(p->lds_chunks.end(), cs.begin(), cs.end());
      }
    }
  auto lds_launch = [&](const PlanGroup& g) { return kClsLds + interp_index(g.order[0], g.kind == 2 ? 2 : 0); };
  // Interpolation streamed along z (family option lds 2): the groups whose z
  // windows overlap at most stream_ka(P, shift) to a plane and whose boxes
  // cut into chunks that fit, in any layout; their chunks and plane records.
  p->stream_chunks.clear();
  p->stream_planes.clear();
  std::vector<std::array<int32_t, 2>> str_of(p->groups.size(), {0, 0});  // first chunk, count
  const int32_t stream_nvt = std::min<int32_t>(nvars, 2 * kThreads);
  for (size_t gi = 0; gi < p->groups.size(); ++gi) {
    const PlanGroup& g = p->groups[gi];
    if (g.cls != kClsInterp || fam_lds(g) != 2 || g.n[2] < 1) continue;
    const int32_t plane0 = int32_t(p->stream_planes.size());
    int32_t zlo = 0;
    if (!stream_planes_of(g, p->stream_planes, &zlo)) continue;
    const int32_t nplanes = int32_t(p->stream_planes.size()) - plane0;
    const auto cs = stream_chunks_of(g, int32_t(gi), stream_nvt, plane0, nplanes, zlo);
    if (cs.empty()) {
      p->stream_planes.resize(size_t(plane0));
      continue;
    }
    str_of[gi] = {int32_t(p->stream_chunks.size()), int32_t(cs.size())};
    p->stream_chunks.insert(p->stream_chunks.end(), cs.begin(), cs.end());
  }
  auto stream_launch = [&](const PlanGroup& g) { return kClsStream + family_of(g); };
  // LDS-staged prolongation by units: the unit classes of at least 64 units
  // (a face's inner units); a group's other classes stay with the unit
  // kernels.
  p->lds_uchunks.clear();
  p->lds_waves.clear();
  p->lds_units.clear();
  std::vector<char> lds_cls(p->wclasses.size(), 0);
  std::vector<std::array<int32_t, 2>> ulds_of(p->groups.size(), {0, 0});  // first chunk, count
  if (pair_capable)
    for (size_t gi = 0; gi < p->groups.size(); ++gi) {
      const PlanGroup& g = p->groups[gi];
      if (g.cls != kClsInterp || g.kind != 3 || fam_lds(g) != 1) continue;
      std::vector<int32_t> cls;
      for (int32_t ci : g.wclasses)
        if (p->wclasses[size_t(ci)].nunits() >= 64) cls.push_back(ci);
      if (cls.empty()) continue;
      const size_t c0 = p->lds_uchunks.size();
      if (!lds_unit_chunks_of(p, g, int32_t(gi), cls, p->lds_uchunks, p->lds_waves, p->lds_units)) continue;
      ulds_of[gi] = {int32_t(c0), int32_t(p->lds_uchunks.size() - c0)};
      for (int32_t ci : cls) lds_cls[size_t(ci)] = 1;
    }
  auto ulds_launch = [&](const PlanGroup& g) { return kClsLdsProlong + interp_index(g.order[0], 0); };
  for (int k = 0; k < nsg; ++k) {
    std::vector<Launch> ls;
    std::vector<int> classes;
    for (size_t gi = 0; gi < p->groups.size(); ++gi) {
      const PlanGroup& g = p->groups[gi];
      if (g.points == 0 || std::find(run_in[gi].begin(), run_in[gi].end(), k) == run_in[gi].end()) continue;
      if (str_of[gi][1] > 0) {
        classes.push_back(stream_launch(g));
      } else if (lds_of[gi][1] > 0) {
        classes.push_back(lds_launch(g));
      } else if (g.cls == kClsInterp) {
        if (ulds_of[gi][1] > 0) classes.push_back(ulds_launch(g));
        for (int32_t ci : g.wclasses)
          if (!lds_cls[size_t(ci)]) classes.push_back(launch_of(p->wclasses[ci].cls));
      } else {
        classes.push_back(group_launch(g.cls));
      }
    }
    std::sort(classes.begin(), classes.end());
    classes.erase(std::unique(classes.begin(), classes.end()), classes.end());
    struct Built {
      int cls;
      std::vector<std::pair<int32_t, Tile>> tiles;  // (ordering block, tile)
      int64_t values;
      std::vector<std::pair<int32_t, LdsTile>> ltiles;  // LDS-staged: (source block, tile)
      int32_t lds_bytes = 0;
      int32_t threads = kThreads;  // LDS-staged: the workgroup size
    };
    std::vector<Built> built;
    for (int cls : classes) {
      if (cls >= kClsStream) {
        // A transfer's chunks one after another, each for every tile of
        // variables in turn, ordered by source block (restriction) or
        // target block (prolongation) as the unit tiles are; one launch
        // per workgroup size.
        for (int32_t T : {kThreads, 128, 64}) {
          Built B{cls, {}, 0, {}, 0};
          B.threads = T;
          for (size_t gi = 0; gi < p->groups.size(); ++gi) {
            const PlanGroup& g = p->groups[gi];
            if (g.points == 0 || str_of[gi][1] == 0 || stream_launch(g) != cls) continue;
            for (int32_t t = 0; t < int32_t(g.tblocks.size()); ++t) {
              if (run_in[gi][t] != k) continue;
              for (int32_t c = str_of[gi][0]; c < str_of[gi][0] + str_of[gi][1]; ++c) {
                const StreamChunk& ch = p->stream_chunks[size_t(c)];
                if (ch.threads != T) continue;
                for (int32_t v0 = 0; v0 < nvars; v0 += ch.nvt) {
                  const int32_t nv = std::min(ch.nvt, nvars - v0);
                  // two plane buffers
                  B.lds_bytes =
                      std::max(B.lds_bytes, int32_t(2 * ch.sext[0] * ch.sext[1] * 2 * ((nv + 1) / 2) * 8));
                  B.ltiles.push_back({g.kind == 2 ? g.sblocks[t] : g.tblocks[t], LdsTile{0, c, t, v0}});
                  B.values += int64_t(ch.tn[0]) * ch.tn[1] * g.n[2] * nv;
                }
              }
            }
          }
          std::stable_sort(B.ltiles.begin(), B.ltiles.end(),
                           [](const std::pair<int32_t, LdsTile>& x, const std::pair<int32_t, LdsTile>& y) {
                             return x.first < y.first;
                           });
          if (!B.ltiles.empty()) built.push_back(std::move(B));
        }
        continue;
      }
      if (cls >= kClsLdsProlong) {
        // Every chunk of a transfer for every pair of variables in turn,
        // ordered by target block as the unit tiles of prolongation are.
        Built B{cls, {}, 0, {}, 0};
        for (size_t gi = 0; gi < p->groups.size(); ++gi) {
          const PlanGroup& g = p->groups[gi];
          if (g.points == 0 || ulds_of[gi][1] == 0 || ulds_launch(g) != cls) continue;
          for (int32_t t = 0; t < int32_t(g.tblocks.size()); ++t) {
            if (run_in[gi][t] != k) continue;
            for (int32_t c = ulds_of[gi][0]; c < ulds_of[gi][0] + ulds_of[gi][1]; ++c) {
              const LdsUnitChunk& ch = p->lds_uchunks[size_t(c)];
              B.lds_bytes = std::max(B.lds_bytes, int32_t(16 * ch.slots));
              int64_t targets = 0;
              for (int32_t w = ch.wave0; w < ch.wave0 + ch.nwave; ++w) {
                const LdsWave& wv = p->lds_waves[size_t(w)];
                const int rb = (p->wclasses[size_t(wv.wclass)].cls - kClsInterp) % 8;
                targets += int64_t(wv.nu) * (1 + (rb & 1)) * (1 + ((rb >> 1) & 1)) * (1 + ((rb >> 2) & 1));
              }
              for (int32_t v0 = 0; v0 < nvars; v0 += 2) {
                B.ltiles.push_back({g.tblocks[t], LdsTile{0, c, t, v0}});
                B.values += 2 * targets;
              }
            }
          }
        }
        std::stable_sort(B.ltiles.begin(), B.ltiles.end(),
                         [](const std::pair<int32_t, LdsTile>& x, const std::pair<int32_t, LdsTile>& y) {
                           return x.first < y.first;
                         });
        if (!B.ltiles.empty()) built.push_back(std::move(B));
        continue;
      }
      if (cls >= kClsLds) {
        // A transfer's chunks one after another, each for every pair of
        // variables in turn (the pairs of one point share its cache lines),
        // ordered by source block as the unit tiles of restriction are.
        Built B{cls, {}, 0, {}, 0};
        for (size_t gi = 0; gi < p->groups.size(); ++gi) {
          const PlanGroup& g = p->groups[gi];
          if (g.points == 0 || lds_of[gi][1] == 0 || lds_launch(g) != cls) continue;
          for (int32_t t = 0; t < int32_t(g.tblocks.size()); ++t) {
            if (run_in[gi][t] != k) continue;
            for (int32_t c = lds_of[gi][0]; c < lds_of[gi][0] + lds_of[gi][1]; ++c) {
              const LdsChunk& ch = p->lds_chunks[size_t(c)];
              B.lds_bytes = std::max(B.lds_bytes, int32_t(16 * ch.slots));
              for (int32_t v0 = 0; v0 < nvars; v0 += 2) {
                B.ltiles.push_back({g.sblocks[t], LdsTile{0, c, t, v0}});
                B.values += 2 * int64_t(ch.points);
              }
            }
          }
        }
        std::stable_sort(B.ltiles.begin(), B.ltiles.end(),
                         [](const std::pair<int32_t, LdsTile>& x, const std::pair<int32_t, LdsTile>& y) {
                           return x.first < y.first;
                         });
        if (!B.ltiles.empty()) built.push_back(std::move(B));
        continue;
      }
      std::vector<std::pair<int32_t, Tile>> tl;
      int64_t per = class_tile_elems(cls);
      int64_t values = 0;
      for (size_t gi = 0; gi < p->groups.size(); ++gi) {
        const PlanGroup& g = p->groups[gi];
        if (g.points == 0 || lds_of[gi][1] > 0 || str_of[gi][1] > 0) continue;
        if (g.cls == kClsInterp) {
          // A transfer's classes one after another, so that a merged
          // launch runs the tiles that share its windows together.
          // A tile takes nu consecutive units of the list (transfer, unit)
          // over a run of consecutive transfers of this stage group, so
          // that tiles stay full across transfers with few units.
          for (int32_t ci : g.wclasses) {
            const PlanWClass& c = p->wclasses[ci];
            if (launch_of(c.cls) != cls || lds_cls[size_t(ci)]) continue;
            // targets per unit: 2 along every dimension with R = 2
            const int rb = (c.cls - kClsInterp) % 8;
            const int64_t per_unit = (1 + (rb & 1)) * (1 + ((rb >> 1) & 1)) * (1 + ((rb >> 2) & 1));
            const int32_t nun = c.nunits();
            int32_t nu, nvt;
            const size_t fam = size_t((c.cls - kClsInterp) / 8);
            const int32_t T = p->fam_threads[fam];
            if (vfast) {
              // a pair tile's thread holds two variables
              const bool pairs = cls >= kClsMergedPair;
              const int32_t per_wg = pairs ? 2 * T : T;
              nvt = std::min(nvars, per_wg);
              if (p->fam_vars[fam] > 0)
                nvt = std::min(nvt, pairs ? std::max(2, p->fam_vars[fam] + (p->fam_vars[fam] & 1)) : p->fam_vars[fam]);
              // threads per unit: nvt, or (nvt + 1) / 2 over pairs
              nu = std::max(1, T / (pairs ? (nvt + 1) / 2 : nvt));
            } else {
              // units fastest: as many as the class has per transfer, the
              // rest of the workgroup over variables
              nu = std::min(nun, T);
              nvt = std::max(1, std::min(nvars, T / nu));
            }
            const int32_t nt = int32_t(g.tblocks.size());
            for (int32_t t = 0; t < nt;) {
              if (run_in[gi][t] != k) {
                ++t;
                continue;
              }
              int32_t t1 = t;
              while (t1 < nt && run_in[gi][t1] == k) ++t1;
              const int64_t count = int64_t(t1 - t) * nun;  // units of the run
              for (int32_t v0 = 0; v0 < nvars; v0 += nvt)
                for (int64_t i = 0; i < count; i += nu) {
                  const int32_t tt = t + int32_t(i / nun), uu = int32_t(i % nun);
                  const int32_t n = int32_t(std::min<int64_t>(nu, count - i));
                  // restriction by source block: the several small
                  // transfers that read one fine block (faces, edges,
                  // corners) run together and meet in one L2
                  const int32_t key = g.kind == 2 ? g.sblocks[tt] : g.tblocks[tt];
                  tl.push_back({key, Tile{0, ci, tt, uu, n, v0, std::min(nvt, nvars - v0), 0}});
                  values += int64_t(n) * std::min(nvt, nvars - v0) * per_unit;
                }
              t = t1;
            }
          }
        } else if (group_launch(g.cls) == cls) {
          const int64_t elems = int64_t(g.points) * nvars;
          const int32_t nt = int32_t(g.tblocks.size());
          // A copy group's transfers of at most half a tile, several to a
          // tile: consecutive transfers of this stage group (tile.nu of them).
          const int64_t pack = g.cls == kClsCopy ? per / elems : 1;
          if (pack >= 2) {
            for (int32_t t = 0; t < nt;) {
              if (run_in[gi][t] != k) {
                ++t;
                continue;
              }
              int32_t t1 = t;
              while (t1 < nt && run_in[gi][t1] == k && t1 - t < pack) ++t1;
              tl.push_back({copy_by_source ? g.sblocks[t] : g.tblocks[t], Tile{0, int32_t(gi), t, 0, t1 - t, 0, 0, 0}});
              values += elems * (t1 - t);
              t = t1;
            }
          } else {
            for (int32_t t = 0; t < nt; ++t) {
              if (run_in[gi][t] != k) continue;
              for (int64_t e = 0; e < elems; e += per)
                tl.push_back({g.cls == kClsCopy && copy_by_source ? g.sblocks[t] : g.tblocks[t],
                              Tile{0, int32_t(gi), t, int32_t(e), 0, 0, 0, 0}});
              values += elems;
            }
          }
        }
      }
      std::stable_sort(tl.begin(), tl.end(),
                       [](const std::pair<int32_t, Tile>& x, const std::pair<int32_t, Tile>& y) {
                         return x.first < y.first;
                       });
      if (!tl.empty()) built.push_back(Built{cls, std::move(tl), values, {}, 0});
    }
    for (Built& B : built) {
      if (!B.ltiles.empty()) {
        Launch L{B.cls, int64_t(p->lds_tiles.size()), int64_t(B.ltiles.size()), B.values, 0, B.lds_bytes, B.threads};
        for (const auto& e : B.ltiles) p->lds_tiles.push_back(e.second);
        ls.push_back(L);
        continue;
      }
      Launch L{B.cls, int64_t(p->tiles.size()), int64_t(B.tiles.size()), B.values, 0};
      {
        // an interpolation family's workgroup size
        int fam = -1;
        if (B.cls >= kClsMergedPair && B.cls < kClsMergedPair + kFamilies) fam = B.cls - kClsMergedPair;
        else if (B.cls >= kClsMerged && B.cls < kClsMerged + kFamilies) fam = B.cls - kClsMerged;
        else if (B.cls >= kClsInterp && B.cls < kClsInterp + 8 * kFamilies) fam = (B.cls - kClsInterp) / 8;
        if (fam >= 0) L.threads = p->fam_threads[size_t(fam)];
      }
      if (p->copy_wgs > 0 && B.cls == copy_launch(p->copy_per_thread) && int64_t(B.tiles.size()) >= 8)
        L.grid = int32_t(std::min<int64_t>(p->copy_wgs, int64_t(B.tiles.size())) / 8 * 8);
      for (const auto& e : B.tiles) p->tiles.push_back(e.second);
      ls.push_back(L);
    }
    order_launches(ls);
    p->launches.push_back(std::move(ls));
  }
  // A new state of the plan: batches built from an earlier one are stale.
  plan_register(p);
  void* old[] = {p->dtiles, p->dchunks, p->dlds_tiles, p->duchunks, p->dwaves, p->dunits, p->dschunks, p->dplanes};
  if (std::any_of(std::begin(old), std::end(old), [](void* q) { return q != nullptr; })) {
    // a previous tile list may still be read by queued fills
    HIPCHECK(hipDeviceSynchronize());
    for (void* q : old)
      if (q) HIPCHECK(hipFree(q));
    p->dtiles = nullptr;
    p->dchunks = nullptr;
    p->dlds_tiles = nullptr;
    p->duchunks = nullptr;
    p->dwaves = nullptr;
    p->dunits = nullptr;
    p->dschunks = nullptr;
    p->dplanes = nullptr;
  }
  auto upload = [](auto*& dev, const auto& host) {
    if (host.empty()) return;
    const size_t bytes = sizeof(host[0]) * host.size();
    HIPCHECK(hipMalloc(&dev, bytes));
    HIPCHECK(hipMemcpy(dev, host.data(), bytes, hipMemcpyHostToDevice));
  };
  upload(p->dtiles, p->tiles);
  upload(p->dchunks, p->lds_chunks);
  upload(p->dlds_tiles, p->lds_tiles);
  upload(p->duchunks, p->lds_uchunks);
  upload(p->dwaves, p->lds_waves);
  upload(p->dunits, p->lds_units);
  upload(p->dschunks, p->stream_chunks);
  upload(p->dplanes, p->stream_planes);
}

Job plan_job(const tah_plan* p, double* work, const double* factors, const double* src) {
  Job job;
  std::memset(&job, 0, sizeof(Job));
  job.groups = p->dgroups;
  job.wclasses = p->dwclasses;
  job.chunks = p->dchunks;
  job.uchunks = p->duchunks;
  job.waves = p->dwaves;
  job.units = p->dunits;
  job.schunks = p->dschunks;
  job.planes = p->dplanes;
  job.work = work;
  job.src = src ? src : work;
  job.factors = factors;
  for (int k = 0; k < 5; ++k) job.stride[k] = p->layout.stride[k];
  job.nvars = p->layout.nvars;
  job.vfast = p->layout.stride[3] < p->layout.stride[0] ? 1 : 0;
  job.dnv = FastDiv(uint32_t(p->layout.nvars));
  const int64_t* st = p->layout.stride;
  job.pairs = st[3] == 1 && p->layout.nvars % 2 == 0 && st[0] % 2 == 0 && st[1] % 2 == 0 && st[2] % 2 == 0 &&
                      st[4] % 2 == 0 && reinterpret_cast<uintptr_t>(work) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(job.src) % 16 == 0
                  ? 1
                  : 0;
  return job;
}

namespace {
bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  HIPCHECK(hipStreamIsCapturing(s, &st));
  return st != hipStreamCaptureStatusNone;
}
}  // namespace

void plan_begin(tah_plan* p, hipStream_t s) {
  if (!p->pending || s == p->last_stream || capturing(s)) return;
  HIPCHECK(hipStreamWaitEvent(s, p->done, 0));
}

void plan_end(tah_plan* p, hipStream_t s) {
  if (capturing(s)) return;
  HIPCHECK(hipEventRecord(p->done, s));
  p->last_stream = s;
  p->pending = true;
}

void plan_upload_job(tah_plan* p, const Job& job, hipStream_t s) {
  if (std::memcmp(&job, &p->last_job, sizeof(Job)) == 0) return;
  // From the plan's own copy: the host buffer outlives the copy.
  p->last_job = job;
  HIPCHECK(hipMemcpyAsync(p->djob, &p->last_job, sizeof(Job), hipMemcpyHostToDevice, s));
}

const std::vector<hipStream_t>& ctx_streams(tah_ctx* c) {
  if (c->streams.empty()) {
    HIPCHECK(hipSetDevice(c->device));
    for (int i = 0; i < 4; ++i) {
      hipStream_t st;
      HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      c->streams.push_back(st);
    }
  }
  return c->streams;
}

void launch_one(const Launch& L, hipStream_t s, const Job* jobs, const Tile* tiles, const LdsTile* ltiles) {
  if (L.lds > 0) launch_lds(L.cls, uint32_t(L.count), s, jobs, ltiles + L.first, L.lds, L.threads);
  else launch_class(L.cls, uint32_t(L.count), s, jobs, tiles + L.first, L.grid, L.threads);
}

void families_from_options(tah_plan* p) {
  for (int f = 0; f < kFamilies; ++f) {
    const bool restriction = kInterpTable[size_t(f)][1] != 0;
    p->fam_threads[size_t(f)] = kThreads;
    p->fam_merge[size_t(f)] = p->merge;
    // pairs by default only over an even count: over an odd one (pairs of
    // single loads, the last variable alone) prolongation measured slower
    // than single variables (the 64-forest batch: 24.4 against 22.7 ms)
    p->fam_pairs[size_t(f)] = restriction || p->layout.nvars % 2 ? 0 : p->interp_pairs;
    p->fam_lds[size_t(f)] = restriction ? ((p->lds & 4) ? 2 : (p->lds & 1)) : ((p->lds & 8) ? 2 : (p->lds >> 1) & 1);
    p->fam_vars[size_t(f)] = 0;
  }
}

void run_launches(tah_ctx* c, const std::vector<Launch>& ls, hipStream_t s, const Job* jobs,
                  const Tile* tiles, const LdsTile* ltiles, hipEvent_t* ev, bool fork) {
  if (!fork || ls.size() <= 1) {
    for (const Launch& L : ls) launch_one(L, s, jobs, tiles, ltiles);
    return;
  }
  const auto& streams = ctx_streams(c);
  HIPCHECK(hipEventRecord(ev[0], s));
  for (size_t i = 0; i < ls.size(); ++i) {
    hipStream_t si = streams[i % streams.size()];
    if (i < streams.size()) HIPCHECK(hipStreamWaitEvent(si, ev[0], 0));
    launch_one(ls[i], si, jobs, tiles, ltiles);
  }
  for (size_t i = 0; i < std::min(ls.size(), streams.size()); ++i) {
    HIPCHECK(hipEventRecord(ev[1 + i], streams[i]));
    HIPCHECK(hipStreamWaitEvent(s, ev[1 + i], 0));
  }
}

}  // namespace tah

tah_ctx::~tah_ctx() {
  for (hipStream_t st : streams) (void)hipStreamDestroy(st);
}

namespace tah {
void free_adjoint(tah_plan* p);
}

tah_plan::~tah_plan() {
  // queued work of this plan, and cached batches that use it, end first
  (void)hipDeviceSynchronize();
  tah::batch_forget_plan(this);
  tah::plan_unregister(this);
  tah::free_adjoint(this);
  for (hipEvent_t e : events) (void)hipEventDestroy(e);
  if (done) (void)hipEventDestroy(done);
  if (dtables) (void)hipFree(dtables);
  if (dtiles) (void)hipFree(dtiles);
  if (dchunks) (void)hipFree(dchunks);
  if (dlds_tiles) (void)hipFree(dlds_tiles);
  if (duchunks) (void)hipFree(duchunks);
  if (dwaves) (void)hipFree(dwaves);
  if (dunits) (void)hipFree(dunits);
  if (dschunks) (void)hipFree(dschunks);
  if (dplanes) (void)hipFree(dplanes);
  if (djob) (void)hipFree(djob);
  if (djob2) (void)hipFree(djob2);
  if (dhooks) (void)hipFree(dhooks);
  if (dhook_tiles) (void)hipFree(dhook_tiles);
}
