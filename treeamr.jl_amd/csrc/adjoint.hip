// The adjoint (transpose) of the fill.
//
// The fill is linear and runs in stages, each overwriting its targets from
// sources it never writes. Its transpose therefore runs the stages in
// reverse: in each, every source point gathers w * f * adj[target] over the
// transfers that read it, and then every target of the stage is zeroed.
//
// An interpolating transfer is a tensor product of one-dimensional stencils
// (weight (w1*w2)*w3 per tap), so its transpose is one too: three passes,
// each applying one dimension's transposed stencil, take the transfer's
// target box to its contribution over its source box, staged in scratch
// memory (z first, then y, then x). Each output value of a pass sums its
// (target row, tap) pairs in a fixed order from a host-built inverted stencil.
//
// The gather that follows is deterministic and uses no float atomics. Each
// source block's read region is cut, along the faces of the boxes its
// incoming transfers read, into cells; every point of a cell is read by one
// and the same list of transfers (group, then transfer order). One thread
// owns one source value and adds each transfer's contribution to it in that
// order -- a copy's target value(s) or an interpolation's staged value, both
// at an affine offset of the point -- so no two threads ever update one point
// and no barrier is needed. A copy's target is read by exactly one thread
// (its source point's), which zeroes it on the spot; the targets of the
// interpolations are zeroed after the gather.
//
// The reference has no adjoint: this is new, pinned by the dot-product
// identity and by the CPU transpose of the oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "api_common.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace tah {

// One dimension of a group's transposed stencil: for source row x in the
// read box lo..lo+ext-1, the (target row, tap) pairs reading it, packed
// r*16 + m, in list[start[x-lo]] .. list[start[x-lo+1]-1].
struct alignas(16) DevGroupT {
  int32_t lo[3], ext[3];
  const int32_t* start[3];
  const int32_t* list[3];
};

// One pass over one transfer: output point (i0, i1, i2) of box o, variable
// v, at out_off + ((i2*o1 + i1)*o0 + i0)*nvars + v, sums weight * input at
// in_off + isv*v + is[d]*i[d] (d not contracted) + is[D]*r over the pairs
// (r, tap) of the contracted dimension D's row i[D]. A pass tile takes rows
// row0 .. row0+kPassRows-1 of D and plane values pe0 .. pe0+kPassTile-1,
// a plane value being (v, then the lower, then the upper other dimension).
struct alignas(16) AdjPass {
  int64_t in_off, out_off;
  int32_t group, plane;  // plane: values per row of D
  int32_t o[3];
  int32_t is[3];
  int32_t isv, pad;
  FastDiv flo;  // division by the lower other dimension's extent
};

struct alignas(16) AdjPassTile {
  int32_t item, row0, pe0, pad;
};

// What one transfer adds to source point p of variable v: the sum over
// rows r < c of buffer[base + a.p + av*v + rs.r], times the factor.
// Buffer 0 is the working array (a copy: c > 1 along a clamped dimension),
// buffer 1 the staged contributions (c = 1).
struct alignas(16) AdjEntry {
  int64_t base;
  int32_t a[3];
  int32_t av;
  int32_t c[3];
  int32_t rs[3];
  int32_t buf;
  int32_t factorcol;
};

// A box of one source block whose points are all read by the same
// transfers: entries[first .. first+count-1]. `target`: the points are also
// targets of the stage, read by a transfer onto themselves (stage P's
// projection), so they start from zero rather than from their own adjoint.
struct alignas(16) AdjCell {
  int32_t sblock, first, count, npts;
  int32_t lo[3], ext[3];
  FastDiv dx[2];  // division by ext[0], ext[1]
  int32_t target, pad[3];
};

// A workgroup's share: kAdjTile values of cell (or pass item) `item` from e0.
struct alignas(8) AdjTile {
  int32_t item, e0;
};

constexpr int kAdjPerThread = 4;
constexpr int kAdjTile = kThreads * kAdjPerThread;
constexpr int kPassRows = 8;              // output rows of one adjoint pass tile
constexpr int kPassRowsAhead = 4;         // input rows loaded ahead of their taps
constexpr int kPassTile = 2 * kThreads;   // and its plane values

struct AdjStage {
  int64_t pass_tiles[3][2];  // per pass (z, y, x): first tile, count
  int64_t gather[2];
  int64_t zero[2];
  int64_t scratch_off[3];   // outputs of the passes, in the scratch buffer
  // Algorithmic bytes of the stage's launches (passes z, y, x, gather,
  // zeroing): 8 per value read and 8 per value written, once each -- a
  // pass reads its input box and writes its output box; the gather reads
  // and writes each source value it owns (writes only, where the cell's
  // points are targets too) and reads each entry's values, zeroing a copy's
  // targets as it reads them; the zeroing writes the interpolated targets.
  int64_t bytes[5];
};

struct AdjointTables {
  char* dmem = nullptr;
  double* scratch = nullptr;
  int64_t scratch_values = 0;
  DevGroupT* dgt = nullptr;
  AdjPass* dpass = nullptr;
  AdjPassTile* dptiles = nullptr;
  AdjCell* dcells = nullptr;
  AdjTile* dtiles = nullptr;
  AdjEntry* dentries = nullptr;
  Tile* dzero = nullptr;
  std::vector<AdjStage> stages;
  // fork and join of a stage's zeroing on a second stream (made at set-up,
  // so that a run creates nothing)
  hipEvent_t zero_fork = nullptr, zero_join = nullptr;
};

void free_adjoint(tah_plan* p) {
  if (!p->adj) return;
  if (p->adj->dmem) (void)hipFree(p->adj->dmem);
  if (p->adj->scratch) (void)hipFree(p->adj->scratch);
  if (p->adj->zero_fork) (void)hipEventDestroy(p->adj->zero_fork);
  if (p->adj->zero_join) (void)hipEventDestroy(p->adj->zero_join);
  delete p->adj;
  p->adj = nullptr;
}

// Pass D (2: z, 1: y, 0: x) of every interpolating transfer of a stage.
// A tile takes kPassRows output rows of D for kPassTile plane values, two
// per thread: one aligned 16-byte pair (variables v, v+1 of one point)
// where input and output hold variables innermost in even number, else two
// values kThreads apart. Scatter form: the tile walks the input rows that
// reach any of its output rows once each, in ascending order, loads the
// thread's values of each once, and adds weight * value into every output
// row of the tile whose stencil holds it (registers). Each output thus sums
// its (input row, tap) pairs in ascending input row, the order of its
// inverted stencil, with one load per input row rather than one per pair.
// The rows, their stencil starts and weights are one for the whole
// workgroup: scalar loads.
template <int D, bool PAIR>
__device__ __forceinline__ void pass_tile(const AdjPass& it, const Job& job, const DevGroup& g, const DevGroupT& T,
                                          const double* __restrict__ in, double* __restrict__ o, uint32_t pe0,
                                          int32_t x0) {
  typedef double pair_t __attribute__((ext_vector_type(2)));
  constexpr int DL = D == 0 ? 1 : 0;  // the lower other dimension
  constexpr int DU = D == 2 ? 1 : 2;  // the upper one
  const int32_t nvars = job.nvars;
  const uint32_t plane = uint32_t(it.plane);
  const uint32_t pe[2] = {PAIR ? pe0 + 2u * threadIdx.x : pe0 + threadIdx.x,
                          PAIR ? pe0 + 2u * threadIdx.x + 1u : pe0 + kThreads + threadIdx.x};
  if (pe[0] >= plane) return;
  const bool ok1 = pe[1] < plane;
  int32_t b[2], ob[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const uint32_t e = s == 1 && !ok1 ? pe[0] : pe[s];
    const uint32_t r = job.dnv.div(e);
    const int32_t v = int32_t(e - r * uint32_t(nvars));
    const uint32_t q = it.flo.div(r);
    int32_t i[3];
    i[DL] = int32_t(r - q * uint32_t(it.o[DL]));
    i[DU] = int32_t(q);
    b[s] = it.isv * v + it.is[DL] * i[DL] + it.is[DU] * i[DU];
    ob[s] = (((D == 2 ? 0 : i[2]) * it.o[1] + (D == 1 ? 0 : i[1])) * it.o[0] + (D == 0 ? 0 : i[0])) * nvars + v;
  }
  const int32_t ostep = (D == 0 ? 1 : D == 1 ? it.o[0] : it.o[0] * it.o[1]) * nvars;
  const auto start = constant(T.start[D]);
  const auto list = constant(T.list[D]);
  const auto ss = constant(g.srcstart[D]);
  const auto w = constant(g.weights[D]);
  const int32_t order = uniform(g.order[D]);
  const int32_t isd = it.is[D];
  const int32_t nr = min(kPassRows, uniform(it.o[D]) - x0);
  // The input rows that reach the tile: each output row's pairs run in
  // ascending input row, so its first and last pair bound them.
  int32_t rlo = INT32_MAX, rhi = -1;
  for (int k = 0; k < nr; ++k) {
    const int32_t j0 = start[x0 + k], j1 = start[x0 + k + 1];
    if (j0 < j1) {
      rlo = min(rlo, list[j0] >> 4);
      rhi = max(rhi, list[j1 - 1] >> 4);
    }
  }
  double a0[kPassRows], a1[kPassRows];
#pragma unroll
  for (int k = 0; k < kPassRows; ++k) a0[k] = a1[k] = 0.0;
  // Input row rr reaches output row x0 + k with tap m = mx0 + k - ss[rr].
  const int32_t mx0 = x0 + uniform(T.lo[D]);
  // Input rows are loaded kPassRowsAhead rows ahead of their taps, into a
  // ring (one slot per row of a group of kPassRowsAhead), so that a row's
  // load latency overlaps the rows before it.
  auto load = [&](int32_t rr, double& u0, double& u1) {
    if constexpr (PAIR) {
      const pair_t uu = global(reinterpret_cast<const pair_t*>(in))[(b[0] + isd * rr) >> 1];
      u0 = uu.x;
      u1 = uu.y;
    } else {
      u0 = global(in)[b[0] + isd * rr];
      u1 = global(in)[b[1] + isd * rr];
    }
  };
  double ring0[kPassRowsAhead], ring1[kPassRowsAhead];
  if (rlo <= rhi) {  // (no input rows: rlo is INT32_MAX)
#pragma unroll
    for (int q = 0; q < kPassRowsAhead; ++q)
      if (q <= rhi - rlo) load(rlo + q, ring0[q], ring1[q]);
  }
  for (int32_t r0 = rlo; r0 <= rhi; r0 += kPassRowsAhead) {
#pragma unroll
    for (int q = 0; q < kPassRowsAhead; ++q) {
      const int32_t rr = r0 + q;
      if (rr > rhi) break;
      const double u0 = ring0[q], u1 = ring1[q];
      if (kPassRowsAhead <= rhi - rr) load(rr + kPassRowsAhead, ring0[q], ring1[q]);
      const int32_t m0 = mx0 - ss[rr];
      const int32_t wr = order * rr;
#pragma unroll
      for (int k = 0; k < kPassRows; ++k) {
        const int32_t m = m0 + k;
        if (k < nr && m >= 0 && m < order) {
          const double ww = w[wr + m];
          a0[k] = __dadd_rn(a0[k], __dmul_rn(ww, u0));
          a1[k] = __dadd_rn(a1[k], __dmul_rn(ww, u1));
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kPassRows; ++k) {
    if (k >= nr) break;
    if constexpr (PAIR) {
      pair_t res;
      res.x = a0[k];
      res.y = a1[k];
      reinterpret_cast<pair_t*>(o)[(ob[0] + ostep * (x0 + k)) >> 1] = res;
    } else {
      o[ob[0] + ostep * (x0 + k)] = a0[k];
      if (ok1) o[ob[1] + ostep * (x0 + k)] = a1[k];
    }
  }
}

template <int D>
__global__ void __launch_bounds__(kThreads) adjoint_pass_kernel(const Job* __restrict__ jobs,
                                                                const DevGroupT* __restrict__ gts,
                                                                const AdjPass* __restrict__ items,
                                                                const AdjPassTile* __restrict__ tiles,
                                                                uint32_t ntiles,
                                                                const double* __restrict__ scratch_in,
                                                                double* __restrict__ out) {
  const AdjPassTile tile = tiles[xcd_tile(blockIdx.x, ntiles)];
  const AdjPass& it = items[uniform(tile.item)];
  const Job& job = jobs[0];
  const int32_t gi = uniform(it.group);
  const DevGroup& g = job.groups[gi];
  const DevGroupT& T = gts[gi];
  const double* __restrict__ in = (D == 2 ? job.work : scratch_in) + it.in_off;
  double* __restrict__ o = out + it.out_off;
  // The staged passes hold variables innermost; the first reads the
  // working array, paired where its layout allows.
  if ((job.nvars & 1) == 0 && (D < 2 || job.pairs))
    pass_tile<D, true>(it, job, g, T, in, o, uint32_t(tile.pe0), uniform(tile.row0));
  else
    pass_tile<D, false>(it, job, g, T, in, o, uint32_t(tile.pe0), uniform(tile.row0));
}

// The gather with variables innermost and even in number: values 2e and
// 2e+1 are one point's neighbouring variables, 16 contiguous bytes wherever
// an entry reads them (copy targets and staged contributions alike have
// even offsets), so each thread moves both with one access.
__device__ __forceinline__ void gather_paired(const Job& job, const AdjCell& c, const AdjTile& tile,
                                              const AdjEntry* __restrict__ entries, double* __restrict__ staged) {
  typedef double pair_t __attribute__((ext_vector_type(2)));
  const pair_t zero2 = {0.0, 0.0};
  constexpr int K = kAdjPerThread / 2;
  const int32_t s0 = int32_t(job.stride[0]), s1 = int32_t(job.stride[1]), s2 = int32_t(job.stride[2]);
  const int32_t nvars = job.nvars;
  const uint32_t total = uint32_t(c.npts) * uint32_t(nvars);
  double* __restrict__ sb = job.work + job.stride[4] * c.sblock;
  int32_t p[K][3], var[K], off[K];
  bool ok[K];
  pair_t tot[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint32_t e = uint32_t(tile.e0) + 2u * (uint32_t(k * kThreads) + threadIdx.x);
    ok[k] = e < total;
    const uint32_t ee = ok[k] ? e : 0;
    const uint32_t r = job.dnv.div(ee);
    var[k] = int32_t(ee - r * uint32_t(nvars));
    const uint32_t q = c.dx[0].div(r);
    const uint32_t zq = c.dx[1].div(q);
    p[k][0] = c.lo[0] + int32_t(r - q * uint32_t(c.ext[0]));
    p[k][1] = c.lo[1] + int32_t(q - zq * uint32_t(c.ext[1]));
    p[k][2] = c.lo[2] + int32_t(zq);
    off[k] = var[k] + s0 * p[k][0] + s1 * p[k][1] + s2 * p[k][2];
    tot[k] = zero2;
    // (each source value is read and written by this thread alone, once:
    // both streamed past the caches)
    if (ok[k] && !c.target) tot[k] = __builtin_nontemporal_load(reinterpret_cast<const pair_t*>(sb + off[k]));
  }
  const int32_t first = uniform(c.first), count = uniform(c.count);
  for (int32_t q = 0; q < count; ++q) {
    const AdjEntry& E = entries[first + q];
    const bool copy = E.buf == 0;
    double* __restrict__ src = (copy ? job.work : staged) + E.base;
    const int32_t c0 = E.c[0], c1 = E.c[1], c2 = E.c[2];
    const bool single = c0 * c1 * c2 == 1;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!ok[k]) continue;
      const int32_t o = E.a[0] * p[k][0] + E.a[1] * p[k][1] + E.a[2] * p[k][2] + E.av * var[k];
      pair_t acc;
      // A copy's targets are read once and zeroed, a staged value read once:
      // streamed past the caches, as the source values are (the large
      // forest's gather: 16.4–16.5 against 17.5–17.6 ms with cached loads and
      // stores, on one box).
      if (single) {
        if (copy) {
          acc = __builtin_nontemporal_load(reinterpret_cast<const pair_t*>(src + o));
          __builtin_nontemporal_store(zero2, reinterpret_cast<pair_t*>(src + o));
        } else {
          acc = __builtin_nontemporal_load(reinterpret_cast<const pair_t*>(src + o));
        }
      } else {
        acc = zero2;
        for (int32_t r2 = 0; r2 < c2; ++r2)
          for (int32_t r1 = 0; r1 < c1; ++r1)
            for (int32_t r0 = 0; r0 < c0; ++r0) {
              const int32_t oo = o + E.rs[0] * r0 + E.rs[1] * r1 + E.rs[2] * r2;
              const pair_t x = __builtin_nontemporal_load(reinterpret_cast<const pair_t*>(src + oo));
              acc.x = __dadd_rn(acc.x, x.x);
              acc.y = __dadd_rn(acc.y, x.y);
              if (copy) __builtin_nontemporal_store(zero2, reinterpret_cast<pair_t*>(src + oo));
            }
      }
      if (E.factorcol) {
        const double* f = job.factors + int64_t(nvars) * (E.factorcol - 1) + var[k];
        acc.x = __dmul_rn(acc.x, f[0]);
        acc.y = __dmul_rn(acc.y, f[1]);
      }
      tot[k].x = __dadd_rn(tot[k].x, acc.x);
      tot[k].y = __dadd_rn(tot[k].y, acc.y);
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (ok[k]) __builtin_nontemporal_store(tot[k], reinterpret_cast<pair_t*>(sb + off[k]));
}

__global__ void __launch_bounds__(kThreads) adjoint_gather_kernel(const Job* __restrict__ jobs,
                                                                  const AdjCell* __restrict__ cells,
                                                                  const AdjTile* __restrict__ tiles,
                                                                  const AdjEntry* __restrict__ entries,
                                                                  uint32_t ntiles,
                                                                  double* __restrict__ staged) {
  const AdjTile tile = tiles[xcd_tile(blockIdx.x, ntiles)];
  const AdjCell& c = cells[uniform(tile.item)];
  const Job& job = jobs[0];
  const int32_t s0 = int32_t(job.stride[0]), s1 = int32_t(job.stride[1]),
                s2 = int32_t(job.stride[2]), s3 = int32_t(job.stride[3]);
  const int32_t nvars = job.nvars;
  if (job.pairs) {
    gather_paired(job, c, tile, entries, staged);
    return;
  }
  const uint32_t npts = uint32_t(c.npts);
  const uint32_t total = npts * uint32_t(nvars);
  double* __restrict__ sb = job.work + job.stride[4] * c.sblock;
  int32_t p[kAdjPerThread][3], var[kAdjPerThread], off[kAdjPerThread];
  bool ok[kAdjPerThread];
  double tot[kAdjPerThread];
#pragma unroll
  for (int k = 0; k < kAdjPerThread; ++k) {
    const uint32_t e = uint32_t(tile.e0) + uint32_t(k * kThreads) + threadIdx.x;
    ok[k] = e < total;
    const uint32_t ee = ok[k] ? e : 0;
    uint32_t r;
    if (job.vfast) {
      r = job.dnv.div(ee);
      var[k] = int32_t(ee - r * uint32_t(nvars));
    } else {
      var[k] = int32_t(ee / npts);
      r = ee - uint32_t(var[k]) * npts;
    }
    const uint32_t q = c.dx[0].div(r);
    const uint32_t zq = c.dx[1].div(q);
    p[k][0] = c.lo[0] + int32_t(r - q * uint32_t(c.ext[0]));
    p[k][1] = c.lo[1] + int32_t(q - zq * uint32_t(c.ext[1]));
    p[k][2] = c.lo[2] + int32_t(zq);
    off[k] = s3 * var[k] + s0 * p[k][0] + s1 * p[k][1] + s2 * p[k][2];
    tot[k] = ok[k] && !c.target ? __builtin_nontemporal_load(sb + off[k]) : 0.0;
  }
  const int32_t first = uniform(c.first), count = uniform(c.count);
  for (int32_t q = 0; q < count; ++q) {
    const AdjEntry& E = entries[first + q];
    const bool copy = E.buf == 0;
    // A copy's targets are read by this thread alone: it zeroes them too.
    double* __restrict__ src = (copy ? job.work : staged) + E.base;
    const int32_t c0 = E.c[0], c1 = E.c[1], c2 = E.c[2];
    const bool single = c0 * c1 * c2 == 1;
#pragma unroll
    for (int k = 0; k < kAdjPerThread; ++k) {
      if (!ok[k]) continue;
      const int32_t o = E.a[0] * p[k][0] + E.a[1] * p[k][1] + E.a[2] * p[k][2] + E.av * var[k];
      double acc;
      if (single) {
        if (copy) {
          acc = __builtin_nontemporal_load(src + o);
          __builtin_nontemporal_store(0.0, src + o);
        } else {
          acc = __builtin_nontemporal_load(src + o);
        }
      } else {
        acc = 0.0;
        for (int32_t r2 = 0; r2 < c2; ++r2)
          for (int32_t r1 = 0; r1 < c1; ++r1)
            for (int32_t r0 = 0; r0 < c0; ++r0) {
              const int32_t oo = o + E.rs[0] * r0 + E.rs[1] * r1 + E.rs[2] * r2;
              acc = __dadd_rn(acc, __builtin_nontemporal_load(src + oo));
              if (copy) __builtin_nontemporal_store(0.0, src + oo);
            }
      }
      if (E.factorcol) acc = __dmul_rn(acc, job.factors[var[k] + int64_t(nvars) * (E.factorcol - 1)]);
      tot[k] = __dadd_rn(tot[k], acc);
    }
  }
#pragma unroll
  for (int k = 0; k < kAdjPerThread; ++k)
    if (ok[k]) __builtin_nontemporal_store(tot[k], sb + off[k]);
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
  if (job.pairs) {
    // variables innermost in even number, 16-byte aligned: values 2e and
    // 2e+1 are one point's neighbouring variables, zeroed by one 16-byte
    // store that streams past the caches (nothing reads them again)
    typedef double pair_t __attribute__((ext_vector_type(2)));
    const pair_t zero2 = {0.0, 0.0};
    for (int r = 0; r < kCopyPerThread / 2; ++r) {
      const uint32_t e = uint32_t(tile.e0) + 2u * (uint32_t(r * kThreads) + threadIdx.x);
      if (e >= total) break;
      int32_t i, j, k, v;
      split_element(g, job, e, i, j, k, v);
      __builtin_nontemporal_store(
          zero2, reinterpret_cast<pair_t*>(tbase + (v + s0 * (g.tfirst[0] + i) + s1 * (g.tfirst[1] + j) +
                                                    s2 * (g.tfirst[2] + k))));
    }
    return;
  }
  for (int r = 0; r < kCopyPerThread; ++r) {
    const uint32_t e = uint32_t(tile.e0) + uint32_t(r * kThreads) + threadIdx.x;
    if (e >= total) break;
    int32_t i, j, k, v;
    split_element(g, job, e, i, j, k, v);
    tbase[s3 * v + s0 * (g.tfirst[0] + i) + s1 * (g.tfirst[1] + j) + s2 * (g.tfirst[2] + k)] = 0.0;
  }
}

namespace {

struct Box {
  int32_t lo[3], hi[3];  // half-open
};

struct Reader {
  int32_t group, transfer;
  AdjEntry entry;
  bool self;  // a copy of every point onto itself
};

// Cut the read region of one source block (the boxes of the groups that
// read it) into cells along every box face; each cell with at least one
// reader gets the readers' entries in their given order.
void cut_cells(int32_t sblock, const std::vector<Reader>& readers, const std::vector<Box>& boxes,
               std::vector<AdjCell>& cells, std::vector<AdjEntry>& entries) {
  std::vector<int32_t> cuts[3];
  for (const auto& r : readers)
    for (int d = 0; d < 3; ++d) {
      cuts[d].push_back(boxes[r.group].lo[d]);
      cuts[d].push_back(boxes[r.group].hi[d]);
    }
  for (auto& c : cuts) {
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
  }
  for (size_t iz = 0; iz + 1 < cuts[2].size(); ++iz)
    for (size_t iy = 0; iy + 1 < cuts[1].size(); ++iy)
      for (size_t ix = 0; ix + 1 < cuts[0].size(); ++ix) {
        const int32_t lo[3] = {cuts[0][ix], cuts[1][iy], cuts[2][iz]};
        const int32_t hi[3] = {cuts[0][ix + 1], cuts[1][iy + 1], cuts[2][iz + 1]};
        const int32_t first = int32_t(entries.size());
        bool target = false;
        for (const auto& r : readers) {
          const Box& b = boxes[r.group];
          bool in = true;
          for (int d = 0; d < 3; ++d) in = in && b.lo[d] <= lo[d] && hi[d] <= b.hi[d];
          if (!in) continue;
          entries.push_back(r.entry);
          target = target || r.self;
        }
        const int32_t count = int32_t(entries.size()) - first;
        if (count == 0) continue;
        AdjCell c{};
        c.sblock = sblock;
        c.first = first;
        c.count = count;
        for (int d = 0; d < 3; ++d) {
          c.lo[d] = lo[d];
          c.ext[d] = hi[d] - lo[d];
        }
        c.npts = c.ext[0] * c.ext[1] * c.ext[2];
        c.dx[0] = FastDiv(uint32_t(c.ext[0]));
        c.dx[1] = FastDiv(uint32_t(c.ext[1]));
        c.target = target ? 1 : 0;
        cells.push_back(c);
      }
}

void push_tiles(std::vector<AdjTile>& tiles, int32_t item, int64_t elems) {
  if (elems > INT32_MAX) fail("an adjoint work item exceeds 2^31 values", kUnsupported);
  for (int64_t e = 0; e < elems; e += kAdjTile) tiles.push_back(AdjTile{item, int32_t(e)});
}

void build_adjoint_tables(tah_plan* p) {
  auto* A = new AdjointTables();
  p->adj = A;
  std::vector<char> host;
  auto put = [&](const void* src, int64_t bytes) {
    int64_t off = (int64_t(host.size()) + 255) / 256 * 256;
    host.resize(off + std::max<int64_t>(bytes, 8));
    if (bytes) std::memcpy(host.data() + off, src, bytes);
    return off;
  };
  const tah_layout& L = p->layout;
  const int32_t nvars = L.nvars;
  // Transposed stencils and read boxes of every group.
  struct GT {
    int32_t lo[3], ext[3];
    int64_t start[3], list[3];
  };
  std::vector<GT> gts(p->groups.size());
  std::vector<Box> boxes(p->groups.size());
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
      boxes[gi].lo[d] = lo;
      boxes[gi].hi[d] = hi + 1;
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
  std::vector<AdjPass> passes;
  std::vector<AdjPassTile> ptiles;
  std::vector<AdjTile> tiles;
  std::vector<AdjCell> cells;
  std::vector<AdjEntry> entries;
  std::vector<Tile> zeros;
  int64_t scratch_max = 0;
  for (int32_t st = 0; st < p->nstages; ++st) {
    AdjStage S{};
    std::vector<std::array<int32_t, 2>> interp;  // (group, transfer)
    std::vector<std::array<int32_t, 3>> es;      // sblock, group, transfer
    S.zero[0] = int64_t(zeros.size());
    for (size_t gi = 0; gi < p->groups.size(); ++gi) {
      const PlanGroup& g = p->groups[gi];
      if (g.stage != st || g.points == 0) continue;
      int32_t ca[3], cs[3];
      const bool copy = affine_copy(g, ca, cs, 1);
      for (int32_t t = 0; t < int32_t(g.tblocks.size()); ++t) {
        es.push_back({g.sblocks[t], int32_t(gi), t});
        if (copy) continue;  // the gather zeroes a copy's targets
        interp.push_back({int32_t(gi), t});
        const int64_t elems = int64_t(g.points) * nvars;
        S.bytes[4] += 8 * elems;
        for (int64_t e = 0; e < elems; e += kCopyTile) zeros.push_back(Tile{0, int32_t(gi), t, int32_t(e), 0, 0, 0, 0});
      }
    }
    S.zero[1] = int64_t(zeros.size()) - S.zero[0];
    // Scratch: the outputs of the z, y and x passes of every interpolating
    // transfer, one after the other.
    std::vector<int64_t> off_of(interp.size() * 3);
    int64_t tot[3] = {0, 0, 0};
    for (size_t k = 0; k < interp.size(); ++k) {
      const PlanGroup& g = p->groups[interp[k][0]];
      const Box& b = boxes[interp[k][0]];
      const int64_t e0 = b.hi[0] - b.lo[0], e1 = b.hi[1] - b.lo[1], e2 = b.hi[2] - b.lo[2];
      const int64_t sz[3] = {int64_t(g.n[0]) * g.n[1] * e2, int64_t(g.n[0]) * e1 * e2, e0 * e1 * e2};
      for (int q = 0; q < 3; ++q) {
        off_of[3 * k + q] = tot[q];
        tot[q] += sz[q] * nvars;
      }
    }
    S.scratch_off[0] = 0;
    S.scratch_off[1] = tot[0];
    S.scratch_off[2] = tot[0] + tot[1];
    scratch_max = std::max(scratch_max, tot[0] + tot[1] + tot[2]);
    for (int q = 0; q < 3; ++q) {
      S.pass_tiles[q][0] = int64_t(ptiles.size());
      for (size_t k = 0; k < interp.size(); ++k) {
        const int32_t gi = interp[k][0], t = interp[k][1];
        const PlanGroup& g = p->groups[gi];
        const Box& b = boxes[gi];
        const int32_t e[3] = {b.hi[0] - b.lo[0], b.hi[1] - b.lo[1], b.hi[2] - b.lo[2]};
        AdjPass P{};
        P.group = gi;
        if (q == 0) {  // z: target box of the working array -> (n0, n1, e2)
          P.in_off = L.stride[4] * g.tblocks[t] + L.stride[0] * g.tfirst[0] + L.stride[1] * g.tfirst[1] +
                     L.stride[2] * g.tfirst[2];
          for (int d = 0; d < 3; ++d) P.is[d] = int32_t(L.stride[d]);
          P.isv = int32_t(L.stride[3]);
          P.o[0] = g.n[0];
          P.o[1] = g.n[1];
          P.o[2] = e[2];
        } else if (q == 1) {  // y: (n0, n1, e2) -> (n0, e1, e2)
          P.in_off = off_of[3 * k + 0];
          P.is[0] = nvars;
          P.is[1] = nvars * g.n[0];
          P.is[2] = nvars * g.n[0] * g.n[1];
          P.isv = 1;
          P.o[0] = g.n[0];
          P.o[1] = e[1];
          P.o[2] = e[2];
        } else {  // x: (n0, e1, e2) -> (e0, e1, e2)
          P.in_off = S.scratch_off[1] + off_of[3 * k + 1];
          P.is[0] = nvars;
          P.is[1] = nvars * g.n[0];
          P.is[2] = nvars * g.n[0] * e[1];
          P.isv = 1;
          P.o[0] = e[0];
          P.o[1] = e[1];
          P.o[2] = e[2];
        }
        P.out_off = S.scratch_off[q] + off_of[3 * k + q];
        {
          const int64_t in = q == 0 ? int64_t(g.n[0]) * g.n[1] * g.n[2]
                             : q == 1 ? int64_t(g.n[0]) * g.n[1] * e[2]
                                      : int64_t(g.n[0]) * e[1] * e[2];
          S.bytes[q] += 8 * nvars * (in + int64_t(P.o[0]) * P.o[1] * P.o[2]);
        }
        const int D = 2 - q, DL = D == 0 ? 1 : 0, DU = D == 2 ? 1 : 2;
        const int64_t plane = int64_t(P.o[DL]) * P.o[DU] * nvars;
        if (plane > INT32_MAX) fail("an adjoint pass plane exceeds 2^31 values", kUnsupported);
        P.plane = int32_t(plane);
        P.flo = FastDiv(uint32_t(std::max(1, P.o[DL])));
        passes.push_back(P);
        const int32_t item = int32_t(passes.size() - 1);
        for (int32_t r0 = 0; r0 < P.o[D]; r0 += kPassRows)
          for (int64_t e = 0; e < plane; e += kPassTile) ptiles.push_back(AdjPassTile{item, r0, int32_t(e), 0});
      }
      S.pass_tiles[q][1] = int64_t(ptiles.size()) - S.pass_tiles[q][0];
    }
    // The gather: cells of every source block, entries in (group, transfer) order.
    std::stable_sort(es.begin(), es.end(), [](const std::array<int32_t, 3>& a, const std::array<int32_t, 3>& b) {
      return a[0] < b[0];
    });
    std::vector<int64_t> interp_index(p->groups.size(), -1);  // first interp slot of a group
    for (size_t k = 0; k < interp.size(); ++k)
      if (interp_index[interp[k][0]] < 0) interp_index[interp[k][0]] = int64_t(k);
    S.gather[0] = int64_t(tiles.size());
    std::vector<Reader> readers;
    for (size_t i = 0; i < es.size();) {
      size_t j = i;
      readers.clear();
      for (; j < es.size() && es[j][0] == es[i][0]; ++j) {
        const int32_t gi = es[j][1], t = es[j][2];
        const PlanGroup& g = p->groups[gi];
        const Box& b = boxes[gi];
        Reader R{gi, t, AdjEntry{}, false};
        AdjEntry& E = R.entry;
        E.factorcol = g.factorcol;
        int32_t ca[3], cs[3];
        if (affine_copy(g, ca, cs, 1)) {
          R.self = g.tblocks[t] == g.sblocks[t];
          for (int d = 0; d < 3; ++d) R.self = R.self && ca[d] == g.tfirst[d] && (cs[d] == 1 || g.n[d] == 1);
          // target row r = (p - ca)*cs along a stepped dimension; all rows
          // of a clamped one
          E.buf = 0;
          E.base = L.stride[4] * g.tblocks[t];
          E.av = int32_t(L.stride[3]);
          for (int d = 0; d < 3; ++d) {
            E.base += L.stride[d] * g.tfirst[d];
            if (cs[d] == 0) {
              E.a[d] = 0;
              E.c[d] = g.n[d];
              E.rs[d] = int32_t(L.stride[d]);
            } else {
              // the box is the exact image of the rows: every point has its row
              for (int32_t x : {b.lo[d], b.hi[d] - 1}) {
                const int32_t r = (x - ca[d]) * cs[d];
                if (r < 0 || r >= g.n[d]) fail("adjoint: a copy's read box exceeds its rows");
              }
              E.a[d] = int32_t(L.stride[d]) * cs[d];
              E.base -= int64_t(L.stride[d]) * cs[d] * ca[d];
              E.c[d] = 1;
              E.rs[d] = 0;
            }
          }
        } else {
          // staged value of point p at ((p2-lo2)*e1 + (p1-lo1))*e0 + (p0-lo0)
          const int64_t k = interp_index[gi] + t;  // transfers of a group are consecutive
          const int32_t e0 = b.hi[0] - b.lo[0], e1 = b.hi[1] - b.lo[1];
          E.buf = 1;
          E.av = 1;
          E.a[0] = nvars;
          E.a[1] = nvars * e0;
          E.a[2] = nvars * e0 * e1;
          E.base = S.scratch_off[2] + off_of[3 * k + 2] -
                   (int64_t(E.a[0]) * b.lo[0] + int64_t(E.a[1]) * b.lo[1] + int64_t(E.a[2]) * b.lo[2]);
          for (int d = 0; d < 3; ++d) {
            E.c[d] = 1;
            E.rs[d] = 0;
          }
        }
        readers.push_back(R);
      }
      const size_t c0 = cells.size();
      cut_cells(es[i][0], readers, boxes, cells, entries);
      for (size_t c = c0; c < cells.size(); ++c) {
        const AdjCell& C = cells[c];
        push_tiles(tiles, int32_t(c), int64_t(C.npts) * nvars);
        int64_t per = C.target ? 8 : 16;
        for (int32_t q = C.first; q < C.first + C.count; ++q) {
          const AdjEntry& E = entries[size_t(q)];
          per += E.buf == 0 ? 16 * int64_t(E.c[0]) * E.c[1] * E.c[2] : 8;
        }
        S.bytes[3] += per * int64_t(C.npts) * nvars;
      }
      i = j;
    }
    S.gather[1] = int64_t(tiles.size()) - S.gather[0];
    if (S.gather[1] > INT32_MAX / 2 || S.zero[1] > INT32_MAX / 2)
      fail("too many adjoint tiles in one stage", kUnsupported);
    A->stages.push_back(S);
  }
  if (cells.size() > size_t(INT32_MAX) || passes.size() > size_t(INT32_MAX) || entries.size() > size_t(INT32_MAX))
    fail("adjoint tables too large", kUnsupported);
  const int64_t off_pass = put(passes.data(), int64_t(sizeof(AdjPass) * passes.size()));
  const int64_t off_ptiles = put(ptiles.data(), int64_t(sizeof(AdjPassTile) * ptiles.size()));
  const int64_t off_cells = put(cells.data(), int64_t(sizeof(AdjCell) * cells.size()));
  const int64_t off_tiles = put(tiles.data(), int64_t(sizeof(AdjTile) * tiles.size()));
  const int64_t off_entries = put(entries.data(), int64_t(sizeof(AdjEntry) * entries.size()));
  const int64_t off_zero = put(zeros.data(), int64_t(sizeof(Tile) * zeros.size()));
  const int64_t off_gt = put(nullptr, int64_t(sizeof(DevGroupT) * std::max<size_t>(1, gts.size())));
  HIPCHECK(hipMalloc(&A->dmem, host.size()));
  A->scratch_values = scratch_max;
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
    dg[gi] = d;
  }
  if (!dg.empty()) std::memcpy(host.data() + off_gt, dg.data(), sizeof(DevGroupT) * dg.size());
  HIPCHECK(hipMemcpy(A->dmem, host.data(), host.size(), hipMemcpyHostToDevice));
  A->dgt = reinterpret_cast<DevGroupT*>(base + off_gt);
  A->dpass = reinterpret_cast<AdjPass*>(base + off_pass);
  A->dptiles = reinterpret_cast<AdjPassTile*>(base + off_ptiles);
  A->dcells = reinterpret_cast<AdjCell*>(base + off_cells);
  A->dtiles = reinterpret_cast<AdjTile*>(base + off_tiles);
  A->dentries = reinterpret_cast<AdjEntry*>(base + off_entries);
  A->dzero = reinterpret_cast<Tile*>(base + off_zero);
}

// The inverted schedule, built once per plan; on failure nothing is kept.
void build_adjoint(tah_plan* p) {
  try {
    build_adjoint_tables(p);
    HIPCHECK(hipEventCreateWithFlags(&p->adj->zero_fork, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&p->adj->zero_join, hipEventDisableTiming));
    ctx_streams(p->ctx);
  } catch (...) {
    free_adjoint(p);
    throw;
  }
}

// staged: the passes' staging buffer, at least A.scratch_values doubles;
// pass items address it from its start.
template <int D>
void launch_pass(tah_plan* p, const AdjStage& S, int q, hipStream_t s, double* staged) {
  const AdjointTables& A = *p->adj;
  const int64_t n = S.pass_tiles[q][1];
  if (n == 0) return;
  const double* in = q == 0 ? nullptr : staged;
  hipLaunchKernelGGL(adjoint_pass_kernel<D>, dim3(uint32_t(n)), dim3(kThreads), 0, s, p->djob, A.dgt, A.dpass,
                     A.dptiles + S.pass_tiles[q][0], uint32_t(n), in, staged);
  HIPCHECK(hipGetLastError());
}

void check_adjoint_call(tah_plan* p, double* adj, const double* factors) {
  if (!p || !adj) fail("null argument");
  if (p->regrid) fail("the adjoint of a regrid transfer is not provided");
  if (p->needs_factors && !factors)
    fail("this schedule mirrors ghosts across a reflecting face, but no factor table was given");
}

// Every stage in reverse: passes, gather, zero. Each pass writes every
// staged value the gather after it reads, so the buffer needs no clearing.
// `after` (may be null) is called after each launch with its kind (0-2 the
// z, y, x passes, 3 the gather, 4 the zeroing) and algorithmic bytes.
// fork: a stage's zeroing runs on a second stream once the z pass (the one
// launch that reads the interpolated targets) is done, beside the y and x
// passes and the gather, which neither read nor write those targets (a
// stage's targets are never its sources); the stage ends when both have.
template <class After>
void run_adjoint_with(tah_plan* p, double* adj, const double* factors, hipStream_t s, double* staged,
                      const After& after, bool fork = false) {
  plan_upload_job(p, plan_job(p, adj, factors), s);
  const AdjointTables& A = *p->adj;
  for (int32_t st = p->nstages - 1; st >= 0; --st) {
    const AdjStage& S = A.stages[st];
    launch_pass<2>(p, S, 0, s, staged);
    if (S.pass_tiles[0][1] > 0) after(0, S.bytes[0]);
    const bool forked = fork && S.zero[1] > 0;
    if (forked) {
      const hipStream_t zs = ctx_streams(p->ctx)[0];
      HIPCHECK(hipEventRecord(A.zero_fork, s));
      HIPCHECK(hipStreamWaitEvent(zs, A.zero_fork, 0));
      hipLaunchKernelGGL(zero_kernel, dim3(uint32_t(S.zero[1])), dim3(kThreads), 0, zs, p->djob,
                         A.dzero + S.zero[0]);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipEventRecord(A.zero_join, zs));
    }
    launch_pass<1>(p, S, 1, s, staged);
    if (S.pass_tiles[1][1] > 0) after(1, S.bytes[1]);
    launch_pass<0>(p, S, 2, s, staged);
    if (S.pass_tiles[2][1] > 0) after(2, S.bytes[2]);
    if (S.gather[1] > 0) {
      hipLaunchKernelGGL(adjoint_gather_kernel, dim3(uint32_t(S.gather[1])), dim3(kThreads), 0, s, p->djob,
                         A.dcells, A.dtiles + S.gather[0], A.dentries, uint32_t(S.gather[1]), staged);
      HIPCHECK(hipGetLastError());
      after(3, S.bytes[3]);
    }
    if (forked) {
      HIPCHECK(hipStreamWaitEvent(s, A.zero_join, 0));
    } else if (S.zero[1] > 0) {
      hipLaunchKernelGGL(zero_kernel, dim3(uint32_t(S.zero[1])), dim3(kThreads), 0, s, p->djob,
                         A.dzero + S.zero[0]);
      HIPCHECK(hipGetLastError());
      after(4, S.bytes[4]);
    }
  }
}

void run_adjoint(tah_plan* p, double* adj, const double* factors, hipStream_t s, double* staged) {
  run_adjoint_with(p, adj, factors, s, staged, [](int, int64_t) {}, true);
}

}  // namespace
}  // namespace tah

// The staging buffer of tah_fill_adjoint_noalloc: one array per GPU in the
// code object, so that the call allocates nothing. Host code never uses this
// name as an address (on the host it stands in for the symbol only): the
// device address of each GPU's copy is taken once per device by
// noalloc_scratch_init and used from there on.
static __device__ __align__(256) double scratch[tah::kNoallocScratchValues];

namespace tah {
namespace {
constexpr int kMaxDevices = 64;
std::mutex noalloc_mutex;
// Per device: the array's address, and the stream and completion event of
// the last no-allocation adjoint on that device.
double* noalloc_ptr[kMaxDevices] = {};
hipEvent_t noalloc_done[kMaxDevices] = {};
hipStream_t noalloc_stream[kMaxDevices] = {};
bool noalloc_pending[kMaxDevices] = {};
}  // namespace

void noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) fail("device index out of range for the no-allocation adjoint");
  std::lock_guard<std::mutex> lock(noalloc_mutex);
  if (noalloc_ptr[device]) return;
  void* ptr = nullptr;
  HIPCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) fail("no device address for the no-allocation adjoint's staging buffer", kHip);
  HIPCHECK(hipEventCreateWithFlags(&noalloc_done[device], hipEventDisableTiming));
  noalloc_ptr[device] = static_cast<double*>(ptr);
}

}  // namespace tah

using namespace tah;

extern "C" tah_status tah_fill_adjoint(tah_plan* p, double* adj, const double* factors, void* stream) {
  return guard([&] {
    check_adjoint_call(p, adj, factors);
    HIPCHECK(hipSetDevice(p->ctx->device));
    if (!p->adj) build_adjoint(p);
    AdjointTables& A = *p->adj;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    plan_begin(p, s);
    if (!A.scratch && A.scratch_values > 0) HIPCHECK(hipMalloc(&A.scratch, sizeof(double) * A.scratch_values));
    run_adjoint(p, adj, factors, s, A.scratch);
    plan_end(p, s);
  });
}

extern "C" tah_status tah_fill_adjoint_timed(tah_plan* p, double* adj, const double* factors, void* stream,
                                             int32_t* n, int32_t* kind, int64_t* bytes, float* ms) {
  return guard([&] {
    if (!n) fail("null argument");
    check_adjoint_call(p, adj, factors);
    HIPCHECK(hipSetDevice(p->ctx->device));
    if (!p->adj) build_adjoint(p);
    AdjointTables& A = *p->adj;
    int32_t total = 0;
    for (const AdjStage& S : A.stages)
      total += (S.pass_tiles[0][1] > 0) + (S.pass_tiles[1][1] > 0) + (S.pass_tiles[2][1] > 0) + (S.gather[1] > 0) +
               (S.zero[1] > 0);
    if (!kind && !bytes && !ms) {
      *n = total;
      return;
    }
    if (!kind || !bytes || !ms) fail("null argument");
    if (*n < total)
      fail("the record arrays hold " + std::to_string(*n) + " launches; the adjoint makes " + std::to_string(total));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    plan_begin(p, s);
    if (!A.scratch && A.scratch_values > 0) HIPCHECK(hipMalloc(&A.scratch, sizeof(double) * A.scratch_values));
    std::vector<hipEvent_t> ev(size_t(total) + 1);
    for (auto& e : ev) HIPCHECK(hipEventCreate(&e));
    try {
      HIPCHECK(hipEventRecord(ev[0], s));
      int32_t k = 0;
      run_adjoint_with(p, adj, factors, s, A.scratch, [&](int kd, int64_t b) {
        if (k >= total) fail("adjoint launch count changed", kHip);
        HIPCHECK(hipEventRecord(ev[size_t(k) + 1], s));
        kind[k] = kd;
        bytes[k] = b;
        ++k;
      });
      plan_end(p, s);
      HIPCHECK(hipEventSynchronize(ev[size_t(total)]));
      for (int32_t i = 0; i < total; ++i) HIPCHECK(hipEventElapsedTime(&ms[i], ev[size_t(i)], ev[size_t(i) + 1]));
    } catch (...) {
      for (auto& e : ev) (void)hipEventDestroy(e);
      throw;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    *n = total;
  });
}

extern "C" tah_status tah_fill_adjoint_noalloc(tah_plan* p, double* adj, const double* factors, void* stream) {
  return guard([&] {
    check_adjoint_call(p, adj, factors);
    if (!p->adj)
      fail("the plan's adjoint tables are not built yet: they are built (and allocated) at set-up, by the "
           "plan's first tah_fill_adjoint call");
    const int dev = p->ctx->device;
    if (p->adj->scratch_values > kNoallocScratchValues)
      fail("this adjoint stages " + std::to_string(p->adj->scratch_values) +
               " values; the no-allocation staging buffer holds at most " + std::to_string(kNoallocScratchValues) +
               " (64 MiB): use tah_fill_adjoint",
           kUnsupported);
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (dev < 0 || dev >= kMaxDevices || !noalloc_ptr[dev])
      fail("the no-allocation staging buffer of device " + std::to_string(dev) + " was not set up", kHip);
    HIPCHECK(hipSetDevice(dev));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    plan_begin(p, s);
    // The buffer is shared by every call on this GPU: a call on another
    // stream than the last one waits for it.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIPCHECK(hipStreamIsCapturing(s, &cap));
    const bool captured = cap != hipStreamCaptureStatusNone;
    if (!captured && noalloc_pending[dev] && noalloc_stream[dev] != s)
      HIPCHECK(hipStreamWaitEvent(s, noalloc_done[dev], 0));
    run_adjoint(p, adj, factors, s, noalloc_ptr[dev]);
    if (!captured) {
      HIPCHECK(hipEventRecord(noalloc_done[dev], s));
      noalloc_stream[dev] = s;
      noalloc_pending[dev] = true;
    }
    plan_end(p, s);
  });
}
