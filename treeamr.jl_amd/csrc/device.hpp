// Device-side types shared by the fill kernels and the plan builder.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tah {

// Threads per workgroup of every tile kernel (four wave64s).
constexpr int kThreads = 256;

// Division of a 32-bit index by a run-time constant through a
// multiply-high, as the reference's LinearShape does with multiplicative
// inverses (src/threading.jl:240-267). Exact for every 32-bit numerator.
struct FastDiv {
  uint32_t d = 1, m = 0, s = 0;
  __host__ __device__ FastDiv() {}
  __host__ __device__ explicit FastDiv(uint32_t div) : d(div) {
    s = 0;
    while ((uint64_t(1) << s) < div) ++s;
    m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << s) - div)) / div + 1);
  }
  __host__ __device__ inline uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t hi = __umulhi(n, m);
#else
    uint32_t hi = uint32_t((uint64_t(n) * m) >> 32);
#endif
    return uint32_t((uint64_t(hi) + n) >> s);
  }
};

// One transfer group as the kernels read it: the flattened Stencil1D of
// each dimension and the block lists, in device memory.
struct alignas(16) DevGroup {
  int32_t kind;
  int32_t factorcol;
  int32_t ntransfers;
  int32_t unit;       // 1: every stencil one point of weight one; 2: and affine rows
  int32_t n[3];       // box extents
  int32_t order[3];   // stencil widths
  int32_t tfirst[3];  // first target row, 0-based
  int32_t points;     // n0*n1*n2
  int32_t ca[3];      // a unit copy's source row of target row 0 ...
  int32_t cs[3];      // ... and its step per row (1, -1 or 0)
  FastDiv dn[3];      // division by n0, n1, n2
  FastDiv dper;       // division by points*nvars, the values of one transfer
  const int32_t* tblocks;
  const int32_t* sblocks;
  const int32_t* srcstart[3];
  const double* weights[3];
};

// Interpolation units (up to 2x2x2 targets of one transfer that share a
// source window) whose rows carry one and the same set of weights, so that
// every tap's weight product ((w1*w2)*w3) is one table, W, for the whole
// class, laid out in the order the kernel visits the taps.
struct alignas(16) DevWClass {
  int32_t group;
  int32_t nunits;
  const int32_t* units;  // 8 per unit: x rows (2), y rows (2), z rows (2), -, -
  const double* W;
  int32_t rbits;  // bit d set where the unit has two rows along d
  int32_t pad;
};

// A chunk of an interpolating group's target box whose transfers a
// workgroup fills from an LDS copy of the chunk's source box: target rows
// tlo .. tlo+tn-1 per dimension (0-based within the group's box), read
// from source rows slo .. slo+sext-1 of the source block (the union of the
// rows' windows), one pair of variables at a time.
// In LDS each dimension of the box is split into its even and its odd rows:
// source row x along dimension d sits at X = (x & 1)*half[d] + (x >> 1), and
// point (X0, X1, X2) is 16-byte pair X0 + X1*pitch[0] + X2*pitch[1], both
// pitches odd. Every window of the chunk starts an even number of rows into
// the box (the windows of neighbouring targets two rows apart), so
// neighbouring targets' windows start in neighbouring pairs of LDS. A
// thread's targets are numbered with dimension ord[0] fastest, then ord[1],
// then ord[2] (ord[0] a dimension of at least 8 rows where there is one):
// the lanes that LDS serves together then read different banks.
struct alignas(16) LdsChunk {
  int32_t group;
  int32_t tlo[3], tn[3];
  int32_t slo[3], sext[3];
  int32_t points;    // tn0*tn1*tn2
  FastDiv dt[2];     // division by tn[ord[0]], tn[ord[1]]
  FastDiv dsx[2];    // division by sext0, sext1
  int32_t staged;    // sext0*sext1*sext2 (points copied)
  int32_t ord[3];
  int32_t half[3];   // (sext[d] + 1) / 2
  int32_t pitch[2];
  int32_t slots;     // 16-byte slots of LDS the box takes
};

// Units of one class in a chunk of a prolongation's box, for one wave:
// units[unit0 .. unit0+nu-1] of the plan's LDS unit list, each two int32:
// the rows (ra0, ra1, rb0, rb1) and (rc0, rc1) of the unit's targets along
// x, y, z as bytes (a second row 255 where the unit has one row there).
struct alignas(16) LdsWave {
  int32_t wclass;  // index into the job's DevWClass list
  int32_t unit0;
  int32_t nu;
  int32_t pad;
};

// A chunk of a prolongation's box: target rows as in LdsChunk, its source
// box slo .. slo+sext-1 per dimension (point (x, y, z) at 16-byte pair
// x + y*pitch[0] + z*pitch[1] of LDS, odd pitches), and its class waves
// waves[wave0 .. wave0+nwave-1].
struct alignas(16) LdsUnitChunk {
  int32_t group;
  int32_t slo[3], sext[3];
  int32_t pitch[2];
  int32_t slots;
  int32_t staged;
  int32_t wave0, nwave;
  FastDiv dsx[2];  // division by sext0, sext1
  int32_t pad[3];
};

// The most LDS an LDS-staged workgroup takes (a chunk's source box, as
// 16-byte pairs of variables).
constexpr int kLdsMaxBytes = 64 * 1024;

// The staging slots of a streamed chunk's plane per thread: a plane holds at
// most kStreamSlots pairs of variables per thread of its workgroup.
constexpr int kStreamSlots = 8;
// The most z windows of a streamed interpolation of order P that may hold
// one plane (the kernel's accumulators per thread). Restriction (shift 2):
// P/2 where windows start two rows apart, one more where the first two
// windows share a start (clamped into the fine block's interior).
// Prolongation (shift 0): the two fine rows of a coarse row share its
// window, and windows start one coarse row apart: 2P. kStreamKA bounds them
// all.
constexpr int stream_ka(int P, int shift) { return shift != 0 ? P / 2 + 1 : 2 * P; }
constexpr int kStreamKA = 12;

// One source plane of a streamed restriction's group, the planes zlo ..
// zlo + nplanes - 1 in ascending order: the z rows whose windows hold it are
// slots 0 .. nact-1, in window order, and c2[q] is slot q's weight along z
// at this plane; slots had .. nact-1 start here (at +0); the first nret
// slots end here -- rows k[0 .. nret-1], stored after the plane -- and the
// rest then move down by nret.
struct alignas(16) StreamPlane {
  double c2[kStreamKA];
  int32_t k[kStreamKA];
  int32_t had, nact, nret, pad;
};

// A chunk of a restriction's target box streamed along z: target rows
// tlo[d] .. tlo[d]+tn[d]-1 along x and y (d = 0, 1) and every row along z.
// Each source plane z (rows slo[d] .. slo[d]+sext[d]-1 along x and y, the
// union of the chunk's windows) is staged in LDS with a tile's variables
// innermost, point (x, y) at (y*sext[0] + x) * nvp doubles, nvp the tile's
// variable count rounded up to even. Its group's planes are
// planes[plane0 .. plane0 + nplanes - 1] of the job's list, the first at
// source row zlo. nvt: variables per tile.
struct alignas(16) StreamChunk {
  int32_t group;
  int32_t tlo[2], tn[2];
  int32_t slo[2], sext[2];
  int32_t plane0, nplanes, zlo;
  int32_t nvt;
  int32_t threads;  // the workgroup size its launch takes: 64, 128 or 256
  int32_t pad[2];
};

// One workgroup's share of an LDS-staged launch: chunk `chunk` (of the job's
// chunk list) of transfer `transfer`, variables v0 and v0+1.
struct alignas(16) LdsTile {
  int32_t job;
  int32_t chunk;
  int32_t transfer;
  int32_t v0;
};

// One fill of one working array. Element (i, j, k, v) of block b is at
// stride[0]*i + stride[1]*j + stride[2]*k + stride[3]*v + stride[4]*b.
struct alignas(16) Job {
  const DevGroup* groups;
  const DevWClass* wclasses;
  const LdsChunk* chunks;  // the plan's LDS-staged chunks
  const LdsUnitChunk* uchunks;  // ... of prolongation, their class waves and units
  const LdsWave* waves;
  const uint32_t* units;
  const StreamChunk* schunks;  // the plan's restriction chunks streamed along z
  const StreamPlane* planes;   // ... and their groups' planes
  double* work;        // the targets' array
  const double* src;   // the sources' array: work, or a regrid's old array
  const double* factors;
  int64_t stride[5];
  int32_t nvars;
  int32_t vfast;  // stride[3] < stride[0]: variables innermost
  FastDiv dnv;    // division by nvars
  // Variables innermost (stride[3] == 1), nvars and the other strides even
  // and both arrays 16-byte aligned: values 2e, 2e+1 of any point are one
  // aligned 16-byte pair.
  int32_t pairs;
};

// An outer-face hook region in HBM: its block, box (first stored index lo,
// extents n), the nearest evolved point's index along each clamped
// dimension (-1: the ghost's own index) and its first point in the caller's
// value arrays.
struct alignas(16) HookRegion {
  int32_t block;
  int32_t lo[3];
  int32_t n[3];
  int32_t clamp[3];
  int64_t first;
};

// A hook launch's share of one region: values e0 .. e0 + kHookTile - 1 of
// its point-major, variable-innermost numbering (256 threads, four each).
constexpr int kHookPerThread = 4;
constexpr int kHookTile = 256 * kHookPerThread;
struct alignas(8) HookTile {
  int32_t region;
  int32_t e0;
};

// A slice of work. Copy and generic tiles: group `item`, transfer, flat
// elements from e0. Interpolation tiles: class `item`, nu units from unit e0
// of transfer `transfer` on through the next transfers (unit e0+i is unit
// (e0+i) mod nunits of transfer transfer + (e0+i) div nunits), times
// variables v0 .. v0+nvt-1.
struct alignas(16) Tile {
  int32_t job;
  int32_t item;
  int32_t transfer;
  int32_t e0;
  int32_t nu;
  int32_t v0;
  int32_t nvt;
  int32_t pad;
};

// Workgroups are dealt round-robin to the eight XCDs, each with its own L2:
// give XCD x the x-th contiguous share of the tiles, so that neighbouring
// tiles (the same transfer, the same blocks) meet in one L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
  constexpr uint32_t X = 8;
  const uint32_t per = n / X, rem = n % X;
  const uint32_t x = b % X, i = b / X;
  return x * per + (x < rem ? x : rem) + i;
}

}  // namespace tah
