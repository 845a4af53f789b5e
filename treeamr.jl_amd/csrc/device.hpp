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
  int32_t unit;
  int32_t n[3];       // box extents
  int32_t order[3];   // stencil widths
  int32_t tfirst[3];  // first target row, 0-based
  int32_t points;     // n0*n1*n2
  int32_t ca[3];      // a unit copy's source row of target row 0 ...
  int32_t cs[3];      // ... and its step per row (1, -1 or 0)
  FastDiv dn[3];      // division by n0, n1, n2
  const int32_t* tblocks;
  const int32_t* sblocks;
  const int32_t* srcstart[3];
  const double* weights[3];
};

// A sub-box of an interpolating group in which every thread computes the
// same shape of targets: R_d rows along dimension d (R_d = 2: a pair of
// rows whose windows start SHIFT apart), rows[d] listing them unit by unit.
struct alignas(16) DevSub {
  int32_t group;
  int32_t units;   // cnt0*cnt1*cnt2
  int32_t cnt[3];
  int32_t pad;
  FastDiv dc[2];   // division by cnt0, cnt1
  FastDiv du;      // division by units
  const int32_t* rows[3];
};

// One fill of one working array.
struct alignas(16) Job {
  const DevGroup* groups;
  const DevSub* subs;
  double* work;
  const double* factors;
  int64_t stride[5];
  int32_t nvars;
  int32_t pad;
};

// A slice of work: job, group (or sub-box), transfer, first flat element.
struct alignas(16) Tile {
  int32_t job;
  int32_t group;
  int32_t transfer;
  int32_t e0;
};

// What an interpolation tile stages in LDS: the source box lo..lo+ext-1
// (stored indices) of variables v0..v0+nv-1, and how many of the tile's
// flat elements it computes.
struct alignas(16) TileBox {
  int32_t lo[3];
  int32_t ext[3];
  int32_t v0, nv;
  int32_t count;
  int32_t pad[3];
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
