// The fill kernels for gfx950.
//
// Every value is computed as the reference's transfer_kernel! computes it
// (src/ghosts.jl:372-402): a unit copy is 0 + u (copy_load, :127-128); a
// stencil is one accumulator from +0 with m1 innermost and the tap weight
// ((w1*w2)*w3) added as acc += w*u (stencil_sum, :58-118), every product and
// sum rounded on its own (_rn intrinsics; the file is also compiled with
// -ffp-contract=off); a factor is acc*f + 0 (scaled, :369-370). Kernels that
// compute several targets per thread keep one accumulator per target and
// visit each target's taps in that same order, so they change no bit; a
// weight product formed on the host in double, without contraction, is the
// same number the device would form.
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace tah {

// Waves per SIMD the merged interpolation kernels are compiled for, and how
// many rows ahead restriction loads its window. Measured on the large forest
// at order 6: restriction, whose rows wait on HBM, takes 3.3 ms with its
// loads three rows ahead at four waves against 3.7 ms with none at seven;
// prolongation gained nothing from loads ahead at any depth, single (seven
// waves) or over variable pairs (five).
constexpr int kRestrictWaves = 4;
constexpr int kRestrictRowsAhead = 3;
constexpr int kProlongWaves = 7;
constexpr int kPairWaves = 5;

// A value every lane holds alike, moved into scalar registers, so that what
// is loaded through it is loaded by the scalar unit.
__device__ __forceinline__ int32_t uniform(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
template <class T>
__device__ __forceinline__ T* uniform(T* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint64_t lo = uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(v)));
  const uint64_t hi = uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(v >> 32)));
  return reinterpret_cast<T*>((hi << 32) | lo);
}

// A uniform table pointer in the constant address space, so that what is
// read through it arrives by scalar loads into SGPRs.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* constant(const T* p) {
  return (const __attribute__((address_space(4))) T*)(uniform(const_cast<T*>(p)));
}
// A pointer in the global address space (plain vector loads, not flat ones).
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* global(const T* p) {
  return (const __attribute__((address_space(1))) T*)(p);
}

__device__ __forceinline__ double scaled(double acc, const double* __restrict__ factors,
                                         int32_t factorcol, int32_t nvars, int32_t v) {
  if (factorcol == 0) return acc;
  double f = factors[v + int64_t(nvars) * (factorcol - 1)];
  return __dadd_rn(__dmul_rn(acc, f), 0.0);
}

// The full tensor sum of one target, any widths (the device form: no
// one-tap skip, as a flat launch in the reference takes it). sb points at
// the source block's variable; s0..s2 are the point strides.
__device__ __forceinline__ double stencil_value(const DevGroup& g, const double* __restrict__ sb,
                                                int64_t s0, int64_t s1, int64_t s2, int32_t i,
                                                int32_t j, int32_t k) {
  const int32_t b0 = g.srcstart[0][i], b1 = g.srcstart[1][j], b2 = g.srcstart[2][k];
  if (g.unit) return __dadd_rn(0.0, sb[s0 * b0 + s1 * b1 + s2 * b2]);
  const int32_t o0 = g.order[0], o1 = g.order[1], o2 = g.order[2];
  const double* __restrict__ w0 = g.weights[0] + int64_t(o0) * i;
  const double* __restrict__ w1 = g.weights[1] + int64_t(o1) * j;
  const double* __restrict__ w2 = g.weights[2] + int64_t(o2) * k;
  double acc = 0.0;
  for (int32_t m2 = 0; m2 < o2; ++m2) {
    const double c2 = w2[m2];
    for (int32_t m1 = 0; m1 < o1; ++m1) {
      const double c1 = w1[m1];
      const double* __restrict__ row = sb + s2 * (b2 + m2) + s1 * (b1 + m1) + s0 * b0;
      for (int32_t m0 = 0; m0 < o0; ++m0) {
        double w = __dmul_rn(__dmul_rn(w0[m0], c1), c2);
        acc = __dadd_rn(acc, __dmul_rn(w, row[s0 * m0]));
      }
    }
  }
  return acc;
}

// ---------------------------------------------------------------------------
// The plain fill: one launch per group, one thread per written value, the
// flat index split into (i, j, k, v, t) with i fastest (launch_transfer!'s
// flat form, src/ghosts.jl:507-521).
#ifdef TAH_DEFINE_KERNELS
__global__ void __launch_bounds__(256) plain_kernel(DevGroup g, double* __restrict__ work,
                                                    const double* __restrict__ factors, int64_t s0,
                                                    int64_t s1, int64_t s2, int64_t s3, int64_t s4,
                                                    int32_t nvars, FastDiv dv, uint32_t total, int32_t t0) {
  uint32_t e = blockIdx.x * 256u + threadIdx.x;
  if (e >= total) return;
  uint32_t r = g.dn[0].div(e);
  int32_t i = int32_t(e - r * uint32_t(g.n[0]));
  uint32_t r2 = g.dn[1].div(r);
  int32_t j = int32_t(r - r2 * uint32_t(g.n[1]));
  uint32_t r3 = g.dn[2].div(r2);
  int32_t k = int32_t(r2 - r3 * uint32_t(g.n[2]));
  const uint32_t tl = dv.div(r3);
  int32_t v = int32_t(r3 - tl * uint32_t(nvars));
  const int32_t t = t0 + int32_t(tl);  // transfers t0 .. of this launch
  const double* sb = work + s4 * g.sblocks[t] + s3 * v;
  double acc = stencil_value(g, sb, s0, s1, s2, i, j, k);
  acc = scaled(acc, factors, g.factorcol, nvars, v);
  double* tb = work + s4 * g.tblocks[t] + s3 * v;
  tb[s0 * (g.tfirst[0] + i) + s1 * (g.tfirst[1] + j) + s2 * (g.tfirst[2] + k)] = acc;
}

#endif  // TAH_DEFINE_KERNELS

// The outer-face hook (the reference's boundary_kernel! and relative
// kernels, src/ghosts.jl:1177-1260) with its values computed by the caller:
// value e of a tile's region is variable e mod nvars of point e div nvars,
// the point (i, j, k) of the region's box with i fastest, the caller's
// value at index (first + point)*nvars + variable. Relative: the ghost
// becomes values + (work[r] - fr), r the point with the clamped indices
// replaced; r never lies in a region of the same launch (an evolved point,
// or a ghost a transfer filled).

#ifdef TAH_DEFINE_KERNELS
__global__ void __launch_bounds__(256) hook_kernel(const Job* __restrict__ jobs,
                                                   const HookRegion* __restrict__ regions,
                                                   const HookTile* __restrict__ tiles,
                                                   const double* __restrict__ values,
                                                   const double* __restrict__ fr, int32_t relative) {
  const Job& job = jobs[0];
  const HookTile t = tiles[blockIdx.x];
  const HookRegion R = regions[t.region];
  const int32_t nv = job.nvars;
  const uint32_t total = uint32_t(R.n[0]) * uint32_t(R.n[1]) * uint32_t(R.n[2]) * uint32_t(nv);
  const int64_t s0 = job.stride[0], s1 = job.stride[1], s2 = job.stride[2], s3 = job.stride[3];
  double* b = job.work + job.stride[4] * R.block;
  for (int r = 0; r < kHookPerThread; ++r) {
    const uint32_t e = uint32_t(t.e0) + uint32_t(r * 256) + threadIdx.x;
    if (e >= total) break;
    const uint32_t p = e / uint32_t(nv);
    const int32_t v = int32_t(e - p * uint32_t(nv));
    const uint32_t q = p / uint32_t(R.n[0]);
    const int32_t i = R.lo[0] + int32_t(p - q * uint32_t(R.n[0]));
    const int32_t j = R.lo[1] + int32_t(q % uint32_t(R.n[1]));
    const int32_t k = R.lo[2] + int32_t(q / uint32_t(R.n[1]));
    const int64_t at = (R.first + int64_t(p)) * nv + v;
    double x = values[at];
    if (relative) {
      const int32_t ri = R.clamp[0] >= 0 ? R.clamp[0] : i;
      const int32_t rj = R.clamp[1] >= 0 ? R.clamp[1] : j;
      const int32_t rk = R.clamp[2] >= 0 ? R.clamp[2] : k;
      x = __dadd_rn(x, __dsub_rn(b[s0 * ri + s1 * rj + s2 * rk + s3 * v], fr[at]));
    }
    b[s0 * i + s1 * j + s2 * k + s3 * v] = x;
  }
}
#endif  // TAH_DEFINE_KERNELS

// Split a flat element of a (box x nvars) space: variables fastest when
// they are innermost in memory, else x fastest.
__device__ __forceinline__ void split_element(const DevGroup& g, const Job& job, uint32_t e,
                                              int32_t& i, int32_t& j, int32_t& k, int32_t& v) {
  if (job.vfast) {
    const uint32_t r = job.dnv.div(e);
    v = int32_t(e - r * uint32_t(job.nvars));
    const uint32_t q = g.dn[0].div(r);
    i = int32_t(r - q * uint32_t(g.n[0]));
    const uint32_t q2 = g.dn[1].div(q);
    j = int32_t(q - q2 * uint32_t(g.n[1]));
    k = int32_t(q2 - g.dn[2].div(q2) * uint32_t(g.n[2]));
  } else {
    const uint32_t q = g.dn[0].div(e);
    i = int32_t(e - q * uint32_t(g.n[0]));
    const uint32_t q2 = g.dn[1].div(q);
    j = int32_t(q - q2 * uint32_t(g.n[1]));
    const uint32_t v32 = g.dn[2].div(q2);
    k = int32_t(q2 - v32 * uint32_t(g.n[2]));
    v = int32_t(v32);
  }
}

// ---------------------------------------------------------------------------
// Generic tiles: any group, one value per thread-iteration (the fallback
// for stencils no specialised kernel covers).
constexpr int kGenericPerThread = 4;
constexpr int kGenericTile = kThreads * kGenericPerThread;

#ifdef TAH_DEFINE_KERNELS
__global__ void __launch_bounds__(kThreads) generic_kernel(const Job* __restrict__ jobs,
                                                           const Tile* __restrict__ tiles) {
  const Tile tile = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const Job& job = jobs[tile.job];
  const DevGroup& g = job.groups[tile.item];
  const int64_t s0 = job.stride[0], s1 = job.stride[1], s2 = job.stride[2], s3 = job.stride[3],
                s4 = job.stride[4];
  const int32_t nvars = job.nvars;
  const uint32_t total = uint32_t(g.points) * uint32_t(nvars);
  const double* sbase = job.src + s4 * g.sblocks[tile.transfer];
  double* tbase = job.work + s4 * g.tblocks[tile.transfer];
  for (int r = 0; r < kGenericPerThread; ++r) {
    uint32_t e = uint32_t(tile.e0) + uint32_t(r * kThreads) + threadIdx.x;
    if (e >= total) break;
    int32_t i, j, k, v;
    split_element(g, job, e, i, j, k, v);
    double acc = stencil_value(g, sbase + s3 * v, s0, s1, s2, i, j, k);
    acc = scaled(acc, job.factors, g.factorcol, nvars, v);
    tbase[s3 * v + s0 * (g.tfirst[0] + i) + s1 * (g.tfirst[1] + j) + s2 * (g.tfirst[2] + k)] = acc;
  }
}

#endif  // TAH_DEFINE_KERNELS

// ---------------------------------------------------------------------------
// Copies (unit stencils: same-level copies, mirrored and clamped rows): the
// source row of target row r along d is ca[d] + cs[d]*r, so no index table
// is read. Each thread issues all of its loads before its first store.
constexpr int kCopyPerThread = 8;
constexpr int kCopyTile = kThreads * kCopyPerThread;

// A tile of tile.nu > 1 small transfers of one group, packed: its values
// are the nu transfers' values one after another (transfers tile.transfer
// ..), so that a tile of small boxes (corners, edges of small blocks) keeps
// its threads busy. Each value finds its transfer by one division.
template <int CPT>
__device__ __forceinline__ void copy_packed_body(const Job& job, const Tile& tile) {
  typedef double pair_t __attribute__((ext_vector_type(2)));
  const DevGroup& g = job.groups[tile.item];
  const int32_t s0 = int32_t(job.stride[0]), s1 = int32_t(job.stride[1]),
                s2 = int32_t(job.stride[2]), s3 = int32_t(job.stride[3]);
  const int64_t s4 = job.stride[4];
  const int32_t nvars = job.nvars;
  const uint32_t per = uint32_t(g.points) * uint32_t(nvars);
  const uint32_t total = per * uint32_t(tile.nu);
  const int32_t tb0 = s0 * g.tfirst[0] + s1 * g.tfirst[1] + s2 * g.tfirst[2];
  const int32_t sb0 = s0 * g.ca[0] + s1 * g.ca[1] + s2 * g.ca[2];
  const int32_t ds0 = s0 * g.cs[0], ds1 = s1 * g.cs[1], ds2 = s2 * g.cs[2];
  constexpr int STEP = 2;  // pairs where the array allows, else single values twice
#pragma unroll
  for (int r = 0; r < CPT / STEP; ++r) {
    const uint32_t q = STEP * (uint32_t(r * kThreads) + threadIdx.x);
    if (q >= total) break;
    if (job.pairs) {
      // values q, q+1: one point's variables v, v+1 of one transfer (the
      // count of a transfer's values is even)
      const uint32_t tl = g.dper.div(q);
      const int32_t tr = tile.transfer + int32_t(tl);
      const double* __restrict__ sb = job.src + s4 * g.sblocks[tr];
      double* __restrict__ tb = job.work + s4 * g.tblocks[tr];
      int32_t i, j, k, v;
      split_element(g, job, q - tl * per, i, j, k, v);
      const pair_t x = *reinterpret_cast<const pair_t*>(sb + (v + sb0 + ds0 * i + ds1 * j + ds2 * k));
      pair_t out;
      out.x = scaled(__dadd_rn(0.0, x.x), job.factors, g.factorcol, nvars, v);
      out.y = scaled(__dadd_rn(0.0, x.y), job.factors, g.factorcol, nvars, v + 1);
      __builtin_nontemporal_store(out, reinterpret_cast<pair_t*>(tb + (v + tb0 + s0 * i + s1 * j + s2 * k)));
    } else {
#pragma unroll
      for (int h = 0; h < STEP; ++h) {
        const uint32_t qh = q + uint32_t(h);
        if (qh >= total) break;
        const uint32_t tlh = g.dper.div(qh);
        const int32_t trh = tile.transfer + int32_t(tlh);
        int32_t ih, jh, kh, vh;
        split_element(g, job, qh - tlh * per, ih, jh, kh, vh);
        const double* __restrict__ sbh = job.src + s4 * g.sblocks[trh];
        double* __restrict__ tbh = job.work + s4 * g.tblocks[trh];
        const double x = sbh[s3 * vh + sb0 + ds0 * ih + ds1 * jh + ds2 * kh];
        tbh[s3 * vh + tb0 + s0 * ih + s1 * jh + s2 * kh] =
            scaled(__dadd_rn(0.0, x), job.factors, g.factorcol, nvars, vh);
      }
    }
  }
}

// CPT values per thread (the tile shape a plan option picks: 4, 8 or 16).
template <int CPT>
__device__ __forceinline__ void copy_body(const Job& job, const Tile& tile) {
  if (tile.nu > 1) {
    copy_packed_body<CPT>(job, tile);
    return;
  }
  const DevGroup& g = job.groups[tile.item];
  // Offsets within one block fit 32 bits (plans refuse larger blocks).
  const int32_t s0 = int32_t(job.stride[0]), s1 = int32_t(job.stride[1]),
                s2 = int32_t(job.stride[2]), s3 = int32_t(job.stride[3]);
  const int64_t s4 = job.stride[4];
  const int32_t nvars = job.nvars;
  const uint32_t total = uint32_t(g.points) * uint32_t(nvars);
  const double* __restrict__ sbase = job.src + s4 * g.sblocks[tile.transfer];
  double* __restrict__ tbase = job.work + s4 * g.tblocks[tile.transfer];
  const int32_t tb0 = s0 * g.tfirst[0] + s1 * g.tfirst[1] + s2 * g.tfirst[2];
  const int32_t sb0 = s0 * g.ca[0] + s1 * g.ca[1] + s2 * g.ca[2];
  const int32_t ds0 = s0 * g.cs[0], ds1 = s1 * g.cs[1], ds2 = s2 * g.cs[2];
  if (job.pairs) {
    // Values 2e and 2e+1 are one point's neighbouring variables, an aligned
    // 16-byte pair in source and target alike: one 16-byte load and store
    // for both. The ghosts are not read again within the
    // fill: the stores stream past the caches.
    constexpr int CP = CPT / 2;
    typedef double pair_t __attribute__((ext_vector_type(2)));
    int32_t dst2[CP], var2[CP];
    pair_t val2[CP];
#pragma unroll
    for (int r = 0; r < CP; ++r) {
      uint32_t e = uint32_t(tile.e0) + 2u * (uint32_t(r * kThreads) + threadIdx.x);
      const bool ok = e < total;
      e = ok ? e : 0u;
      int32_t i, j, k, v;
      split_element(g, job, e, i, j, k, v);
      dst2[r] = ok ? v + tb0 + s0 * i + s1 * j + s2 * k : -1;
      var2[r] = v;
      val2[r] = *reinterpret_cast<const pair_t*>(sbase + (v + sb0 + ds0 * i + ds1 * j + ds2 * k));
    }
#pragma unroll
    for (int r = 0; r < CP; ++r) {
      if (dst2[r] < 0) continue;
      pair_t out;
      out.x = scaled(__dadd_rn(0.0, val2[r].x), job.factors, g.factorcol, nvars, var2[r]);
      out.y = scaled(__dadd_rn(0.0, val2[r].y), job.factors, g.factorcol, nvars, var2[r] + 1);
      __builtin_nontemporal_store(out, reinterpret_cast<pair_t*>(tbase + dst2[r]));
    }
    return;
  }
  int32_t dst[CPT], var[CPT];
  double val[CPT];
#pragma unroll
  for (int r = 0; r < CPT; ++r) {
    uint32_t e = uint32_t(tile.e0) + uint32_t(r * kThreads) + threadIdx.x;
    const bool ok = e < total;
    e = ok ? e : 0u;
    int32_t i, j, k, v;
    split_element(g, job, e, i, j, k, v);
    const int32_t vo = s3 * v;
    dst[r] = ok ? vo + tb0 + s0 * i + s1 * j + s2 * k : -1;
    var[r] = v;
    val[r] = sbase[vo + sb0 + ds0 * i + ds1 * j + ds2 * k];
  }
#pragma unroll
  for (int r = 0; r < CPT; ++r) {
    if (dst[r] < 0) continue;
    tbase[dst[r]] = scaled(__dadd_rn(0.0, val[r]), job.factors, g.factorcol, nvars, var[r]);
  }
}

#ifdef TAH_DEFINE_KERNELS
template <int CPT>
__global__ void __launch_bounds__(kThreads) copy_kernel(const Job* __restrict__ jobs,
                                                        const Tile* __restrict__ tiles) {
  const Tile tile = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  copy_body<CPT>(jobs[tile.job], tile);
}

// The same copies from a bounded grid (a multiple of eight workgroups): the
// workgroups of XCD x walk x's contiguous share of the tiles, so that the
// copies hold only a few wave slots per CU and leave the rest to the
// interpolations launched beside them on another stream.
template <int CPT>
__global__ void __launch_bounds__(kThreads) copy_loop_kernel(const Job* __restrict__ jobs,
                                                             const Tile* __restrict__ tiles, uint32_t ntiles) {
  constexpr uint32_t X = 8;
  const uint32_t x = blockIdx.x % X, slot = blockIdx.x / X, nslots = gridDim.x / X;
  const uint32_t per = ntiles / X, rem = ntiles % X;
  const uint32_t lo = x * per + (x < rem ? x : rem), hi = lo + per + (x < rem ? 1u : 0u);
  for (uint32_t t = lo + slot; t < hi; t += nslots) {
    const Tile tile = tiles[t];
    copy_body<CPT>(jobs[tile.job], tile);
  }
}
#endif  // TAH_DEFINE_KERNELS

// ---------------------------------------------------------------------------
// Interpolation (restriction and prolongation) of order P in every
// dimension. A thread computes R0 x R1 x R2 targets of one variable: along a
// dimension with R_d = 2 the two target rows' source windows start SHIFT
// apart (SHIFT = 0 for prolongation, whose pairs of fine points share a
// coarse window; 2 for restriction), so every source value is loaded once
// for all of them. Each target keeps its own accumulator and visits its own
// taps in the reference's order (m3, m2, m1 ascending); the weight product
// of every (target, tap) is the class's table W, read in visit order through
// uniform loads, identical for every lane.
//
// MODE 0: one variable per thread. MODE 1 and 2 (variables innermost, an
// even count): two neighbouring variables per thread, which share every
// weight and every address; MODE 2 loads and stores the two as one aligned
// 16-byte pair, MODE 1 (an array not 16-byte aligned) one at a time.
template <int P, int SHIFT, int R0, int R1, int R2, int MODE = 0>
__device__ __forceinline__ void interp_body(const Job& job, const Tile& tile, const DevWClass& wc) {
  constexpr int U0 = P + (R0 == 2 ? SHIFT : 0);
  constexpr int U1 = P + (R1 == 2 ? SHIFT : 0);
  constexpr int U2 = P + (R2 == 2 ? SHIFT : 0);
  constexpr int NV = MODE == 0 ? 1 : 2;
  typedef double pair_t __attribute__((ext_vector_type(2)));
  const DevGroup& g = job.groups[wc.group];
  const int32_t t = int32_t(threadIdx.x);
  // threads per unit: one per variable, or per pair (the last one of an odd
  // count alone)
  const int32_t lanes = MODE == 0 ? tile.nvt : (tile.nvt + 1) / 2;
  if (t >= tile.nu * lanes) return;
  int32_t ul, vl;
  if (MODE != 0 || job.vfast) {
    ul = t / lanes;
    vl = t - ul * lanes;
  } else {
    vl = t / tile.nu;
    ul = t - vl * tile.nu;
  }
  const int32_t v = tile.v0 + NV * vl;
  // MODE 1 over an odd count: the last lane's second variable is none; it
  // loads its own variable twice and stores one.
  const bool has2 = MODE != 1 || v + 1 < tile.v0 + tile.nvt;
  const int32_t o1 = has2 ? 1 : 0;
  // The tile's units run on across consecutive transfers of the group.
  int32_t u = tile.e0 + ul, tr = tile.transfer;
  const int32_t nun = wc.nunits;
  if (u >= nun) {
    const int32_t q = u / nun;
    tr += q;
    u -= q * nun;
  }
  const int32_t* __restrict__ un = wc.units + 8 * u;
  const int32_t s0 = int32_t(job.stride[0]), s1 = int32_t(job.stride[1]),
                s2 = int32_t(job.stride[2]), s3 = int32_t(job.stride[3]);
  const int64_t s4 = job.stride[4];
  const int32_t ra[2] = {un[0], un[1]}, rb[2] = {un[2], un[3]}, rc[2] = {un[4], un[5]};
  const double* __restrict__ src = job.src + s4 * g.sblocks[tr] +
                                   (s3 * v + s0 * g.srcstart[0][ra[0]] + s1 * g.srcstart[1][rb[0]] +
                                    s2 * g.srcstart[2][rc[0]]);
  // The table is the same for every lane: read it through the constant
  // address space, so that it arrives by scalar loads into SGPRs.
  uint64_t wbase = reinterpret_cast<uint64_t>(uniform(wc.W));
  double acc[NV][R0][R1][R2];
#pragma unroll
  for (int n = 0; n < NV; ++n)
#pragma unroll
    for (int a = 0; a < R0; ++a)
#pragma unroll
      for (int b = 0; b < R1; ++b)
#pragma unroll
        for (int c = 0; c < R2; ++c) acc[n][a][b][c] = 0.0;
  // One row of the window (m2, m1): U0 values of each variable.
  auto load_row = [&](int m2, int m1, double (&row)[NV][U0]) {
    const __attribute__((address_space(1))) double* __restrict__ p =
        (const __attribute__((address_space(1))) double*)(src + (s2 * m2 + s1 * m1));
#pragma unroll
    for (int m0 = 0; m0 < U0; ++m0) {
      if (MODE == 2) {
        const pair_t x = *(const __attribute__((address_space(1))) pair_t*)(p + s0 * m0);
        row[0][m0] = x.x;
        row[NV - 1][m0] = x.y;
      } else {
        row[0][m0] = p[s0 * m0];
        if (MODE == 1) row[NV - 1][m0] = p[s0 * m0 + o1];
      }
    }
  };
  // PF rows ahead (restriction): the loads of row r + PF are issued before
  // row r's taps, into a ring of PF + 1 row buffers, so that each row's
  // memory latency overlaps the rows before it within the wave.
  constexpr int PF = SHIFT != 0 ? kRestrictRowsAhead : 0;
  constexpr int NR = U1 * U2;
  double rows[PF + 1][NV][U0];
#pragma unroll
  for (int q = 0; q < PF && q < NR; ++q) load_row(q / U1, q % U1, rows[q]);
  int idx = 0;
#pragma unroll
  for (int m2 = 0; m2 < U2; ++m2) {
#pragma unroll
    for (int m1 = 0; m1 < U1; ++m1) {
      // The table pointer passes through an empty asm per row, so that each
      // row's weights are loaded (by the scalar unit) when the row is summed,
      // not all hoisted to the top of the kernel.
      asm volatile("" : "+s"(wbase));
      const __attribute__((address_space(4))) double* __restrict__ W =
          (const __attribute__((address_space(4))) double*)(wbase) + idx;
      int widx = 0;
      const int r = m2 * U1 + m1;
      if (r + PF < NR) load_row((r + PF) / U1, (r + PF) % U1, rows[(r + PF) % (PF + 1)]);
      double(&row)[NV][U0] = rows[r % (PF + 1)];
#pragma unroll
      for (int c = 0; c < R2; ++c) {
        const int k2 = m2 - c * SHIFT;
        if (k2 < 0 || k2 >= P) continue;
#pragma unroll
        for (int b = 0; b < R1; ++b) {
          const int k1 = m1 - b * SHIFT;
          if (k1 < 0 || k1 >= P) continue;
#pragma unroll
          for (int a = 0; a < R0; ++a) {
#pragma unroll
            for (int m0 = 0; m0 < U0; ++m0) {
              const int k0 = m0 - a * SHIFT;
              if (k0 < 0 || k0 >= P) continue;
              const double w = W[widx];
#pragma unroll
              for (int n = 0; n < NV; ++n) acc[n][a][b][c] = __dadd_rn(acc[n][a][b][c], __dmul_rn(w, row[n][m0]));
              ++widx;
            }
          }
        }
      }
      idx += widx;
      // Every accumulator is brought up to date row by row: without this
      // the compiler defers whole chains to the end and keeps every row's
      // source values and weights live until then, spilling them.
#pragma unroll
      for (int n = 0; n < NV; ++n)
#pragma unroll
        for (int a = 0; a < R0; ++a)
#pragma unroll
          for (int b = 0; b < R1; ++b)
#pragma unroll
            for (int c = 0; c < R2; ++c) asm volatile("" : "+v"(acc[n][a][b][c]));
    }
  }
  double* __restrict__ dst = job.work + s4 * g.tblocks[tr] + s3 * v;
  const int32_t nvars = job.nvars;
#pragma unroll
  for (int c = 0; c < R2; ++c)
#pragma unroll
    for (int b = 0; b < R1; ++b)
#pragma unroll
      for (int a = 0; a < R0; ++a) {
        double* q = dst + (s0 * (g.tfirst[0] + ra[a]) + s1 * (g.tfirst[1] + rb[b]) + s2 * (g.tfirst[2] + rc[c]));
        const double x = scaled(acc[0][a][b][c], job.factors, g.factorcol, nvars, v);
        if (MODE == 2) {
          pair_t out;
          out.x = x;
          out.y = scaled(acc[NV - 1][a][b][c], job.factors, g.factorcol, nvars, v + 1);
          *reinterpret_cast<pair_t*>(q) = out;
        } else {
          q[0] = x;
          if (MODE == 1 && has2) q[1] = scaled(acc[NV - 1][a][b][c], job.factors, g.factorcol, nvars, v + 1);
        }
      }
}

#ifdef TAH_DEFINE_KERNELS
// One unit class per launch.
template <int P, int SHIFT, int R0, int R1, int R2>
__global__ void __launch_bounds__(kThreads) interp_kernel(const Job* __restrict__ jobs,
                                                          const Tile* __restrict__ tiles) {
  const Tile tile = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const Job& job = jobs[tile.job];
  interp_body<P, SHIFT, R0, R1, R2>(job, tile, job.wclasses[tile.item]);
}

// Every unit class of one order and shift in one launch, the class picked
// per tile (a branch uniform over the workgroup), so that the tiles of all
// classes of one transfer -- which read the same source windows -- can run
// side by side and meet in one L2.
// Prolongation's registers are capped for seven waves per SIMD (72 VGPRs:
// no spills at order 6, where the unconstrained octet body takes 94);
// restriction's for four (its ring of row buffers: 115 VGPRs at order 6,
// no spills).
template <int P, int SHIFT>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(SHIFT != 0 ? kRestrictWaves : kProlongWaves, 8)))
interp_all_kernel(const Job* __restrict__ jobs, const Tile* __restrict__ tiles) {
  const Tile tile = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const Job& job = jobs[tile.job];
  const DevWClass& wc = job.wclasses[tile.item];
  switch (uniform(wc.rbits)) {
    case 0: interp_body<P, SHIFT, 1, 1, 1>(job, tile, wc); break;
    case 1: interp_body<P, SHIFT, 2, 1, 1>(job, tile, wc); break;
    case 2: interp_body<P, SHIFT, 1, 2, 1>(job, tile, wc); break;
    case 3: interp_body<P, SHIFT, 2, 2, 1>(job, tile, wc); break;
    case 4: interp_body<P, SHIFT, 1, 1, 2>(job, tile, wc); break;
    case 5: interp_body<P, SHIFT, 2, 1, 2>(job, tile, wc); break;
    case 6: interp_body<P, SHIFT, 1, 2, 2>(job, tile, wc); break;
    default: interp_body<P, SHIFT, 2, 2, 2>(job, tile, wc); break;
  }
}

// The merged launch over tiles of variable pairs (variables innermost, an
// even count): 16-byte loads and stores where the arrays are 16-byte
// aligned. A pair halves the addresses the vector memory unit processes per
// value, the unit that bounds the single-variable kernel, and each weight
// read serves both variables. Registers capped for five waves per SIMD.
// Prolongation only (SHIFT 0): over pairs, restriction was slower.
template <int MODE, int P, int SHIFT>
__device__ __forceinline__ void interp_pair_dispatch(const Job& job, const Tile& tile, const DevWClass& wc) {
  switch (uniform(wc.rbits)) {
    case 0: interp_body<P, SHIFT, 1, 1, 1, MODE>(job, tile, wc); break;
    case 1: interp_body<P, SHIFT, 2, 1, 1, MODE>(job, tile, wc); break;
    case 2: interp_body<P, SHIFT, 1, 2, 1, MODE>(job, tile, wc); break;
    case 3: interp_body<P, SHIFT, 2, 2, 1, MODE>(job, tile, wc); break;
    case 4: interp_body<P, SHIFT, 1, 1, 2, MODE>(job, tile, wc); break;
    case 5: interp_body<P, SHIFT, 2, 1, 2, MODE>(job, tile, wc); break;
    case 6: interp_body<P, SHIFT, 1, 2, 2, MODE>(job, tile, wc); break;
    default: interp_body<P, SHIFT, 2, 2, 2, MODE>(job, tile, wc); break;
  }
}

template <int P, int SHIFT>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kPairWaves, 8)))
interp_pair_kernel(const Job* __restrict__ jobs, const Tile* __restrict__ tiles) {
  static_assert(SHIFT == 0, "variable pairs are for prolongation");
  const Tile tile = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const Job& job = jobs[tile.job];
  const DevWClass& wc = job.wclasses[tile.item];
  if (job.pairs) interp_pair_dispatch<2, P, SHIFT>(job, tile, wc);
  else interp_pair_dispatch<1, P, SHIFT>(job, tile, wc);
}

#endif  // TAH_DEFINE_KERNELS

// ---------------------------------------------------------------------------
// The work queue of a batched fill: every tile of one stage position of every
// plan of the batch -- copies, interpolations of the orders below, generic
// transfers -- in one launch of a persistent grid. The tiles are cut into
// eight contiguous shares, one per XCD (a workgroup's XCD is its index mod 8,
// as the dispatcher deals them); the workgroups of an XCD take the tiles of
// its share in order from an atomic counter of its own until the share is
// used up, so that small classes fill the gaps of large ones instead of each
// class's launch ending on a tail. Tile.pad names the tile's kernel: a
// kQueue* code.
constexpr int kQueueCopy = 0;     // copy_body<8>
constexpr int kQueueGeneric = 1;  // the generic tile
constexpr int kQueueCopy4 = 2;    // copy_body<4>
constexpr int kQueueCopy16 = 3;   // copy_body<16>
constexpr int kQueueMerged = 16;  // + index into kInterpTable: one variable per thread
constexpr int kQueuePair = 32;    // + index: two variables per thread (prolongation)
constexpr int kQueueWaves = 4;    // waves per SIMD the queue kernel is compiled for

#ifdef TAH_DEFINE_KERNELS
__device__ __forceinline__ void generic_body(const Job& job, const Tile& tile) {
  const DevGroup& g = job.groups[tile.item];
  const int64_t s0 = job.stride[0], s1 = job.stride[1], s2 = job.stride[2], s3 = job.stride[3],
                s4 = job.stride[4];
  const int32_t nvars = job.nvars;
  const uint32_t total = uint32_t(g.points) * uint32_t(nvars);
  const double* sbase = job.src + s4 * g.sblocks[tile.transfer];
  double* tbase = job.work + s4 * g.tblocks[tile.transfer];
  for (int r = 0; r < kGenericPerThread; ++r) {
    uint32_t e = uint32_t(tile.e0) + uint32_t(r * kThreads) + threadIdx.x;
    if (e >= total) break;
    int32_t i, j, k, v;
    split_element(g, job, e, i, j, k, v);
    double acc = stencil_value(g, sbase + s3 * v, s0, s1, s2, i, j, k);
    acc = scaled(acc, job.factors, g.factorcol, nvars, v);
    tbase[s3 * v + s0 * (g.tfirst[0] + i) + s1 * (g.tfirst[1] + j) + s2 * (g.tfirst[2] + k)] = acc;
  }
}

__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kQueueWaves, 8)))
queue_kernel(const Job* __restrict__ jobs, const Tile* __restrict__ tiles, const uint32_t* __restrict__ share,
             uint32_t* __restrict__ counters) {
  __shared__ uint32_t slot;
  const uint32_t x = blockIdx.x % 8;
  const uint32_t lo = share[x], hi = share[x + 1];
  for (;;) {
    if (threadIdx.x == 0) slot = lo + atomicAdd(counters + x, 1u);
    __syncthreads();
    const uint32_t it = slot;
    __syncthreads();  // every thread has read slot before thread 0 writes it again
    if (it >= hi) break;  // the same for the whole workgroup
    const Tile tile = tiles[it];
    const Job& job = jobs[tile.job];
    const int32_t kind = uniform(tile.pad);
    if (kind == kQueueCopy) {
      copy_body<8>(job, tile);
    } else if (kind == kQueueCopy4) {
      copy_body<4>(job, tile);
    } else if (kind == kQueueCopy16) {
      copy_body<16>(job, tile);
    } else if (kind == kQueueGeneric) {
      generic_body(job, tile);
    } else {
      const DevWClass& wc = job.wclasses[tile.item];
      switch (kind) {
        case kQueueMerged + 0: interp_pair_dispatch<0, 2, 0>(job, tile, wc); break;
        case kQueueMerged + 2: interp_pair_dispatch<0, 4, 0>(job, tile, wc); break;
        case kQueueMerged + 4: interp_pair_dispatch<0, 6, 0>(job, tile, wc); break;
        case kQueueMerged + 5: interp_pair_dispatch<0, 2, 2>(job, tile, wc); break;
        case kQueueMerged + 6: interp_pair_dispatch<0, 4, 2>(job, tile, wc); break;
        case kQueueMerged + 7: interp_pair_dispatch<0, 6, 2>(job, tile, wc); break;
        case kQueuePair + 0:
          if (job.pairs) interp_pair_dispatch<2, 2, 0>(job, tile, wc);
          else interp_pair_dispatch<1, 2, 0>(job, tile, wc);
          break;
        case kQueuePair + 2:
          if (job.pairs) interp_pair_dispatch<2, 4, 0>(job, tile, wc);
          else interp_pair_dispatch<1, 4, 0>(job, tile, wc);
          break;
        default:
          if (job.pairs) interp_pair_dispatch<2, 6, 0>(job, tile, wc);
          else interp_pair_dispatch<1, 6, 0>(job, tile, wc);
          break;
      }
    }
  }
}

// The queue's counters to zero (a kernel, not a memset, so that the batch
// can be captured in a graph and replayed).
__global__ void queue_reset_kernel(uint32_t* __restrict__ counters, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) counters[i] = 0u;
}
#endif  // TAH_DEFINE_KERNELS

// ---------------------------------------------------------------------------
// Interpolation from an LDS copy of the source box (restriction and
// prolongation of order P in every dimension, variables innermost in even
// number). A workgroup takes one chunk of one transfer for one pair of
// variables: it copies the chunk's source box into LDS once, with 16-byte
// loads in the array's own order, and then every thread computes targets of
// the chunk for both variables, reading its taps from LDS. Each source value
// thus crosses the vector memory path once per chunk instead of once per
// tap, and neighbouring targets' overlapping windows are read from LDS.
// A target's weight is formed as the reference forms it, ((w1*w2)*w3) from
// the three one-dimensional stencils, and its taps are summed in the
// reference's order (m3, m2, m1 ascending, m1 innermost) from acc = +0 with
// acc += w*u, every operation rounded on its own: bit for bit the other
// kernels' results.

#ifdef TAH_DEFINE_KERNELS
template <int P>
__global__ void __launch_bounds__(kThreads) interp_lds_kernel(const Job* __restrict__ jobs,
                                                              const LdsTile* __restrict__ tiles) {
  typedef double pair_t __attribute__((ext_vector_type(2)));
  extern __shared__ __align__(16) double lds[];
  const LdsTile tile = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const Job& job = jobs[tile.job];
  const LdsChunk& ch = job.chunks[tile.chunk];
  const DevGroup& g = job.groups[ch.group];
  const int32_t s0 = int32_t(job.stride[0]), s1 = int32_t(job.stride[1]), s2 = int32_t(job.stride[2]);
  const int64_t s4 = job.stride[4];
  const int32_t e0 = ch.sext[0], e1 = ch.sext[1];
  const int32_t p1 = ch.pitch[0], p2 = ch.pitch[1];
  const int32_t h0 = ch.half[0], h1 = ch.half[1], h2 = ch.half[2];
  // Stage the source box: point (x, y, z) of the box at pair X + Y*p1 + Z*p2,
  // X = (x & 1)*h0 + (x >> 1) and so on (even rows first, then odd ones).
  {
    const double* __restrict__ src = job.src + s4 * g.sblocks[tile.transfer] +
                                     (tile.v0 + s0 * ch.slo[0] + s1 * ch.slo[1] + s2 * ch.slo[2]);
    const uint32_t n = uint32_t(ch.staged);
    for (uint32_t q = threadIdx.x; q < n; q += kThreads) {
      const uint32_t r = ch.dsx[0].div(q);
      const int32_t x = int32_t(q - r * uint32_t(e0));
      const uint32_t z = ch.dsx[1].div(r);
      const int32_t y = int32_t(r - z * uint32_t(e1));
      const double* p = src + (s2 * int32_t(z) + s1 * y + s0 * x);
      pair_t v;
      if (job.pairs) {
        v = *reinterpret_cast<const pair_t*>(p);
      } else {
        v.x = p[0];
        v.y = p[1];
      }
      const int32_t zi = int32_t(z);
      reinterpret_cast<pair_t*>(lds)[((x & 1) * h0 + (x >> 1)) + ((y & 1) * h1 + (y >> 1)) * p1 +
                                     ((zi & 1) * h2 + (zi >> 1)) * p2] = v;
    }
  }
  __syncthreads();
  const pair_t* __restrict__ box = reinterpret_cast<const pair_t*>(lds);
  double* __restrict__ dst = job.work + s4 * g.tblocks[tile.transfer] + tile.v0;
  const int32_t nvars = job.nvars;
  const int32_t d0 = ch.ord[0], d1 = ch.ord[1];
  const int32_t n0 = ch.tn[d0], n1 = ch.tn[d1];
  for (uint32_t q = threadIdx.x; q < uint32_t(ch.points); q += kThreads) {
    // target rows (i, j, k) of the group's box: a, b, c along ord[0..2]
    const uint32_t r = ch.dt[0].div(q);
    const int32_t a = int32_t(q - r * uint32_t(n0));
    const uint32_t c = ch.dt[1].div(r);
    const int32_t b = int32_t(r - c * uint32_t(n1));
    const int32_t i = ch.tlo[0] + (d0 == 0 ? a : d1 == 0 ? b : int32_t(c));
    const int32_t j = ch.tlo[1] + (d0 == 1 ? a : d1 == 1 ? b : int32_t(c));
    const int32_t k = ch.tlo[2] + (d0 == 2 ? a : d1 == 2 ? b : int32_t(c));
    double w0[P], w1[P], w2[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
      w0[m] = g.weights[0][P * i + m];
      w1[m] = g.weights[1][P * j + m];
      w2[m] = g.weights[2][P * k + m];
    }
    const int32_t o0 = g.srcstart[0][i] - ch.slo[0], o1 = g.srcstart[1][j] - ch.slo[1],
                  o2 = g.srcstart[2][k] - ch.slo[2];
    // every window starts an even number of rows into the box (the chunk
    // builder checks), so tap m of a dimension is at half (m & 1), row
    // o/2 + m/2 there
    const pair_t* __restrict__ base = box + ((o2 >> 1) * p2 + (o1 >> 1) * p1 + (o0 >> 1));
    double ax = 0.0, ay = 0.0;
#pragma unroll
    for (int m2 = 0; m2 < P; ++m2) {
      const double c2 = w2[m2];
#pragma unroll
      for (int m1 = 0; m1 < P; ++m1) {
        const double c1 = w1[m1];
        const pair_t* __restrict__ row =
            base + (((m2 & 1) * h2 + (m2 >> 1)) * p2 + ((m1 & 1) * h1 + (m1 >> 1)) * p1);
        const pair_t* __restrict__ odd = row + h0;
#pragma unroll
        for (int m0 = 0; m0 < P; ++m0) {
          const double w = __dmul_rn(__dmul_rn(w0[m0], c1), c2);
          const pair_t u = (m0 & 1) ? odd[m0 >> 1] : row[m0 >> 1];
          ax = __dadd_rn(ax, __dmul_rn(w, u.x));
          ay = __dadd_rn(ay, __dmul_rn(w, u.y));
        }
        // Row by row: without this the compiler hoists every row's LDS
        // reads to the top and keeps them all live, spilling them.
        asm volatile("" : "+v"(ax), "+v"(ay));
      }
    }
    double* __restrict__ t = dst + (s0 * (g.tfirst[0] + i) + s1 * (g.tfirst[1] + j) + s2 * (g.tfirst[2] + k));
    pair_t out;
    out.x = scaled(ax, job.factors, g.factorcol, nvars, tile.v0);
    out.y = scaled(ay, job.factors, g.factorcol, nvars, tile.v0 + 1);
    if (job.pairs) {
      *reinterpret_cast<pair_t*>(t) = out;
    } else {
      t[0] = out.x;
      t[1] = out.y;
    }
  }
}
#endif  // TAH_DEFINE_KERNELS

// Prolongation from an LDS copy of a chunk's coarse source box, by units. A
// lane computes one unit (R0 x R1 x R2 fine targets sharing a coarse
// window) for a pair of variables, exactly as interp_body does -- the same
// taps in the same order, the weight products of its class's table W read
// by scalar loads (each wave holds units of one class) -- but reads the
// window's rows from LDS.
template <int P, int R0, int R1, int R2>
__device__ __forceinline__ void lds_unit_body(const Job& job, const DevGroup& g, const DevWClass& wc,
                                              const LdsUnitChunk& ch, const double* __restrict__ lds, uint32_t pk0,
                                              uint32_t pk1, int32_t tr, int32_t v0) {
  typedef double pair_t __attribute__((ext_vector_type(2)));
  const int32_t ra[2] = {int32_t(pk0 & 255u), int32_t((pk0 >> 8) & 255u)};
  const int32_t rb[2] = {int32_t((pk0 >> 16) & 255u), int32_t(pk0 >> 24)};
  const int32_t rc[2] = {int32_t(pk1 & 255u), int32_t((pk1 >> 8) & 255u)};
  const int32_t p1 = ch.pitch[0], p2 = ch.pitch[1];
  const pair_t* __restrict__ src = reinterpret_cast<const pair_t*>(lds) +
                                   ((g.srcstart[2][rc[0]] - ch.slo[2]) * p2 + (g.srcstart[1][rb[0]] - ch.slo[1]) * p1 +
                                    (g.srcstart[0][ra[0]] - ch.slo[0]));
  uint64_t wbase = reinterpret_cast<uint64_t>(uniform(wc.W));
  double acc[2][R0][R1][R2];
#pragma unroll
  for (int n = 0; n < 2; ++n)
#pragma unroll
    for (int a = 0; a < R0; ++a)
#pragma unroll
      for (int b = 0; b < R1; ++b)
#pragma unroll
        for (int c = 0; c < R2; ++c) acc[n][a][b][c] = 0.0;
  int idx = 0;
#pragma unroll
  for (int m2 = 0; m2 < P; ++m2) {
#pragma unroll
    for (int m1 = 0; m1 < P; ++m1) {
      asm volatile("" : "+s"(wbase));
      const __attribute__((address_space(4))) double* __restrict__ W =
          (const __attribute__((address_space(4))) double*)(wbase) + idx;
      const pair_t* __restrict__ rowp = src + (m2 * p2 + m1 * p1);
      double row[2][P];
#pragma unroll
      for (int m0 = 0; m0 < P; ++m0) {
        const pair_t x = rowp[m0];
        row[0][m0] = x.x;
        row[1][m0] = x.y;
      }
      // the order of interp_body with SHIFT 0: targets (c, b, a), then m0
      int widx = 0;
#pragma unroll
      for (int c = 0; c < R2; ++c)
#pragma unroll
        for (int b = 0; b < R1; ++b)
#pragma unroll
          for (int a = 0; a < R0; ++a)
#pragma unroll
            for (int m0 = 0; m0 < P; ++m0) {
              const double w = W[widx];
#pragma unroll
              for (int n = 0; n < 2; ++n) acc[n][a][b][c] = __dadd_rn(acc[n][a][b][c], __dmul_rn(w, row[n][m0]));
              ++widx;
            }
      idx += widx;
#pragma unroll
      for (int n = 0; n < 2; ++n)
#pragma unroll
        for (int a = 0; a < R0; ++a)
#pragma unroll
          for (int b = 0; b < R1; ++b)
#pragma unroll
            for (int c = 0; c < R2; ++c) asm volatile("" : "+v"(acc[n][a][b][c]));
    }
  }
  const int32_t s0 = int32_t(job.stride[0]), s1 = int32_t(job.stride[1]), s2 = int32_t(job.stride[2]);
  double* __restrict__ dst = job.work + job.stride[4] * g.tblocks[tr] + v0;
  const int32_t nvars = job.nvars;
#pragma unroll
  for (int c = 0; c < R2; ++c)
#pragma unroll
    for (int b = 0; b < R1; ++b)
#pragma unroll
      for (int a = 0; a < R0; ++a) {
        double* q = dst + (s0 * (g.tfirst[0] + ra[a]) + s1 * (g.tfirst[1] + rb[b]) + s2 * (g.tfirst[2] + rc[c]));
        pair_t out;
        out.x = scaled(acc[0][a][b][c], job.factors, g.factorcol, nvars, v0);
        out.y = scaled(acc[1][a][b][c], job.factors, g.factorcol, nvars, v0 + 1);
        if (job.pairs) {
          *reinterpret_cast<pair_t*>(q) = out;
        } else {
          q[0] = out.x;
          q[1] = out.y;
        }
      }
}

constexpr int kLdsProlongThreads = 128;  // two waves: a chunk's class waves in turn

#ifdef TAH_DEFINE_KERNELS
template <int P>
__global__ void __launch_bounds__(kLdsProlongThreads) __attribute__((amdgpu_waves_per_eu(kPairWaves, 8)))
interp_lds_prolong_kernel(const Job* __restrict__ jobs, const LdsTile* __restrict__ tiles) {
  typedef double pair_t __attribute__((ext_vector_type(2)));
  extern __shared__ __align__(16) double lds[];
  const LdsTile tile = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const Job& job = jobs[tile.job];
  const LdsUnitChunk& ch = job.uchunks[tile.chunk];
  const DevGroup& g = job.groups[ch.group];
  {
    const int32_t s0 = int32_t(job.stride[0]), s1 = int32_t(job.stride[1]), s2 = int32_t(job.stride[2]);
    const double* __restrict__ src = job.src + job.stride[4] * g.sblocks[tile.transfer] +
                                     (tile.v0 + s0 * ch.slo[0] + s1 * ch.slo[1] + s2 * ch.slo[2]);
    const int32_t e0 = ch.sext[0], e1 = ch.sext[1], p1 = ch.pitch[0], p2 = ch.pitch[1];
    for (uint32_t q = threadIdx.x; q < uint32_t(ch.staged); q += kLdsProlongThreads) {
      const uint32_t r = ch.dsx[0].div(q);
      const int32_t x = int32_t(q - r * uint32_t(e0));
      const uint32_t z = ch.dsx[1].div(r);
      const int32_t y = int32_t(r - z * uint32_t(e1));
      const double* p = src + (s2 * int32_t(z) + s1 * y + s0 * x);
      pair_t v;
      if (job.pairs) {
        v = *reinterpret_cast<const pair_t*>(p);
      } else {
        v.x = p[0];
        v.y = p[1];
      }
      reinterpret_cast<pair_t*>(lds)[x + y * p1 + int32_t(z) * p2] = v;
    }
  }
  __syncthreads();
  const int32_t hw = int32_t(threadIdx.x) / 64, lane = int32_t(threadIdx.x) % 64;
  for (int32_t w = hw; w < ch.nwave; w += kLdsProlongThreads / 64) {
    const LdsWave& wv = job.waves[ch.wave0 + w];
    if (lane >= wv.nu) continue;
    const uint32_t pk0 = job.units[2 * (wv.unit0 + lane)], pk1 = job.units[2 * (wv.unit0 + lane) + 1];
    const DevWClass& wc = job.wclasses[wv.wclass];
    switch (uniform(wc.rbits)) {
      case 0: lds_unit_body<P, 1, 1, 1>(job, g, wc, ch, lds, pk0, pk1, tile.transfer, tile.v0); break;
      case 1: lds_unit_body<P, 2, 1, 1>(job, g, wc, ch, lds, pk0, pk1, tile.transfer, tile.v0); break;
      case 2: lds_unit_body<P, 1, 2, 1>(job, g, wc, ch, lds, pk0, pk1, tile.transfer, tile.v0); break;
      case 3: lds_unit_body<P, 2, 2, 1>(job, g, wc, ch, lds, pk0, pk1, tile.transfer, tile.v0); break;
      case 4: lds_unit_body<P, 1, 1, 2>(job, g, wc, ch, lds, pk0, pk1, tile.transfer, tile.v0); break;
      case 5: lds_unit_body<P, 2, 1, 2>(job, g, wc, ch, lds, pk0, pk1, tile.transfer, tile.v0); break;
      case 6: lds_unit_body<P, 1, 2, 2>(job, g, wc, ch, lds, pk0, pk1, tile.transfer, tile.v0); break;
      default: lds_unit_body<P, 2, 2, 2>(job, g, wc, ch, lds, pk0, pk1, tile.transfer, tile.v0); break;
    }
  }
}
#endif  // TAH_DEFINE_KERNELS

// ---------------------------------------------------------------------------
// Restriction of order P streamed along z through LDS. A workgroup takes one
// chunk of one transfer (target rows along x and y, every row along z) for a
// tile of variables and walks the source planes of its windows in ascending
// z. Each plane -- the union of the chunk's windows along x and y, every
// variable of the tile of each point, so that the global loads read whole
// runs of a point's variables -- is staged in LDS once; a thread holds one
// target (i, j) for a pair of variables, with one accumulator per z row
// whose window holds the plane (at most KA: the rows sorted by window
// start enter and leave in order, as the group's plane records say). For
// each such row k the plane is tap m3 = z - srcstart[2][k] of its window;
// the thread adds the plane's taps in the order m2, then m1, each as
// acc += ((w1*w2)*w3)*u. So every target sees its taps in the reference's
// order (m3, m2, m1 ascending, m1 innermost), from +0, every operation
// rounded on its own (stencil_sum, src/ghosts.jl:58-118), and each fine
// value crosses the vector memory path once per chunk instead of once per
// tap.
//
// ASYNC (16-byte aligned pairs of variables innermost): two plane buffers;
// the plane after next is loaded straight into LDS (global_load_lds, no
// registers) while the threads sum the current one, one barrier a plane.
// Otherwise one buffer, staged through registers between two barriers.
#ifdef TAH_DEFINE_KERNELS
template <int P, int KA, bool ASYNC>
__device__ __forceinline__ void stream_body(const Job& job, const LdsTile& tile, double* __restrict__ lds) {
  typedef double pair_t __attribute__((ext_vector_type(2)));
  const StreamChunk& ch = job.schunks[tile.chunk];
  const DevGroup& g = job.groups[ch.group];
  const int32_t s0 = int32_t(job.stride[0]), s1 = int32_t(job.stride[1]), s2 = int32_t(job.stride[2]),
                s3 = int32_t(job.stride[3]);
  const int64_t s4 = job.stride[4];
  const int32_t nvars = job.nvars;
  const int32_t v0 = tile.v0;
  const int32_t nv = min(ch.nvt, nvars - v0);
  const int32_t L = (nv + 1) >> 1;  // threads per target: one per pair of variables
  const int32_t nvp = 2 * L;        // doubles per staged point
  const int32_t e0 = ch.sext[0], e1 = ch.sext[1];
  const int32_t t = int32_t(threadIdx.x), T = int32_t(blockDim.x);
  const bool pairs = job.pairs != 0;

  // The plane's pairs, numbered with a point's pairs fastest where variables
  // are innermost, else x fastest; thread t stages pairs t, t + T, ... (T
  // the workgroup's threads)
  // (at most kStreamSlots), pair vp of point (x, y) at LDS double
  // (y*e0 + x)*nvp + 2*vp -- in the numbering's own order where variables
  // are innermost. Odd counts: the last pair's second value is staged as
  // zero.
  const int32_t npairs = e0 * e1 * L;
  const int32_t plane = e0 * e1 * nvp;  // doubles of one plane buffer
  const bool vfast = job.vfast != 0;
  const int32_t A = vfast ? L : e0, B = vfast ? e0 : e1;
  const FastDiv dA{uint32_t(A)}, dB{uint32_t(B)};
  const double* __restrict__ src = job.src + s4 * g.sblocks[tile.transfer] +
                                   (s3 * v0 + s0 * ch.slo[0] + s1 * ch.slo[1]);
  // pair q's offset from the plane's first staged point and the tile's
  // first variable, and its LDS double
  auto slot_of = [&](int32_t q, int32_t& go, int32_t& at, int32_t& vp) {
    const int32_t r = int32_t(dA.div(uint32_t(q))), c0 = q - r * A;
    const int32_t c2 = int32_t(dB.div(uint32_t(r))), c1 = r - c2 * B;
    vp = vfast ? c0 : c2;
    const int32_t x = vfast ? c1 : c0, y = vfast ? c2 : c1;
    go = s1 * y + s0 * x + s3 * (2 * vp);
    at = (y * e0 + x) * nvp + 2 * vp;
  };
  auto stage_sync = [&](int32_t z) {
    const __attribute__((address_space(1))) double* __restrict__ p =
        (const __attribute__((address_space(1))) double*)(src + int64_t(s2) * z);
    pair_t val[kStreamSlots];
    int32_t at[kStreamSlots];
    // (the slots' offsets formed anew on every plane: kept across planes,
    // they would hold three registers a slot)
    int32_t tt = t;
    asm volatile("" : "+v"(tt));
#pragma unroll
    for (int s = 0; s < kStreamSlots; ++s) {
      const int32_t q = tt + s * T;
      at[s] = -1;
      if (q >= npairs) continue;
      int32_t go, vp;
      slot_of(q, go, at[s], vp);
      if (pairs) {
        val[s] = *(const __attribute__((address_space(1))) pair_t*)(p + go);
      } else {
        val[s].x = p[go];
        val[s].y = 2 * vp + 1 < nv ? p[go + s3] : 0.0;
      }
    }
#pragma unroll
    for (int s = 0; s < kStreamSlots; ++s)
      if (at[s] >= 0) *reinterpret_cast<pair_t*>(lds + at[s]) = val[s];
  };
  // ASYNC: each wave loads its 64 consecutive pairs of a slot straight into
  // LDS, lane l's 16 bytes at the wave's base + 16*l (the numbering's order).
  auto stage_async = [&](int32_t z, int32_t buf) {
    const double* __restrict__ p = src + int64_t(s2) * z;
    int32_t tt = t;
    asm volatile("" : "+v"(tt));
    const int32_t wave = tt & ~63;
#pragma unroll
    for (int s = 0; s < kStreamSlots; ++s) {
      const int32_t q0 = wave + s * T;  // the wave's first pair: uniform
      if (q0 >= npairs) break;
      const int32_t q = tt + s * T;
      if (q < npairs) {
        int32_t go, at, vp;
        slot_of(q, go, at, vp);
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(p + go),
                                         (__attribute__((address_space(3))) void*)(lds + buf * plane + 2 * q0), 16, 0,
                                         0);
      }
    }
  };

  // This thread's target and pair of variables; threads past the chunk's
  // targets only stage.
  const int32_t item = t / L, vl = t - item * L;
  const bool active = item < ch.tn[0] * ch.tn[1];
  const int32_t b = active ? item / ch.tn[0] : 0;
  const int32_t a = active ? item - b * ch.tn[0] : 0;
  const int32_t i = ch.tlo[0] + a, j = ch.tlo[1] + b;
  const int32_t ox = g.srcstart[0][i] - ch.slo[0], oy = g.srcstart[1][j] - ch.slo[1];
  double w0[P], w1[P];
#pragma unroll
  for (int m = 0; m < P; ++m) {
    w0[m] = g.weights[0][P * i + m];
    w1[m] = g.weights[1][P * j + m];
  }
  const int32_t tap0 = (oy * e0 + ox) * nvp + 2 * vl;  // the window's first tap in a plane buffer (doubles)
  const int32_t v = v0 + 2 * vl;
  const bool has2 = 2 * vl + 1 < nv;
  double* __restrict__ dst = job.work + s4 * g.tblocks[tile.transfer] +
                             (s3 * v + s0 * (g.tfirst[0] + i) + s1 * (g.tfirst[1] + j));
  int32_t tz = g.tfirst[2];
  asm volatile("" : "+v"(tz));  // read once, not at each store
  // the parity factors of the thread's variables, read once (scaled():
  // acc*f + 0 where the group has a factor column)
  const int32_t fcol = g.factorcol;
  double fx = 1.0, fy = 1.0;
  if (fcol != 0) {
    fx = job.factors[v + int64_t(nvars) * (fcol - 1)];
    fy = has2 ? job.factors[v + 1 + int64_t(nvars) * (fcol - 1)] : 0.0;
  }
  // the plane records by scalar loads (uniform)
  const __attribute__((address_space(4))) StreamPlane* __restrict__ planes = constant(job.planes + ch.plane0);
  const int32_t np = ch.nplanes, zlo = ch.zlo;

  double ax[KA], ay[KA];
#pragma unroll
  for (int q = 0; q < KA; ++q) ax[q] = ay[q] = 0.0;
  if (ASYNC) stage_async(zlo, 0);
  for (int32_t pz = 0; pz < np; ++pz) {
    const int32_t z = zlo + pz;
    const __attribute__((address_space(4))) StreamPlane& pl = planes[pz];
    const int32_t had = pl.had, nact = pl.nact, nret = pl.nret;
    // rows whose windows start here join, at +0
#pragma unroll
    for (int q = 0; q < KA; ++q)
      if (q >= had && q < nact) ax[q] = ay[q] = 0.0;
    int32_t buf = 0;
    if (ASYNC) {
      // this thread's loads of plane z are done; after the barrier every
      // thread's are, and every thread is done with plane z - 1, whose
      // buffer plane z + 1 then fills
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
      __syncthreads();
      buf = pz & 1;
      if (pz + 1 < np) stage_async(z + 1, buf ^ 1);
    } else {
      __syncthreads();  // every thread is done with the previous plane
      if (nact > 0) stage_sync(z);
      __syncthreads();
    }
    if (active && nact > 0) {
      // (the tap addresses formed anew on every plane: kept across planes,
      // they would hold P*P registers)
      int32_t taps = tap0 + buf * plane;
      asm volatile("" : "+v"(taps));
      double c2[KA];
#pragma unroll
      for (int q = 0; q < KA; ++q) c2[q] = pl.c2[q];
#pragma unroll
      for (int m2 = 0; m2 < P; ++m2) {
        const double* __restrict__ row = lds + (taps + m2 * e0 * nvp);
        pair_t u[P];
        double w01[P];
        // (formed anew on every plane: kept across planes, the row
        // products would hold 2*P*P registers)
        double c1 = w1[m2];
        asm volatile("" : "+v"(c1));
#pragma unroll
        for (int m1 = 0; m1 < P; ++m1) {
          u[m1] = *reinterpret_cast<const pair_t*>(row + m1 * nvp);
          w01[m1] = __dmul_rn(w0[m1], c1);
        }
#pragma unroll
        for (int q = 0; q < KA; ++q) {
          if (q >= nact) continue;
#pragma unroll
          for (int m1 = 0; m1 < P; ++m1) {
            const double w = __dmul_rn(w01[m1], c2[q]);
            ax[q] = __dadd_rn(ax[q], __dmul_rn(w, u[m1].x));
            ay[q] = __dadd_rn(ay[q], __dmul_rn(w, u[m1].y));
          }
        }
        // Row by row: without this the compiler hoists every row's LDS
        // reads to the top of the plane and keeps them all live.
#pragma unroll
        for (int q = 0; q < KA; ++q) asm volatile("" : "+v"(ax[q]), "+v"(ay[q]));
      }
    }
    // rows whose windows end here are done: the first nret slots, stored,
    // then the rest move down
    if (nret > 0) {
#pragma unroll
      for (int q = 0; q < KA; ++q) {
        if (q >= nret || !active) continue;
        double* __restrict__ o = dst + int64_t(s2) * (tz + pl.k[q]);
        pair_t out;
        out.x = fcol != 0 ? __dadd_rn(__dmul_rn(ax[q], fx), 0.0) : ax[q];
        out.y = fcol != 0 ? __dadd_rn(__dmul_rn(ay[q], fy), 0.0) : ay[q];
        if (pairs) {
          *reinterpret_cast<pair_t*>(o) = out;
        } else {
          o[0] = out.x;
          if (has2) o[s3] = out.y;
        }
      }
#pragma unroll
      for (int r = 0; r < KA; ++r) {
        if (r >= nret) break;
#pragma unroll
        for (int q = 0; q + 1 < KA; ++q) {
          ax[q] = ax[q + 1];
          ay[q] = ay[q + 1];
        }
      }
    }
  }
}

// LDS: two plane buffers (the launch's largest plane twice).
template <int P, int KA>
__global__ void __launch_bounds__(kThreads)
interp_stream_kernel(const Job* __restrict__ jobs, const LdsTile* __restrict__ tiles) {
  extern __shared__ __align__(16) double lds[];
  const LdsTile tile = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const Job& job = jobs[tile.job];
  if (job.pairs && job.vfast) stream_body<P, KA, true>(job, tile, lds);
  else stream_body<P, KA, false>(job, tile, lds);
}
#endif  // TAH_DEFINE_KERNELS

}  // namespace tah
