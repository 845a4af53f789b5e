// The fill kernels for gfx950.
//
// Every value is computed as the reference's transfer_kernel! computes it
// (src/ghosts.jl:372-402): a unit copy is 0 + u (copy_load, :127-128); a
// stencil is one accumulator from +0 with m1 innermost and the tap weight
// ((w1*w2)*w3) added as acc += w*u (stencil_sum, :58-118), every product and
// sum rounded on its own (_rn intrinsics; the file is also compiled with
// -ffp-contract=off); a factor is acc*f + 0 (scaled, :369-370). Kernels that
// compute several targets per thread keep one accumulator per target and
// visit each target's taps in that same order, so they change no bit.
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace tah {

__device__ __forceinline__ double scaled(double acc, const double* __restrict__ factors,
                                         int32_t factorcol, int32_t nvars, int32_t v) {
  if (factorcol == 0) return acc;
  double f = factors[v + int64_t(nvars) * (factorcol - 1)];
  return __dadd_rn(__dmul_rn(acc, f), 0.0);
}

// The full tensor sum of one target, any widths (the device form: no
// one-tap skip, as a flat launch in the reference takes it).
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
// Generic tiles: any group, one value per thread-iteration (the fallback
// for stencils no specialised kernel covers).
constexpr int kGenericPerThread = 4;
constexpr int kGenericTile = kThreads * kGenericPerThread;

__global__ void __launch_bounds__(kThreads) generic_kernel(const Job* __restrict__ jobs,
                                                           const Tile* __restrict__ tiles,
                                                           const TileBox* __restrict__) {
  const Tile tile = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const Job& job = jobs[tile.job];
  const DevGroup& g = job.groups[tile.group];
  const int64_t s1 = job.stride[1], s2 = job.stride[2], s3 = job.stride[3], s4 = job.stride[4];
  const int32_t nvars = job.nvars;
  const uint32_t total = uint32_t(g.points) * uint32_t(nvars);
  const double* sbase = job.work + s4 * g.sblocks[tile.transfer];
  double* tbase = job.work + s4 * g.tblocks[tile.transfer];
  for (int r = 0; r < kGenericPerThread; ++r) {
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

// ---------------------------------------------------------------------------
// Copies (unit stencils: same-level copies, mirrored and clamped rows): the
// source row of target row r along d is ca[d] + cs[d]*r, so no index table
// is read. Each thread issues all of its loads before its first store.
constexpr int kCopyPerThread = 8;
constexpr int kCopyTile = kThreads * kCopyPerThread;

__global__ void __launch_bounds__(kThreads) copy_kernel(const Job* __restrict__ jobs,
                                                        const Tile* __restrict__ tiles,
                                                        const TileBox* __restrict__) {
  const Tile tile = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const Job& job = jobs[tile.job];
  const DevGroup& g = job.groups[tile.group];
  // Offsets within one block fit 32 bits (plans refuse larger blocks).
  const int32_t s1 = int32_t(job.stride[1]), s2 = int32_t(job.stride[2]), s3 = int32_t(job.stride[3]);
  const int64_t s4 = job.stride[4];
  const int32_t nvars = job.nvars;
  const uint32_t total = uint32_t(g.points) * uint32_t(nvars);
  const double* __restrict__ sbase = job.work + s4 * g.sblocks[tile.transfer];
  double* __restrict__ tbase = job.work + s4 * g.tblocks[tile.transfer];
  const int32_t tb0 = g.tfirst[0] + s1 * g.tfirst[1] + s2 * g.tfirst[2];
  const int32_t sb0 = g.ca[0] + s1 * g.ca[1] + s2 * g.ca[2];
  const int32_t ds0 = g.cs[0], ds1 = s1 * g.cs[1], ds2 = s2 * g.cs[2];
  int32_t dst[kCopyPerThread];
  double val[kCopyPerThread];
#pragma unroll
  for (int r = 0; r < kCopyPerThread; ++r) {
    uint32_t e = uint32_t(tile.e0) + uint32_t(r * kThreads) + threadIdx.x;
    const bool ok = e < total;
    e = ok ? e : 0u;
    const uint32_t q = g.dn[0].div(e);
    const int32_t i = int32_t(e - q * uint32_t(g.n[0]));
    const uint32_t q2 = g.dn[1].div(q);
    const int32_t j = int32_t(q - q2 * uint32_t(g.n[1]));
    const uint32_t v32 = g.dn[2].div(q2);
    const int32_t k = int32_t(q2 - v32 * uint32_t(g.n[2]));
    const int32_t vo = s3 * int32_t(v32);
    dst[r] = ok ? vo + tb0 + i + s1 * j + s2 * k : -1;
    val[r] = sbase[vo + sb0 + ds0 * i + ds1 * j + ds2 * k];
  }
  if (g.factorcol == 0) {
#pragma unroll
    for (int r = 0; r < kCopyPerThread; ++r)
      if (dst[r] >= 0) tbase[dst[r]] = __dadd_rn(0.0, val[r]);
  } else {
#pragma unroll
    for (int r = 0; r < kCopyPerThread; ++r) {
      if (dst[r] < 0) continue;
      const int32_t v = (dst[r] - tb0) / s3;
      tbase[dst[r]] = scaled(__dadd_rn(0.0, val[r]), job.factors, g.factorcol, nvars, v);
    }
  }
}

// ---------------------------------------------------------------------------
// Interpolation (restriction and prolongation) of order P in every
// dimension. A thread computes R0 x R1 x R2 targets of one variable: along a
// dimension with R_d = 2 the two target rows' source windows start SHIFT
// apart (SHIFT = 0 for prolongation, whose pairs of fine points share a
// coarse window; 2 for restriction), so the thread loads the union window
// once and every target keeps its own accumulator, visiting its own taps in
// the reference's order (m3, m2, m1 ascending) with its own weight product.
template <int P, int SHIFT, int R0, int R1, int R2>
__global__ void __launch_bounds__(kThreads) interp_kernel(const Job* __restrict__ jobs,
                                                          const Tile* __restrict__ tiles,
                                                          const TileBox* __restrict__ boxes) {
  extern __shared__ double lds[];
  constexpr int U0 = P + (R0 == 2 ? SHIFT : 0);
  constexpr int U1 = P + (R1 == 2 ? SHIFT : 0);
  constexpr int U2 = P + (R2 == 2 ? SHIFT : 0);
  const uint32_t ti = xcd_tile(blockIdx.x, gridDim.x);
  const Tile tile = tiles[ti];
  const TileBox box = boxes[ti];
  const Job& job = jobs[tile.job];
  const DevSub& sub = job.subs[tile.group];
  const DevGroup& g = job.groups[sub.group];
  const int32_t s1 = int32_t(job.stride[1]), s2 = int32_t(job.stride[2]), s3 = int32_t(job.stride[3]);
  const int64_t s4 = job.stride[4];
  const int32_t nvars = job.nvars;

  // Stage the tile's source box in LDS, x fastest, coalesced rows.
  const int32_t e0 = box.ext[0], e1 = box.ext[1], e2 = box.ext[2];
  {
    const double* __restrict__ sblk = job.work + s4 * g.sblocks[tile.transfer] +
                                      (s3 * box.v0 + box.lo[0] + s1 * box.lo[1] + s2 * box.lo[2]);
    const FastDiv dx{uint32_t(e0)}, dy{uint32_t(e1)}, dz{uint32_t(e2)};
    const uint32_t n = uint32_t(e0 * e1 * e2 * box.nv);
    for (uint32_t idx = threadIdx.x; idx < n; idx += kThreads) {
      const uint32_t r = dx.div(idx);
      const int32_t x = int32_t(idx - r * uint32_t(e0));
      const uint32_t r2 = dy.div(r);
      const int32_t y = int32_t(r - r2 * uint32_t(e1));
      const uint32_t vv = dz.div(r2);
      const int32_t z = int32_t(r2 - vv * uint32_t(e2));
      lds[idx] = sblk[s3 * int32_t(vv) + x + s1 * y + s2 * z];
    }
  }
  __syncthreads();
  if (int32_t(threadIdx.x) >= box.count) return;

  const uint32_t e = uint32_t(tile.e0) + threadIdx.x;
  const uint32_t v32 = sub.du.div(e);
  const uint32_t u = e - v32 * uint32_t(sub.units);
  const uint32_t q = sub.dc[0].div(u);
  const int32_t ux = int32_t(u - q * uint32_t(sub.cnt[0]));
  const uint32_t uz = sub.dc[1].div(q);
  const int32_t uy = int32_t(q - uz * uint32_t(sub.cnt[1]));
  const int32_t v = int32_t(v32);

  int32_t ra[R0], rb[R1], rc[R2];
#pragma unroll
  for (int a = 0; a < R0; ++a) ra[a] = sub.rows[0][ux * R0 + a];
#pragma unroll
  for (int b = 0; b < R1; ++b) rb[b] = sub.rows[1][uy * R1 + b];
#pragma unroll
  for (int c = 0; c < R2; ++c) rc[c] = sub.rows[2][int32_t(uz) * R2 + c];

  double w0[R0][P], w1[R1][P], w2[R2][P];
#pragma unroll
  for (int a = 0; a < R0; ++a)
#pragma unroll
    for (int m = 0; m < P; ++m) w0[a][m] = g.weights[0][P * ra[a] + m];
#pragma unroll
  for (int b = 0; b < R1; ++b)
#pragma unroll
    for (int m = 0; m < P; ++m) w1[b][m] = g.weights[1][P * rb[b] + m];
#pragma unroll
  for (int c = 0; c < R2; ++c)
#pragma unroll
    for (int m = 0; m < P; ++m) w2[c][m] = g.weights[2][P * rc[c] + m];

  const double* __restrict__ src =
      lds + (g.srcstart[0][ra[0]] - box.lo[0]) +
      e0 * ((g.srcstart[1][rb[0]] - box.lo[1]) +
            e1 * ((g.srcstart[2][rc[0]] - box.lo[2]) + e2 * (v - box.v0)));
  double acc[R0][R1][R2];
#pragma unroll
  for (int a = 0; a < R0; ++a)
#pragma unroll
    for (int b = 0; b < R1; ++b)
#pragma unroll
      for (int c = 0; c < R2; ++c) acc[a][b][c] = 0.0;

#pragma unroll 1
  for (int m2 = 0; m2 < U2; ++m2) {
    double c2[R2];
#pragma unroll
    for (int c = 0; c < R2; ++c) {
      const int k2 = m2 - c * SHIFT;
      double x = 0.0;
#pragma unroll
      for (int m = 0; m < P; ++m) x = (m == k2) ? w2[c][m] : x;
      c2[c] = x;
    }
#pragma unroll
    for (int m1 = 0; m1 < U1; ++m1) {
      double row[U0];
      const double* __restrict__ p = src + e0 * (m1 + e1 * m2);
#pragma unroll
      for (int m0 = 0; m0 < U0; ++m0) row[m0] = p[m0];
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
              const double w = __dmul_rn(__dmul_rn(w0[a][k0], w1[b][k1]), c2[c]);
              acc[a][b][c] = __dadd_rn(acc[a][b][c], __dmul_rn(w, row[m0]));
            }
          }
        }
      }
    }
  }
  double* __restrict__ dst = job.work + s4 * g.tblocks[tile.transfer] + s3 * v;
#pragma unroll
  for (int c = 0; c < R2; ++c)
#pragma unroll
    for (int b = 0; b < R1; ++b)
#pragma unroll
      for (int a = 0; a < R0; ++a)
        dst[(g.tfirst[0] + ra[a]) + s1 * (g.tfirst[1] + rb[b]) + s2 * (g.tfirst[2] + rc[c])] =
            scaled(acc[a][b][c], job.factors, g.factorcol, nvars, v);
}

}  // namespace tah
