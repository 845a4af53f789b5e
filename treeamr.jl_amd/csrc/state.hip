// scatter! and gather! (src/state.jl:101-157): the copies between a state
// vector, which holds every own point of every block and no ghosts, and the
// working array's own range. The state is laid out (x, y, z, variable,
// block), x fastest (statearray, src/state.jl:54-58); the working array
// follows its tah_layout, variables innermost in the arena layout.
//
// A workgroup takes a tile of K pencils (a run of at most XC own points of
// an x row at one y, z and block) times VC variables. It reads the tile in
// the order of the side it reads (the state: x fastest, then variable; an
// arena array: variable fastest, then x), stages it in LDS, and writes it
// in the order of the side it writes, so that both sides move in runs of
// consecutive doubles. Where the working array holds variables innermost a
// tile keeps every variable and cuts long rows into balanced runs of x;
// elsewhere it keeps whole rows and cuts the variables, so that each side's
// runs stay whole.
// The LDS rows have an odd pitch, so that reading a column of them (one x,
// consecutive variables) meets no bank conflicts. The kernels are plain
// copies: bit for bit what the reference's loops produce.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "api_common.hpp"
#include "device.hpp"
#include "plan.hpp"

using namespace tah;

namespace tah {
namespace {

constexpr int kStateThreads = 256;
// Values of one tile: 16 KiB of LDS, eight workgroups per CU (tiles of 1024
// to 8192 values measured 0.079-0.148 ns per owned point at 32^3, 20
// variables; 2048 was fastest).
constexpr int kStateTile = 2048;

struct StateJob {
  double* work;
  double* state;
  int64_t stride[5];
  int32_t nvars;
  int32_t first[3], ext[3];
  int32_t pitch;       // LDS row length: XC, made odd
  int32_t K, VC, XC;   // pencils, variables and points of a pencil in a tile
  int32_t nxc;         // pencils per x row
  int32_t vfirst;      // the working array holds variables innermost
  uint32_t npencils;
  FastDiv dxc, d1, d2;  // division by nxc, ext[1], ext[2]
  FastDiv dX, dVC;      // division by XC, VC
};

// The pencil, variable and point of a tile element, in either side's order.
struct Elem {
  int32_t pl, v, x;
};

// Tile element i of np pencils x VC variable slots x XC point slots: x
// fastest (the state's order) or the variable fastest (an arena array's).
__device__ __forceinline__ Elem state_order(const StateJob& J, uint32_t i) {
  const uint32_t r = J.dX.div(i), pl = J.dVC.div(r);
  return Elem{int32_t(pl), int32_t(r - pl * uint32_t(J.VC)), int32_t(i - r * uint32_t(J.XC))};
}

__device__ __forceinline__ Elem var_first_order(const StateJob& J, uint32_t i) {
  const uint32_t r = J.dVC.div(i), pl = J.dX.div(r);
  return Elem{int32_t(pl), int32_t(i - r * uint32_t(J.VC)), int32_t(r - pl * uint32_t(J.XC))};
}

// Pencil p: its first point x0 and its y, z and block.
__device__ __forceinline__ void pencil_of(const StateJob& J, uint32_t p, int32_t& x0, int32_t& y, int32_t& z,
                                          int32_t& b) {
  const uint32_t s = J.dxc.div(p);
  x0 = int32_t(p - s * uint32_t(J.nxc)) * J.XC;
  p = s;
  const uint32_t t = J.d1.div(p);
  y = int32_t(p - t * uint32_t(J.ext[1]));
  const uint32_t q = J.d2.div(t);
  z = int32_t(t - q * uint32_t(J.ext[2]));
  b = int32_t(q);
}

__device__ __forceinline__ int64_t state_off(const StateJob& J, int32_t y, int32_t z, int32_t b, int32_t v,
                                             int32_t x) {
  return (((int64_t(b) * J.nvars + v) * J.ext[2] + z) * J.ext[1] + y) * int64_t(J.ext[0]) + x;
}

__device__ __forceinline__ int64_t work_off(const StateJob& J, int32_t y, int32_t z, int32_t b, int32_t v,
                                            int32_t x) {
  return J.stride[4] * b + J.stride[0] * (J.first[0] + x) + J.stride[1] * (J.first[1] + y) +
         J.stride[2] * (J.first[2] + z) + J.stride[3] * v;
}

template <bool SCATTER>
__global__ void __launch_bounds__(kStateThreads) state_copy_kernel(const StateJob J) {
  extern __shared__ double lds[];
  const uint32_t p0 = blockIdx.x * uint32_t(J.K);
  const int32_t v0 = int32_t(blockIdx.y) * J.VC;
  const int32_t nv = min(J.VC, J.nvars - v0);
  const int32_t np = int32_t(min(uint32_t(J.K), J.npencils - p0));
  const uint32_t n = uint32_t(np) * uint32_t(J.VC) * uint32_t(J.XC);
  // Read: the state in its own order, or the working array in its own. The
  // last variable tile may hold fewer than VC variables, the last pencil of
  // a row fewer than XC points: their other slots are skipped.
#pragma unroll 4
  for (uint32_t i = threadIdx.x; i < n; i += kStateThreads) {
    const Elem e = (SCATTER || !J.vfirst) ? state_order(J, i) : var_first_order(J, i);
    int32_t x0, y, z, b;
    pencil_of(J, p0 + uint32_t(e.pl), x0, y, z, b);
    const int32_t x = x0 + e.x;
    if (e.v >= nv || x >= J.ext[0]) continue;
    lds[(e.pl * J.VC + e.v) * J.pitch + e.x] =
        SCATTER ? J.state[state_off(J, y, z, b, v0 + e.v, x)] : J.work[work_off(J, y, z, b, v0 + e.v, x)];
  }
  __syncthreads();
  // Write: the other side, in its own order.
#pragma unroll 4
  for (uint32_t i = threadIdx.x; i < n; i += kStateThreads) {
    const Elem e = (!SCATTER || !J.vfirst) ? state_order(J, i) : var_first_order(J, i);
    int32_t x0, y, z, b;
    pencil_of(J, p0 + uint32_t(e.pl), x0, y, z, b);
    const int32_t x = x0 + e.x;
    if (e.v >= nv || x >= J.ext[0]) continue;
    const double val = lds[(e.pl * J.VC + e.v) * J.pitch + e.x];
    if (SCATTER)
      J.work[work_off(J, y, z, b, v0 + e.v, x)] = val;
    else
      J.state[state_off(J, y, z, b, v0 + e.v, x)] = val;
  }
}

void state_copy(bool scatter, const tah_layout* L, const int32_t* first, const int32_t* extent, double* work,
                double* state, void* stream) {
  if (!L || !first || !extent) fail("null argument");
  validate_strides(L);
  int64_t points = 1;
  for (int d = 0; d < 3; ++d) {
    if (extent[d] < 1) fail("the own range's extents must be positive");
    if (first[d] < 0 || int64_t(first[d]) + extent[d] > L->stored[d])
      fail("the own range " + std::to_string(first[d]) + " .. " + std::to_string(first[d] + extent[d] - 1) +
               " does not fit the " + std::to_string(L->stored[d]) + " stored points of dimension " +
               std::to_string(d + 1),
           kLayout);
    points *= extent[d];
  }
  if (L->nblocks == 0) return;
  if (!work || !state) fail("null argument");
  StateJob J{};
  // Balanced cuts: n items in pieces of at most cap, as even as they go.
  auto piece = [](int32_t n, int32_t cap) {
    const int32_t pieces = (n + cap - 1) / cap;
    return (n + pieces - 1) / pieces;
  };
  J.vfirst = L->stride[3] < L->stride[0] ? 1 : 0;
  if (J.vfirst) {
    J.VC = piece(L->nvars, kStateTile);
    J.XC = piece(extent[0], std::max(1, kStateTile / J.VC));
  } else {
    J.XC = piece(extent[0], kStateTile);
    J.VC = piece(L->nvars, std::max(1, kStateTile / (J.XC | 1)));
  }
  J.pitch = J.XC | 1;
  J.K = std::max(1, kStateTile / (J.VC * J.pitch));
  J.nxc = (extent[0] + J.XC - 1) / J.XC;
  const int64_t pencils = points / extent[0] * J.nxc * L->nblocks;
  if (pencils > int64_t(UINT32_MAX)) fail("more than 2^32 pencils are not supported", kUnsupported);
  J.work = work;
  J.state = state;
  for (int k = 0; k < 5; ++k) J.stride[k] = L->stride[k];
  J.nvars = L->nvars;
  for (int d = 0; d < 3; ++d) {
    J.first[d] = first[d];
    J.ext[d] = extent[d];
  }
  J.npencils = uint32_t(pencils);
  J.dxc = FastDiv(uint32_t(J.nxc));
  J.d1 = FastDiv(uint32_t(extent[1]));
  J.d2 = FastDiv(uint32_t(extent[2]));
  J.dX = FastDiv(uint32_t(J.XC));
  J.dVC = FastDiv(uint32_t(J.VC));
  const size_t lds = size_t(J.K) * J.VC * J.pitch * sizeof(double);
  if (lds > 64 * 1024) fail("a state tile needs more than 64 KiB of LDS", kUnsupported);
  const dim3 grid(uint32_t((pencils + J.K - 1) / J.K), uint32_t((L->nvars + J.VC - 1) / J.VC));
  if (grid.y > 65535) fail("too many variables for one launch", kUnsupported);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (scatter)
    hipLaunchKernelGGL(state_copy_kernel<true>, grid, dim3(kStateThreads), lds, s, J);
  else
    hipLaunchKernelGGL(state_copy_kernel<false>, grid, dim3(kStateThreads), lds, s, J);
  HIPCHECK(hipGetLastError());
}

}  // namespace
}  // namespace tah

extern "C" {

tah_status tah_state_length(const tah_layout* layout, const int32_t* extent, int64_t* out) {
  return guard([&] {
    if (!layout || !extent || !out) fail("null argument");
    if (layout->nvars < 1 || layout->nblocks < 0) fail("bad layout", kLayout);
    int64_t n = int64_t(layout->nvars) * layout->nblocks;
    for (int d = 0; d < 3; ++d) {
      if (extent[d] < 1) fail("the own range's extents must be positive");
      n *= extent[d];
    }
    *out = n;
  });
}

tah_status tah_scatter(const tah_layout* layout, const int32_t* first, const int32_t* extent, double* work,
                       const double* state, void* stream) {
  return guard([&] { state_copy(true, layout, first, extent, work, const_cast<double*>(state), stream); });
}

tah_status tah_gather(const tah_layout* layout, const int32_t* first, const int32_t* extent, double* state,
                      const double* work, void* stream) {
  return guard([&] { state_copy(false, layout, first, extent, const_cast<double*>(work), state, stream); });
}

}  // extern "C"
