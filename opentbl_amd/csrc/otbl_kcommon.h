// Helpers shared by the kernel translation units: Fortran-order coefficient
// accessors, the default workgroup size and the launch macro (every launch is
// bracketed by kbegin/kend so per-kernel HIP-event timing can be switched on).
#pragma once
#include "otbl_internal.h"

namespace otbl {

// (r, 1:n) coefficient arrays stored with column i at offset r*i
#define C3(a, r, i) (a)[3 * (size_t)(i) + (r)-1]
#define C2(a, r, i) (a)[2 * (size_t)(i) + (r)-1]
// small dense (r, c) arrays, column-major
#define M4(a, r, c) (a)[4 * ((c)-1) + (r)-1]
#define M5(a, r, c) (a)[5 * ((c)-1) + (r)-1]
#define M2(a, r, c) (a)[2 * ((c)-1) + (r)-1]

static constexpr int TPB = 256;

static inline unsigned nblk(size_t n, int t = TPB) { return (unsigned)((n + t - 1) / t); }

// XCD-contiguous block order.  The dispatcher hands workgroup b of a 1-D
// grid to XCD b % 8, and each XCD has its own L2; a kernel whose neighbouring
// blocks share cache lines (row chunks that do not start on a line) takes
// its logical block from here, so that XCD x works through the contiguous
// run of logical blocks [x G/8, (x+1) G/8) and the shared lines meet in one
// L2.  A bijection of [0, gridDim.x) for every grid size.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned G) {
  const unsigned per = G >> 3, rem = G & 7, x = b & 7, k = b >> 3;
  return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}
__device__ __forceinline__ unsigned xcd_block() { return xcd_remap(blockIdx.x, gridDim.x); }
// the same for a 2-D grid (linear block id x-fastest): logical (bx, by)
__device__ __forceinline__ void xcd_block2(unsigned &bx, unsigned &by) {
  const unsigned lb = xcd_remap(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
  by = lb / gridDim.x;
  bx = lb - by * gridDim.x;
}

#define LAUNCH(b, name, grid, block, shm, kern, ...)                          \
  do {                                                                         \
    kbegin(b, name);                                                           \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), shm, (b).s, __VA_ARGS__); \
    kend(b);                                                                   \
  } while (0)

} // namespace otbl
