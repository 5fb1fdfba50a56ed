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

#define LAUNCH(b, name, grid, block, shm, kern, ...)                          \
  do {                                                                         \
    kbegin(b, name);                                                           \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), shm, (b).s, __VA_ARGS__); \
    kend(b);                                                                   \
  } while (0)

} // namespace otbl
