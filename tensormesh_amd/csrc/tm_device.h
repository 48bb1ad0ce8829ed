// Device helpers shared by the element kernels (HIP translation units only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tmh {

// Node pairs an element kernel computes: all B*B (i, j) in row-major order,
// or with `upper` the B(B+1)/2 pairs i <= j (row-major as well).  Entry p of
// the LDS table packs (i << 8) | j; returns the number of pairs.
template <int B>
__device__ __forceinline__ int fill_pairs(uint16_t* pairs, bool upper) {
  if (!upper) {
    for (int p = threadIdx.x; p < B * B; p += blockDim.x) pairs[p] = (uint16_t)(((p / B) << 8) | (p % B));
    return B * B;
  }
  for (int p = threadIdx.x; p < B * (B + 1) / 2; p += blockDim.x) {
    int i = 0, r = p;
    while (r >= B - i) {  // row i holds B - i pairs
      r -= B - i;
      ++i;
    }
    pairs[p] = (uint16_t)((i << 8) | (i + r));
  }
  return B * (B + 1) / 2;
}

}  // namespace tmh
