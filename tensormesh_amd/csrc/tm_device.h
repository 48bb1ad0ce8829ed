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

// Bijection from the hardware dispatch order (workgroup b of a launch runs on
// XCD b mod 8) to a logical order in which every XCD owns one contiguous
// range of [0, nb): neighbouring logical workgroups then share that XCD's L2,
// and so do launches that map the same logical workgroup to the same rows.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t q = nb / 8, r = nb % 8;
  const int64_t x = b % 8, y = b / 8;
  return x < r ? x * (q + 1) + y : r * (q + 1) + (x - r) * q + y;
}

// Float32 geometry of a P1 simplex, and when it must be redone in float64.
// With edges e_k = x_{k+1} - x_0 and det = det[e_0 .. e_{D-1}], the shape
// quality q = |det| / prod_k |e_k| (the volume over that of the box on the
// edges: 1 for right angles, -> 0 for slivers and needles) bounds the float32
// cancellation: the rounding error of the local matrix, relative to its
// largest entry, grows like eps / q (measured over 300,000 random tetrahedra
// and triangles with q down to 1e-4: at most 7.8e-7 for q >= 0.1).  Elements
// with q < kSliverQ, and degenerate ones, are flagged for float64.  Evaluated
// as a product of O(h) ratios, so that no term under- or overflows for edge
// lengths h within float32's exponent range.
constexpr float kSliverQ = 0.1f;

template <int D>
__device__ __forceinline__ bool sliver_f32(const float (&e)[D][D], float det) {
  float l[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < D; ++j) s += e[k][j] * e[k][j];
    l[k] = s;
  }
  const float r0 = det * __builtin_amdgcn_rcpf(l[0]);
  float lhs, rhs;
  if constexpr (D == 3) {
    lhs = r0 * (det * __builtin_amdgcn_rcpf(l[1]));  // det^2 / (l0 l1)
    rhs = kSliverQ * kSliverQ * l[2];
  } else {
    lhs = r0 * det;  // det^2 / l0
    rhs = kSliverQ * kSliverQ * l[1];
  }
  return !(lhs >= rhs);  // NaN (a zero edge) counts as a sliver
}

// Per-point D x D sums of one element, spread over the element's L lanes:
//   out[q][r * D + c] = (diag && r == c) + sum_{b < B} a(q, r, b) * m(q, b, c)
// One (q, r, c) per lane when the MQ D D entries (MQ >= Q, the instance's
// largest rule) fit in one pass of the lanes, else one row (q, r) per lane.
// Both take the nodes in order, so they give the same bits.
template <int D, int B, int L, int MQ, typename A, typename M, typename Out>
__device__ __forceinline__ void point_sums(int lane, int Q, bool diag, A a, M m, Out out) {
  if constexpr (MQ * D * D <= L) {
    if (lane < Q * D * D) {
      const int q = lane / (D * D), r = (lane / D) % D, c = lane % D;
      double s = (diag && r == c) ? 1.0 : 0.0;
      for (int b = 0; b < B; ++b) s += a(q, r, b) * m(q, b, c);
      out(q, r, c) = s;
    }
    return;
  }
  for (int t = lane; t < Q * D; t += L) {
    const int q = t / D, r = t % D;
    double s[D];
#pragma unroll
    for (int c = 0; c < D; ++c) s[c] = (diag && r == c) ? 1.0 : 0.0;
    for (int b = 0; b < B; ++b) {
      const double ab = a(q, r, b);
#pragma unroll
      for (int c = 0; c < D; ++c) s[c] += ab * m(q, b, c);
    }
#pragma unroll
    for (int c = 0; c < D; ++c) out(q, r, c) = s[c];
  }
}

}  // namespace tmh
